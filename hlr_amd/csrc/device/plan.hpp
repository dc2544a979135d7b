// plan.hpp -- device layout of a flattened H-/TLR-matrix (host and device).
//
// The matrix is cut into row panels and column panels of at most 64 indices
// (one wavefront).  Every leaf contributes
//   - to each row panel it covers: its dense rows D[p,:] (width n) or its U
//     rows U[p,:] (width k), stored column-group-interleaved ("R arena");
//   - to each column panel it covers (low-rank only): its V rows V[q,:]
//     (width k), stored row-major ("V arena").
// Columns inside a panel are grouped by storage format; a group of W columns
// (W * sizeof(element) == 16 bytes) of one panel is m x W row-major, so one
// lane (= one row) loads its W values with a single 16-byte load.
//
// Forward  y += a A x :  P = V^T x per column panel (no reduction),
//                        t = a * sum_q P,  y_p += R_p z  (z = a x | t).
// Transposed y += a A^T x: Q = R_p^T x_p per row panel (lane reduction),
//                        s = a * sum_p Q,  y_q += V_q s + a * sum Q_dense.
#pragma once
#include <cstdint>

#if defined(__HIP__)
#define HMVM_HD __host__ __device__
#else
#define HMVM_HD
#endif

namespace hmvm {

enum Fmt : int32_t { F64 = 0, F32 = 1, BF16 = 2, FP8 = 3, NFMT = 4 };

constexpr int kPanel = 64;          // rows per panel = lanes per wave
constexpr int kChunkGroups = 64;    // column groups per R chunk
constexpr int32_t kPadCol = INT32_MIN;

HMVM_HD inline int fmt_bytes(int f) { return f == F64 ? 8 : f == F32 ? 4 : f == BF16 ? 2 : 1; }
HMVM_HD inline int fmt_width(int f) { return 16 / fmt_bytes(f); }

struct RowPanel {
    int64_t row0;       // first global row
    int32_t m;          // rows (<= 64)
    int32_t nchunks;    // chunks of this panel
    int64_t chunk0;     // first chunk in the chunk table
};

struct RChunk {         // up to kChunkGroups groups of one format of one row panel
    int32_t fmt;
    int32_t ngroups;
    int64_t data_off;   // element offset of group 0 in arena[fmt]; group g at data_off + g*m*W
    int64_t col0;       // first column in the per-R-column tables (ngroups*W columns)
    int32_t panel;      // row panel index
    int32_t pad_;
};

struct ColPanel {
    int64_t col0;       // first global column
    int32_t m;          // rows of V in this panel (<= 64)
    int32_t pad_;
    int64_t off[NFMT];  // element offset of the m x kpad[f] row-major block in arena[f]
    int32_t kpad[NFMT]; // padded width per format (multiple of W)
    int64_t vcol0[NFMT];// first column in the per-V-column tables
};

struct VItem {          // forward V work unit: <= 64 column groups of one format of one column panel
    int32_t q;
    int32_t fmt;
    int32_t g0;
    int32_t ng;
};

} // namespace hmvm
