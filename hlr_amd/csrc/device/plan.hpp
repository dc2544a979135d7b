// plan.hpp -- device layout of a flattened H-/TLR-matrix (host and device).
//
// The matrix is cut into row panels and column panels of at most 64 indices
// (one wavefront).  Every leaf contributes
//   - to each row panel it covers: its dense rows D[p,:] (width n) or its U
//     rows U[p,:] (width k)                                     ("R arena");
//   - (low-rank only) its V rows, grouped by the column index set the leaf
//     shares with others and cut into units of up to four column panels
//     (256 rows)                                                ("V arena").
// Columns are grouped by storage format and cut into groups of W columns,
// W * sizeof(element) == 16 bytes, so that a lane loads W values of one row
// with a single 16-byte load.
//   R: a chunk of up to kChunkGroups groups of one format is contiguous;
//      group g of the chunk is an m x W row-major block (lane = row).
//   V: a segment of up to 64 groups of one format of one unit is
//      contiguous; it is an m x (ng*W) row-major block (lane = group, rows
//      streamed); its rows of one column panel are that panel's view.
// Every chunk and segment is one contiguous stream of HBM.
//
// Forward  y += a A x :  P = V^T x per unit (no reduction),
//                        t = a * sum_u P,  y_p += R_p z  (z = a x | t).
// Transposed y += a A^T x: Q = R_p^T x_p per row panel (lane reduction),
//                        s = a * sum_p Q,  y_q += V_q s + a * sum Q_dense
//                        (V_q: the panel views of the segments).
#pragma once
#include <cstdint>

#if defined(__HIP__)
#define HMVM_HD __host__ __device__
#else
#define HMVM_HD
#endif

namespace hmvm {

// Storage formats of factor columns.  F64, F32, BF16 and FP8 (E4M3) hold
// W = 16 / sizeof(element) values per 16-byte row.  B8, B16, B25 are block
// floating point: NB-bit two's-complement integers q, W = floor(128 / NB) of
// them per 16-byte row (value j in bits [j NB, j NB + NB) of the row read as
// a little-endian 128-bit integer), each standing for q * 2^k with one
// power-of-two scale 2^k per panel column (as FP8's scale).  A12, A16, A24
// are adaptive float (AFL) codes packed the same way: NB = 1 + 5 + m bits
// (sign, 5 exponent bits, m = 6, 10, 18 mantissa bits) standing for
// +-(a - 1) s with a = 2^(e + 1) (1 + mant 2^-m) and one scale s per panel
// column; the all-ones code is zero.
enum Fmt : int32_t { F64 = 0, F32 = 1, BF16 = 2, FP8 = 3, B8 = 4, B16 = 5, B25 = 6, A12 = 7, A16 = 8, A24 = 9,
                     NFMT = 10 };
constexpr int kFirstBfp = B8;                 // first packed format (4-word rows in host arenas)
constexpr int kNumBfp = NFMT - kFirstBfp;     // packed formats: block floating point and AFL
constexpr int kFirstAfl = A12;
constexpr int kAflExpBits = 5;

// Format families: a plan holds formats of one family only (the plan
// builder enforces it), so a kernel instantiated for that family carries the
// decode paths, and the registers, of its own formats alone.  FP64 and FP32
// belong to both.
// kFamF64: a plan that stores FP64 values only (the matrix-core kernels
// have an instantiation of their own for it)
// kFamAfl: FP64 / FP32 and the AFL formats
enum FmtFamily : int { kFamAll = 0, kFamHw = 1, kFamBfp = 2, kFamF64 = 3, kFamAfl = 4 };

constexpr int kPanel = 64;          // rows per panel = lanes per wave
constexpr int kWaves = 4;           // waves per workgroup
constexpr int kChunkGroups = 64 * kWaves;   // column groups per R chunk (one per lane of the workgroup)
constexpr int kItemGroups = 64;     // column groups per V item (one per lane of a wave)
constexpr int32_t kPadCol = INT32_MIN;

// bits per stored value, values per 16-byte row
HMVM_HD constexpr int fmt_bits(int f) {
    return f == F64 ? 64 : f == F32 ? 32 : f == BF16 ? 16 : f == FP8 ? 8 : f == B8 ? 8 : f == B16 ? 16
         : f == B25 ? 25 : f == A12 ? 12 : f == A16 ? 16 : 24;
}
HMVM_HD constexpr int fmt_width(int f) { return 128 / fmt_bits(f); }
// packed in 16-byte rows of NB-bit fields (host arenas of 32-bit words)
HMVM_HD constexpr bool fmt_packed(int f) { return f >= kFirstBfp; }
HMVM_HD constexpr bool fmt_is_bfp(int f) { return f >= kFirstBfp && f < kFirstAfl; }
HMVM_HD constexpr bool fmt_is_afl(int f) { return f >= kFirstAfl && f < NFMT; }
// mantissa bits of an AFL format
HMVM_HD constexpr int afl_mant(int f) { return fmt_bits(f) - 1 - kAflExpBits; }
// formats whose panel columns carry a power-of-two scale
HMVM_HD constexpr bool fmt_scaled(int f) { return f >= FP8; }
// bytes of `values` stored values (a multiple of the format's width)
HMVM_HD constexpr int64_t fmt_store_bytes(int f, int64_t values) { return values / fmt_width(f) * 16; }

// NB-bit field j of a 16-byte row w[0..3] (little-endian words), sign-extended
template <int NB>
HMVM_HD inline int32_t bfp_field(const uint32_t (&w)[4], int j) {
    const int bit = j * NB, word = bit >> 5, off = bit & 31;
    uint32_t u = w[word] >> off;
    if (off + NB > 32) u |= w[word + 1] << (32 - off);
    return int32_t(u << (32 - NB)) >> (32 - NB);
}

// NB-bit field j of a 16-byte row, unsigned
template <int NB>
HMVM_HD inline uint32_t afl_field(const uint32_t (&w)[4], int j) {
    const int bit = j * NB, word = bit >> 5, off = bit & 31;
    uint32_t u = w[word] >> off;
    if (off + NB > 32) u |= w[word + 1] << (32 - off);
    return (NB == 32) ? u : (u & ((1u << NB) - 1u));
}

struct RowPanel {
    int64_t row0;       // first global row
    int32_t m;          // rows (<= 64)
    int32_t nchunks;    // chunks of this panel
    int64_t chunk0;     // first chunk in the chunk table
    int32_t zbytes;     // LDS bytes of all z values of the panel (decoded accumulation type)
    int32_t pad_;
};

struct RChunk {         // up to kChunkGroups groups of one format of one row panel
    int32_t fmt;
    int32_t ngroups;
    int64_t data_off;   // element offset of group 0 in arena[fmt]; group g at data_off + g*m*W
    int64_t col0;       // first column in the per-R-column tables (ngroups*W columns)
    int32_t panel;      // row panel index
    int32_t zoff;       // byte offset of the chunk's z values in the panel's LDS block
};

struct ColPanel {
    int64_t col0;       // first global column
    int32_t m;          // rows of V in this panel (<= 64)
    int32_t nitems;     // panel views of V segments (VItem) of this panel
    int64_t item0;      // first view in the panel view table
};

constexpr int kPackSegs = 8;        // V segments per forward pack
constexpr int kSegLen = 16;         // partial sums per thread in the first summation stage
// A V segment covers one unit of a column cluster: up to kUnitPanels
// consecutive column panels (at most kUnitRows rows) of the column index set
// shared by its blocks; a low-rank block has one forward partial per unit.
constexpr int kUnitPanels = 4;
constexpr int kUnitRows = kUnitPanels * kPanel;
constexpr int kPackXs = kPackSegs * 65;      // LDS values of a pack's x slices (per wave)

// V segment (forward table) or its view in one column panel (panel table):
// an m x (ng W) row-major block of one format, lane = group, row stride ng W
// values; its columns are V columns vc0 .. vc0 + ng W - 1 of the per-V-column
// tables, for rows x0 .. x0 + m - 1 of the vectors
struct VItem {
    int32_t q;          // panel views: column panel; segments: first panel of the unit
    int32_t fmt;
    int32_t ng;
    int32_t m;          // rows (segments <= kUnitRows, panel views <= kPanel)
    int64_t data_off;   // element offset of the block's first row in arena[fmt]
    int64_t vc0;        // first V column
    int64_t x0;         // first row's global index
};

// one lane of the single right-hand-side partial sums: entry e (-1:
// padding) with its partials at off + r * stride, r < cnt; lane i of the
// entry's L lanes (il = i | L << 8)
struct SumLane {
    int32_t off, cnt, stride, e;
};

// transposed / gradient R work item: groups [g0, min(g0 + kRtSpan, ngroups)) of a chunk
constexpr int32_t kRtSpan = 32;
struct RtItem {
    int32_t chunk;
    int32_t g0;
};

struct VPack {          // forward V work unit of one wave: consecutive segments of one format,
    int32_t fmt;        // at most 64 groups in total (lane = group); everything a lane
    int32_t nseg;       // needs is stored here, so the pack starts without load chains
    int32_t gofs[kPackSegs + 1];   // prefix sums of the segments' group counts
    int32_t m[kPackSegs];          // rows of each segment (<= kUnitRows)
    int32_t ng[kPackSegs];         // groups of each segment
    int32_t xoff[kPackSegs];       // first value of each segment's x slice in the wave's LDS
    int32_t x0[kPackSegs];         // first x index of each segment
    int64_t vc0[kPackSegs];        // per-V-column table index of each segment's column 0
    int64_t data_off[kPackSegs];   // element offset of each segment's block
};

} // namespace hmvm
