// fwd.hpp -- device routines of the forward product shared by the
// per-plan kernels (hmvm.hip) and the batched kernels (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace hmvm {

constexpr int kThreads = 64 * kWaves;
constexpr int64_t kWholeZBytes = 48 * 1024;   // largest panel z block staged whole in LDS

// ---------------------------------------------------------------------------
// V pass: one wave processes one pack of consecutive segments (lane = group)
// ---------------------------------------------------------------------------

// The x slices of a pack's segments (up to kUnitRows values each) in the
// wave's LDS, as doubles for the FP64 format and as floats for the others
// (the accumulation type of the pack's format); before the format switch,
// so that a kernel holds one copy of it.  The slices lie an odd distance
// apart: the lanes of different segments read different rows i of their
// own slice in the same step, which then fall into different LDS banks.
// Slot p of the slices belongs to the last segment t with xoff[t] <= p
// (xoff[0] = 0) and holds a value when p - xoff[t] < m[t] (the slot after
// each slice is a gap).
//
// The pack pointer is uniform across the wave, so the segment table is
// read with scalar loads issued together, and all x loads of the pack are
// issued before the first store to LDS.  Reading the table field by field
// through per-lane loads had made the start of a pack a chain of dependent
// round trips (m, then xoff and x0, then the x values, for every segment in
// turn).
template <typename XT>
__device__ inline void vpack_stage_x(const VPack* __restrict__ pkp, const XT* __restrict__ x, void* xsraw,
                                     int lane) {
    const VPack& pk = *pkp;     // read through the pointer: the arrays stay out of scratch
    const int nseg = pk.nseg;
    int32_t sm[kPackSegs], sxo[kPackSegs], sx0[kPackSegs];
#pragma unroll
    for (int t = 0; t < kPackSegs; ++t) {
        sm[t] = pk.m[t];
        sxo[t] = pk.xoff[t];
        sx0[t] = pk.x0[t];
    }
    // the slots up to the end of the last slice, 64 per load
    int xend = sxo[0] + sm[0];
#pragma unroll
    for (int t = 1; t < kPackSegs; ++t)
        if (t == nseg - 1) xend = sxo[t] + sm[t];
    constexpr int kXq = (kPackXs + 63) / 64;
    double xv[kXq];
    int xp[kXq];
#pragma unroll
    for (int q = 0; q < kXq; ++q) {
        xp[q] = -1;
        xv[q] = 0;
        if (64 * q >= xend) continue;           // uniform: the loads before stay in flight
        const int p = lane + 64 * q;
        int o = sxo[0], mm = sm[0], b = sx0[0];
#pragma unroll
        for (int t = 1; t < kPackSegs; ++t)
            if (t < nseg && p >= sxo[t]) {
                o = sxo[t];
                mm = sm[t];
                b = sx0[t];
            }
        // an unconditional load (index 0 for gap and spare slots), so that
        // no branch ends the batch of loads
        const bool on = p - o < mm;
        xp[q] = on ? p : -1;
        xv[q] = double(x[on ? b + (p - o) : 0]);
    }
    if (pk.fmt == F64) {
        double* xs = static_cast<double*>(xsraw);
#pragma unroll
        for (int q = 0; q < kXq; ++q)
            if (xp[q] >= 0) xs[xp[q]] = xv[q];
    } else {
        float* xs = static_cast<float*>(xsraw);
#pragma unroll
        for (int q = 0; q < kXq; ++q)
            if (xp[q] >= 0) xs[xp[q]] = float(xv[q]);
    }
}

// The stream of a pack whose x slices vpack_stage_x has staged: lane =
// group of the pack, W column sums per lane.
template <typename T>
__device__ inline void vpack_fwd(const DevPlan& P, const VPack* pkp, double* __restrict__ Pb,
                                 void* xsraw, int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    static_assert(sizeof(A) == (E::F == F64 ? 8 : 4), "x slices are doubles for FP64 only");
    const VPack& pk = *pkp;     // read through the pointer: the arrays stay out of scratch
    const int nseg = pk.nseg;
    int32_t gofs[kPackSegs + 1];
#pragma unroll
    for (int t = 0; t <= kPackSegs; ++t) gofs[t] = pk.gofs[t];
    int ngr = gofs[1];                              // groups of the pack = active lanes
#pragma unroll
    for (int t = 2; t <= kPackSegs; ++t)
        if (t == nseg) ngr = gofs[t];
    A* xs = static_cast<A*>(xsraw);                 // kPackXs values: segment j's slice at xoff[j]
    // the lane's segment: the last t < nseg with gofs[t] <= lane
    int gj = gofs[0], mj = pk.m[0], xoj = pk.xoff[0];
    int32_t ngj = pk.ng[0];
    int64_t dj = pk.data_off[0];
#pragma unroll
    for (int t = 1; t < kPackSegs; ++t)
        if (t < nseg && lane >= gofs[t]) {
            gj = gofs[t];
            mj = pk.m[t];
            xoj = pk.xoff[t];
            ngj = pk.ng[t];
            dj = pk.data_off[t];
        }
    const int g = lane - gj;
    const int m = (lane < ngr) ? mj : 0;            // lanes past the pack's groups load nothing
    const int64_t ldb = int64_t(ngj) * 16;          // bytes per row of the segment
    const bytep base = elem_addr<T>(P.arena[E::F], dj) + int64_t(g) * 16;
    const A* xj = xs + xoj;
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    // rows per unrolled step (measured best; 8 for W >= 8 since the pack's
    // table and epilogue loads no longer wait on each other: AFL V pass
    // 0.179 -> 0.163 ms, the others 0.5-1 % faster, where 4 had been best;
    // 8 for FP32 as well: AFL V pass 0.164 -> 0.158 ms, others unchanged)
    constexpr int kUnroll = (W >= 8 || W == 4) ? 8 : 16;
#pragma unroll kUnroll
    for (int i = 0; i < m; ++i) {
        A v[W];
        E::load(base + int64_t(i) * ldb, v);
        const A xi = xj[i];
#pragma unroll
        for (int w = 0; w < W; ++w) acc[w] += v[w] * xi;
    }
    // The lane's W sums are columns lane W .. lane W + W - 1 of the pack.
    // They go out through LDS (the x slices are no longer read), so that
    // consecutive lanes then handle consecutive columns: the destination
    // and scale tables are read and the partials written coalesced, instead
    // of W strided accesses per lane (one cache line per lane for wide W).
    static_assert(64 * W * sizeof(A) <= kPackXs * sizeof(double), "result staging fits the slices");
    wave_sync();
    A* res = xs;
#pragma unroll
    for (int w = 0; w < W; ++w) res[lane * W + w] = acc[w];
    wave_sync();
    // Column c of the pack is column c - W gofs[s] of segment s: its table
    // index is vc0[s] - W gofs[s] + c.  The segment bases are read once
    // (scalar loads) and selected per column in registers; then all 2 W
    // table loads of a lane are issued, unconditionally (index 0 for
    // columns past the pack), before the first store.  Selecting with
    // per-column loads of vc0 under divergent conditions had made every
    // selection wait for all loads in flight: up to W (1 + segments)
    // dependent round trips per pack.
    // The table is read again here, after the barrier (the pack pointer is
    // not restrict-qualified, so the compiler does not merge these reads
    // with the earlier ones across its fences), and none of it stays in
    // scalar registers through the stream: at the kernel's SGPR limit that
    // had pushed the decode's compares onto VCC, with hazard NOPs.  (V
    // column indices are below 2^31: the plan builder checks.)
    const int nsg = pk.nseg;
    int32_t eg[kPackSegs + 1], vc[kPackSegs];
#pragma unroll
    for (int t = 0; t <= kPackSegs; ++t) eg[t] = pk.gofs[t];
#pragma unroll
    for (int t = 0; t < kPackSegs; ++t) vc[t] = int32_t(pk.vc0[t]) - eg[t] * W;
    int ecol = eg[1] * W;
#pragma unroll
    for (int t = 2; t <= kPackSegs; ++t)
        if (t == nsg) ecol = eg[t] * W;
    int32_t ix[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int c = lane + 64 * k;
        const int gr = c / W;
        int32_t vb = vc[0];
#pragma unroll
        for (int t = 1; t < kPackSegs; ++t) vb = (t < nsg && gr >= eg[t]) ? vc[t] : vb;
        ix[k] = (c < ecol) ? vb + c : 0;
    }
    int32_t vd[W];
    float vs[W];
#pragma unroll
    for (int k = 0; k < W; ++k) vd[k] = P.vdest[ix[k]];
#pragma unroll
    for (int k = 0; k < W; ++k) vs[k] = E::kScaled ? P.vscale[ix[k]] : 1.0f;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int c = lane + 64 * k;
        if (c < ecol && vd[k] >= 0) Pb[vd[k]] = double(res[c]) * double(vs[k]);
    }
}


// ---------------------------------------------------------------------------
// R pass: chunk staging (z values in LDS) and the streaming inner product
// ---------------------------------------------------------------------------

template <typename T, typename XT>
__device__ inline void chunk_stage(const DevPlan& P, const RChunk& C, double alpha, const XT* __restrict__ x,
                                   const double* __restrict__ t, void* zbuf, int tid) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    A* zs = static_cast<A*>(zbuf);
    const int ncol = C.ngroups * W;
    // U columns per thread per step, their index, scale and source loads
    // issued together (one memory latency per step, not one per column)
    constexpr int U = 4;
    for (int j0 = tid; j0 < ncol; j0 += U * kThreads) {
        int32_t s[U];
        float sc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kThreads;
            s[u] = (j < ncol) ? P.colsrc[C.col0 + j] : kPadCol;
            sc[u] = (E::kScaled && j < ncol) ? P.zscale[C.col0 + j] : 1.0f;
        }
        double z[U];
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = (s[u] >= 0) ? alpha * double(x[s[u]]) : (s[u] == kPadCol ? 0.0 : t[~s[u]]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * kThreads;
            if (j < ncol) zs[j] = A(E::kScaled ? z[u] * double(sc[u]) : z[u]);
        }
    }
}

// The z values of all chunks of a row panel, staged in LDS before any
// stream (the panel's z block fits the LDS): one index space over the
// chunks' columns in panel order, U columns per thread per round, so that
// the loads of all chunks are in flight together.  Staged chunk by chunk,
// each chunk's round ended with a wait for its own loads, and a panel of
// four chunks waited on eight dependent round trips before its stream.  The
// values are chunk_stage's: alpha x or t (0 for padding), times the column
// scale in the scaled formats, rounded to the chunk's accumulation type.
template <typename XT>
__device__ inline void panel_stage(const DevPlan& P, const RowPanel& R, double alpha, const XT* __restrict__ x,
                                   const double* __restrict__ t, unsigned char* zdyn, int tid) {
    int total = 0;
    for (int ci = 0; ci < R.nchunks; ++ci) {
        const RChunk& C = P.rchunks[R.chunk0 + ci];
        total += C.ngroups * fmt_width(C.fmt);
    }
    const int64_t col0 = (R.nchunks > 0) ? P.rchunks[R.chunk0].col0 : 0;
    constexpr int U = 4;
    for (int j0 = 0; j0 < total; j0 += U * kThreads) {
        // the chunk of each of the thread's columns (a scan of the panel's
        // chunks, uniform across the workgroup)
        int64_t col[U];
        int zo[U], f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            col[u] = -1;
            zo[u] = 0;
            f[u] = F64;
        }
        int start = 0;
        for (int ci = 0; ci < R.nchunks; ++ci) {
            const RChunk& C = P.rchunks[R.chunk0 + ci];
            const int n = C.ngroups * fmt_width(C.fmt);
            if (start < j0 + U * kThreads && start + n > j0) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int j = j0 + tid + u * kThreads - start;
                    if (j >= 0 && j < n) {
                        col[u] = C.col0 + j;
                        f[u] = C.fmt;
                        zo[u] = C.zoff + j * (C.fmt == F64 ? int(sizeof(double)) : int(sizeof(float)));
                    }
                }
            }
            start += n;
        }
        // unconditional loads at safe indices (the panel's first column,
        // x[0], t[0]), masked afterwards: a load under a divergent
        // condition ends its block with a wait.  zscale exists in plans
        // with scaled formats, t when there are low-rank entries.
        int64_t cl[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cl[u] = (col[u] >= 0) ? col[u] : col0;
        int32_t s[U];
        float sc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t v = P.colsrc[cl[u]];
            s[u] = (col[u] >= 0) ? v : kPadCol;
        }
        if (P.zscale != nullptr) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float v = P.zscale[cl[u]];
                sc[u] = fmt_scaled(f[u]) ? v : 1.0f;
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) sc[u] = 1.0f;
        }
        double xv[U], tv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = double(x[s[u] >= 0 ? s[u] : 0]);
#pragma unroll
        for (int u = 0; u < U; ++u) tv[u] = 0.0;
        if (P.ntent > 0) {
#pragma unroll
            for (int u = 0; u < U; ++u) tv[u] = t[(s[u] < 0 && s[u] != kPadCol) ? ~s[u] : 0];
        }
        double z[U];
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = (s[u] >= 0) ? alpha * xv[u] : (s[u] == kPadCol ? 0.0 : tv[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (col[u] < 0) continue;
            if (f[u] == F64)
                *reinterpret_cast<double*>(zdyn + zo[u]) = z[u];
            else
                *reinterpret_cast<float*>(zdyn + zo[u]) = float(fmt_scaled(f[u]) ? z[u] * double(sc[u]) : z[u]);
        }
    }
}

template <typename T>
__device__ inline double chunk_range(const DevPlan& P, const RChunk& C, int m, const void* zbuf, int gb, int ge,
                                     int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    const A* zs = static_cast<const A*>(zbuf);
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    if (lane < m) {
        const bytep base = elem_addr<T>(P.arena[E::F], C.data_off) + int64_t(lane) * 16;
        const int64_t gs = int64_t(m) * 16;                      // bytes per group
        constexpr int kUnroll = (W >= 8) ? 4 : 8;
#pragma unroll kUnroll
        for (int g = gb; g < ge; ++g) {
            A v[W];
            E::load(base + g * gs, v);
#pragma unroll
            for (int w = 0; w < W; ++w) acc[w] += v[w] * zs[g * W + w];
        }
    }
    double r = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) r += double(acc[w]);
    return r;
}

// the wave's contiguous quarter of one chunk
template <typename T>
__device__ inline double chunk_compute(const DevPlan& P, const RChunk& C, int m, const void* zbuf, int wave,
                                       int lane) {
    const int per = (C.ngroups + kWaves - 1) / kWaves;
    const int gb = wave * per;
    return chunk_range<T>(P, C, m, zbuf, gb, min(gb + per, int(C.ngroups)), lane);
}

// kWhole: the panel's z values fit in dynamic LDS and are staged once;
// otherwise each chunk is staged into a fixed buffer between barriers.

// Whole row panel with chunk-by-chunk staging into a buffer of
// kChunkGroups * 16 floats; all kThreads threads call it; returns the
// lane's partial row sum (rows = lanes, waves split each chunk).
template <typename XT, int FAM = kFamAll>
__device__ inline double rpanel_fwd_chunked(const DevPlan& P, const RowPanel& R, double alpha,
                                            const XT* __restrict__ x, const double* __restrict__ t, void* zbuf,
                                            int tid, int wave, int lane) {
    double acc = 0;
    for (int ci = 0; ci < R.nchunks; ++ci) {
        const RChunk C = P.rchunks[R.chunk0 + ci];
        __syncthreads();
        HMVM_FMT_SWITCH_FAM(FAM, C.fmt, chunk_stage<T>(P, C, alpha, x, t, zbuf, tid))
        __syncthreads();
        HMVM_FMT_SWITCH_FAM(FAM, C.fmt, acc += chunk_compute<T>(P, C, R.m, zbuf, wave, lane))
    }
    return acc;
}

} // namespace hmvm
