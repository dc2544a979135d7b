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

template <typename T, typename XT>
__device__ inline void vpack_fwd(const DevPlan& P, const VPack* __restrict__ pkp, const XT* __restrict__ x,
                                 double* __restrict__ Pb, void* xsraw, int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    const VPack& pk = *pkp;     // read through the pointer: the arrays stay out of scratch
    // x slices of the segments (up to kUnitRows values each), an odd
    // distance apart: the lanes of different segments read different rows i
    // of their own slice in the same step, which then fall into different
    // LDS banks
    A* xs = static_cast<A*>(xsraw);                 // kPackXs values: segment j's slice at xoff[j]
    const int nseg = pk.nseg;
    for (int j = 0; j < nseg; ++j) {
        const int mj = pk.m[j], o = pk.xoff[j], x0 = pk.x0[j];
        for (int i = lane; i < mj; i += 64) xs[o + i] = A(x[x0 + i]);
    }
    wave_sync();
    const int ngr = pk.gofs[nseg];                  // groups of the pack = active lanes
    int j = 0;
    while (j + 1 < nseg && lane >= pk.gofs[j + 1]) ++j;
    const int g = lane - pk.gofs[j];
    const int m = (lane < ngr) ? pk.m[j] : 0;       // lanes past the pack's groups load nothing
    const int64_t ldb = int64_t(pk.ng[j]) * 16;                 // bytes per row of the segment
    const bytep base = elem_addr<T>(P.arena[E::F], pk.data_off[j]) + int64_t(g) * 16;
    const A* xj = xs + pk.xoff[j];
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    constexpr int kUnroll = (W >= 8) ? 4 : 16;      // rows per unrolled step (measured best)
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
    // index is vc0[s] - W gofs[s] + c, selected in registers from the pack's
    // (uniform) segment bounds, so that all W table loads of a lane are
    // issued together (no load waits on another).
    const int ncol = ngr * W;
    int32_t vd[W];
    float vs[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int c = lane + 64 * k;
        const int gr = c / W;
        int64_t vb = pk.vc0[0] - pk.gofs[0] * W;
#pragma unroll
        for (int t = 1; t < kPackSegs; ++t)
            if (t < nseg && gr >= pk.gofs[t]) vb = pk.vc0[t] - pk.gofs[t] * W;
        const bool on = c < ncol;
        vd[k] = on ? P.vdest[int64_t(vb) + c] : -1;
        vs[k] = (E::kScaled && on) ? P.vscale[int64_t(vb) + c] : 1.0f;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int c = lane + 64 * k;
        if (vd[k] >= 0) Pb[vd[k]] = double(res[c]) * double(vs[k]);
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
