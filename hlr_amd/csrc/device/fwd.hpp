// fwd.hpp -- device routines of the forward product shared by the
// per-plan kernels (hmvm.hip) and the batched kernels (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace hmvm {

constexpr int kThreads = 64 * kWaves;
constexpr int64_t kWholeZBytes = 48 * 1024;   // largest panel z block staged whole in LDS
constexpr int kXsStride = 65;                 // LDS row stride of the x slices of a V pack

template <typename T> struct IsFp8 { static constexpr bool value = false; };
template <> struct IsFp8<uint8_t> { static constexpr bool value = true; };

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
    // x slices of the segments, row stride kXsStride (odd): the lanes of
    // different segments read different rows i of their own slice in the
    // same step, which then fall into different LDS banks
    A* xs = static_cast<A*>(xsraw);                 // [kPackSegs][kXsStride]
    const int nseg = pk.nseg;
    for (int idx = lane; idx < nseg * 64; idx += 64) {
        const int j = idx >> 6, i = idx & 63;
        if (i < pk.m[j]) xs[j * kXsStride + i] = A(x[pk.x0[j] + i]);
    }
    wave_sync();
    if (lane >= pk.gofs[nseg]) return;
    int j = 0;
    while (j + 1 < nseg && lane >= pk.gofs[j + 1]) ++j;
    const int g = lane - pk.gofs[j];
    const int m = pk.m[j];
    const int64_t ld = int64_t(pk.ng[j]) * W;
    const T* base = static_cast<const T*>(P.arena[E::F]) + pk.data_off[j] + int64_t(g) * W;
    const A* xj = xs + j * kXsStride;
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    constexpr int kUnroll = (W >= 8) ? 4 : 16;      // rows per unrolled step (measured best)
#pragma unroll kUnroll
    for (int i = 0; i < m; ++i) {
        A v[W];
        E::load(base + int64_t(i) * ld, v);
        const A xi = xj[i];
#pragma unroll
        for (int w = 0; w < W; ++w) acc[w] += v[w] * xi;
    }
    const int64_t vc0 = int64_t(pk.vc0[j]) + int64_t(g) * W;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int32_t d = P.vdest[vc0 + w];
        double v = double(acc[w]);
        if (IsFp8<T>::value) v *= double(P.vscale[vc0 + w]);
        if (d >= 0) Pb[d] = v;
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
    for (int j = tid; j < ncol; j += kThreads) {
        const int32_t s = P.colsrc[C.col0 + j];
        double z = (s >= 0) ? alpha * double(x[s]) : (s == kPadCol ? 0.0 : t[~s]);
        if (IsFp8<T>::value) z *= double(P.zscale[C.col0 + j]);
        zs[j] = A(z);
    }
}

template <typename T>
__device__ inline double chunk_compute(const DevPlan& P, const RChunk& C, int m, const void* zbuf, int wave,
                                       int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    const A* zs = static_cast<const A*>(zbuf);
    const int per = (C.ngroups + kWaves - 1) / kWaves;
    const int gb = wave * per;
    const int ge = min(gb + per, int(C.ngroups));
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    if (lane < m) {
        const T* base = static_cast<const T*>(P.arena[E::F]) + C.data_off + int64_t(lane) * W;
        const int64_t gs = int64_t(m) * W;
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

// kWhole: the panel's z values fit in dynamic LDS and are staged once;
// otherwise each chunk is staged into a fixed buffer between barriers.

// Whole row panel with chunk-by-chunk staging into a buffer of
// kChunkGroups * 16 floats; all kThreads threads call it; returns the
// lane's partial row sum (rows = lanes, waves split each chunk).
template <typename XT>
__device__ inline double rpanel_fwd_chunked(const DevPlan& P, const RowPanel& R, double alpha,
                                            const XT* __restrict__ x, const double* __restrict__ t, void* zbuf,
                                            int tid, int wave, int lane) {
    double acc = 0;
    for (int ci = 0; ci < R.nchunks; ++ci) {
        const RChunk C = P.rchunks[R.chunk0 + ci];
        __syncthreads();
        switch (C.fmt) {
        case F64: chunk_stage<double>(P, C, alpha, x, t, zbuf, tid); break;
        case F32: chunk_stage<float>(P, C, alpha, x, t, zbuf, tid); break;
        case BF16: chunk_stage<uint16_t>(P, C, alpha, x, t, zbuf, tid); break;
        default: chunk_stage<uint8_t>(P, C, alpha, x, t, zbuf, tid); break;
        }
        __syncthreads();
        switch (C.fmt) {
        case F64: acc += chunk_compute<double>(P, C, R.m, zbuf, wave, lane); break;
        case F32: acc += chunk_compute<float>(P, C, R.m, zbuf, wave, lane); break;
        case BF16: acc += chunk_compute<uint16_t>(P, C, R.m, zbuf, wave, lane); break;
        default: acc += chunk_compute<uint8_t>(P, C, R.m, zbuf, wave, lane); break;
        }
    }
    return acc;
}

} // namespace hmvm
