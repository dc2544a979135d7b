// mrhs_dev.hpp -- device routines of the multi-right-hand-side forward
// product on the matrix cores, shared by the per-plan kernels (mrhs.hip)
// and the batched kernels (batch.hip).
//
// Lane layout of the 16x16x4 tiles (checked on the device):
//   A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
//   f32: D[i][j] in lane j + 16 (i / 4), element i % 4
//   f64: D[i][j] in lane j + 16 (i % 4), element i / 4
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace hmvm {

constexpr int kThr = 256;
constexpr int kKB = 32;              // columns per LDS block of the R pass
constexpr int kMaxRp = 64;           // right-hand sides per pass (tiles of 16)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mm;
template <> struct Mm<double> {
    using A = double; using V4 = d4;
    static constexpr int F = F64; static constexpr int W = 2;
    __device__ static inline double get(const double* p, int64_t i) { return p[i]; }
    __device__ static inline V4 mma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // tile row of element e held by lanes with lane / 16 == kk
    __device__ static inline int drow(int kk, int e) { return kk + 4 * e; }
};
template <> struct Mm<float> {
    using A = float; using V4 = f4;
    static constexpr int F = F32; static constexpr int W = 4;
    __device__ static inline float get(const float* p, int64_t i) { return p[i]; }
    __device__ static inline V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int drow(int kk, int e) { return 4 * kk + e; }
};
template <> struct Mm<uint16_t> {
    using A = float; using V4 = f4;
    static constexpr int F = BF16; static constexpr int W = 8;
    __device__ static inline float get(const uint16_t* p, int64_t i) { return __uint_as_float(unsigned(p[i]) << 16); }
    __device__ static inline V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int drow(int kk, int e) { return 4 * kk + e; }
};
template <> struct Mm<uint8_t> {
    using A = float; using V4 = f4;
    static constexpr int F = FP8; static constexpr int W = 16;
    __device__ static inline float get(const uint8_t* p, int64_t i) {
        return __builtin_amdgcn_cvt_f32_fp8(int(p[i]), 0);
    }
    __device__ static inline V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int drow(int kk, int e) { return 4 * kk + e; }
};

template <typename T> struct IsF8 { static constexpr bool v = false; };
template <> struct IsF8<uint8_t> { static constexpr bool v = true; };

// ---------------------------------------------------------------------------
// V pass: one workgroup per segment, out(c, r) = sum_i V(i, c) X(x0 + i, r);
// the 4 waves split the 16-column tiles, x slice (64 x Rp) shared in LDS
// ---------------------------------------------------------------------------

template <typename T, typename XT>
__device__ inline void vseg_mm_wg(const DevPlan& P, const VItem& sg, const XT* __restrict__ X, int64_t ldx, int R,
                                  double* __restrict__ Pb, double* xsraw, int tid, int wave, int lane) {
    using M = Mm<T>;
    using A = typename M::A;
    constexpr int W = M::W;
    const ColPanel C = P.cp[sg.q];
    const int m = C.m;
    const int Rp = (R + 15) & ~15;
    A* xs = reinterpret_cast<A*>(xsraw);                 // [64][Rp]
    for (int idx = tid; idx < 64 * Rp; idx += kThr) {
        const int i = idx / Rp, r = idx % Rp;
        xs[idx] = (i < m && r < R) ? A(X[int64_t(r) * ldx + C.col0 + i]) : A(0);
    }
    __syncthreads();
    const int ncol = sg.ng * W;
    const int64_t ld = ncol;
    const T* base = static_cast<const T*>(P.arena[M::F]) + sg.data_off;
    const int64_t vc0 = C.vcol0[M::F] + int64_t(sg.g0) * W;
    const int ci = lane & 15, kk = lane >> 4;
    for (int c0 = 16 * wave; c0 < ncol; c0 += 16 * 4) {
        A a[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int i = 4 * s + kk, c = c0 + ci;
            a[s] = (i < m && c < ncol) ? A(M::get(base, int64_t(i) * ld + c)) : A(0);
        }
        for (int r0 = 0; r0 < Rp; r0 += 16) {
            typename M::V4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const A b = xs[(4 * s + kk) * Rp + r0 + ci];
                acc = M::mma(a[s], b, acc);
            }
            const int r = r0 + ci;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = c0 + M::drow(kk, e);
                if (c < ncol && r < R) {
                    const int32_t d = P.vdest[vc0 + c];
                    double v = double(acc[e]);
                    if (IsF8<T>::v) v *= double(P.vscale[vc0 + c]);
                    if (d >= 0) Pb[int64_t(d) * R + r] = v;
                }
            }
        }
    }
}

template <typename XT>
__device__ inline void vseg_mm_dispatch(const DevPlan& P, const VItem& sg, const XT* __restrict__ X, int64_t ldx,
                                        int R, double* __restrict__ Pb, double* xs, int tid, int wave, int lane) {
    switch (sg.fmt) {
    case F64: vseg_mm_wg<double>(P, sg, X, ldx, R, Pb, xs, tid, wave, lane); break;
    case F32: vseg_mm_wg<float>(P, sg, X, ldx, R, Pb, xs, tid, wave, lane); break;
    case BF16: vseg_mm_wg<uint16_t>(P, sg, X, ldx, R, Pb, xs, tid, wave, lane); break;
    default: vseg_mm_wg<uint8_t>(P, sg, X, ldx, R, Pb, xs, tid, wave, lane); break;
    }
}

// ---------------------------------------------------------------------------
// partial sums for R right-hand sides (entry-major, rhs innermost)
// ---------------------------------------------------------------------------

__device__ inline void psum1_item(const DevPlan& D, const double* __restrict__ Pb, double* __restrict__ tmp, int R,
                                  int64_t g) {
    const int64_t s = g / R;
    const int r = int(g % R);
    const int32_t e = D.ps_e[s], r0 = D.ps_r0[s];
    const int32_t o = D.pt_off[e], k = D.tstride[e];
    const int32_t r1 = min(r0 + kSegLen, D.pt_cnt[e]);
    double a = 0;
    for (int32_t j = r0; j < r1; ++j) a += Pb[(int64_t(o) + int64_t(j) * k) * R + r];
    tmp[s * R + r] = a;
}

__device__ inline void psum2_item(const DevPlan& D, const double* __restrict__ tmp, double* __restrict__ t, int R,
                                  double alpha, int64_t g) {
    const int64_t e = g / R;
    const int r = int(g % R);
    const int32_t p = D.ps_pos0[e], n = D.ps_nseg[e], k = D.tstride[e];
    double a = 0;
    for (int32_t j = 0; j < n; ++j) a += tmp[(int64_t(p) + int64_t(j) * k) * R + r];
    t[e * R + r] = alpha * D.tmul[e] * a;
}

// ---------------------------------------------------------------------------
// R pass: one workgroup per row panel, wave w owns rows 16w .. 16w + 15
// ---------------------------------------------------------------------------

template <typename T, typename XT>
__device__ void rchunk_mm(const DevPlan& P, const RChunk& Ch, int m, double alpha, const XT* __restrict__ X,
                          int64_t ldx, const double* __restrict__ t, int R, int Rp, void* zraw,
                          typename Mm<T>::V4 (&acc)[kMaxRp / 16], int tid, int wave, int lane) {
    using M = Mm<T>;
    using A = typename M::A;
    constexpr int W = M::W;
    A* zs = static_cast<A*>(zraw);                       // [kKB][Rp]
    const int ncol = Ch.ngroups * W;
    const T* base = static_cast<const T*>(P.arena[M::F]) + Ch.data_off;
    const int row = 16 * wave + (lane & 15);
    const int kk = lane >> 4;
    for (int k0 = 0; k0 < ncol; k0 += kKB) {
        __syncthreads();
        for (int idx = tid; idx < kKB * Rp; idx += kThr) {
            const int c = k0 + idx / Rp, r = idx % Rp;
            double z = 0;
            if (c < ncol && r < R) {
                const int32_t s = P.colsrc[Ch.col0 + c];
                z = (s >= 0) ? alpha * double(X[int64_t(r) * ldx + s]) : (s == kPadCol ? 0.0 : t[int64_t(~s) * R + r]);
                if (IsF8<T>::v) z *= double(P.zscale[Ch.col0 + c]);
            }
            zs[idx] = A(z);
        }
        __syncthreads();
        // A operands of the whole block first (independent loads), then the MFMAs
        A a[kKB / 4];
#pragma unroll
        for (int s = 0; s < kKB / 4; ++s) {
            const int c = k0 + 4 * s + kk;
            a[s] = (row < m && c < ncol) ? A(M::get(base, int64_t(c / W) * m * W + int64_t(row) * W + (c % W)))
                                         : A(0);
        }
#pragma unroll
        for (int s = 0; s < kKB / 4; ++s) {
            for (int nt = 0; nt < Rp / 16; ++nt) {
                const A b = zs[(4 * s + kk) * Rp + nt * 16 + (lane & 15)];
                acc[nt] = M::mma(a[s], b, acc[nt]);
            }
        }
    }
}

// whole row panel, Y(:, 0..R) += R_p Z; all kThr threads call it
template <typename XT, typename YT>
__device__ inline void rpanel_mm(const DevPlan& P, const RowPanel& Rw, double alpha, const XT* __restrict__ X,
                                 int64_t ldx, const double* __restrict__ t, YT* __restrict__ Y, int64_t ldy, int R,
                                 double* zbuf, int tid, int wave, int lane) {
    const int Rp = (R + 15) & ~15;
    d4 acc64[kMaxRp / 16];
    f4 acc32[kMaxRp / 16];
#pragma unroll
    for (int i = 0; i < kMaxRp / 16; ++i) { acc64[i] = d4{0, 0, 0, 0}; acc32[i] = f4{0, 0, 0, 0}; }
    for (int ci = 0; ci < Rw.nchunks; ++ci) {
        const RChunk Ch = P.rchunks[Rw.chunk0 + ci];
        switch (Ch.fmt) {
        case F64: rchunk_mm<double>(P, Ch, Rw.m, alpha, X, ldx, t, R, Rp, zbuf, acc64, tid, wave, lane); break;
        case F32: rchunk_mm<float>(P, Ch, Rw.m, alpha, X, ldx, t, R, Rp, zbuf, acc32, tid, wave, lane); break;
        case BF16: rchunk_mm<uint16_t>(P, Ch, Rw.m, alpha, X, ldx, t, R, Rp, zbuf, acc32, tid, wave, lane); break;
        default: rchunk_mm<uint8_t>(P, Ch, Rw.m, alpha, X, ldx, t, R, Rp, zbuf, acc32, tid, wave, lane); break;
        }
    }
    // bring the FP64 tiles into the FP32 element order (row 4 kk + e), then
    // one writer per output element: row 16 wave + 4 (lane / 16) + e
    const int kk = lane >> 4;
    for (int nt = 0; nt < Rp / 16; ++nt) {
        double v64[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // FP64 element (row 4 kk + e) lives in lane (lane % 16) + 16 e, element kk
            const int src = (lane & 15) + 16 * e;
            const double a0 = __shfl(acc64[nt][0], src, 64), a1 = __shfl(acc64[nt][1], src, 64);
            const double a2 = __shfl(acc64[nt][2], src, 64), a3 = __shfl(acc64[nt][3], src, 64);
            v64[e] = (kk == 0) ? a0 : (kk == 1) ? a1 : (kk == 2) ? a2 : a3;
        }
        const int r = nt * 16 + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = 16 * wave + 4 * kk + e;
            if (i < Rw.m && r < R) {
                YT* yp = Y + int64_t(r) * ldy + Rw.row0 + i;
                *yp = YT(double(*yp) + v64[e] + double(acc32[nt][e]));
            }
        }
    }
}

} // namespace hmvm
