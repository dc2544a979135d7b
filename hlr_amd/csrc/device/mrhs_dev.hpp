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
constexpr int kMaxRp = 64;           // right-hand sides per pass (tiles of 16)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// every format but FP64 is decoded to FP32 and uses the FP32 matrix cores
template <typename T> struct Mm {
    using A = float; using V4 = f4;
    static constexpr int F = Elem<T>::F; static constexpr int W = Elem<T>::W;
    __device__ static inline V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int drow(int kk, int e) { return 4 * kk + e; }
};
template <> struct Mm<double> {
    using A = double; using V4 = d4;
    static constexpr int F = F64; static constexpr int W = 2;
    __device__ static inline V4 mma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // tile row of element e held by lanes with lane / 16 == kk
    __device__ static inline int drow(int kk, int e) { return kk + 4 * e; }
};

// MFMA steps whose factor loads are issued together: 8 for FP64 / FP32,
// fewer for the narrow formats, whose loads each feed more MFMAs
template <int W> struct MmSteps { static constexpr int value = (W <= 4) ? 8 : (W <= 8 ? 4 : 2); };

// V pass accumulator sets per segment pass: the largest divisor of W that
// is at most 5 (W / JB passes over the segment)
constexpr int mm_jb(int W) { return W % 5 == 0 ? 5 : W % 4 == 0 ? 4 : W % 3 == 0 ? 3 : W % 2 == 0 ? 2 : 1; }

// ---------------------------------------------------------------------------
// V pass: one workgroup per segment, out(c, r) = sum_i V(i, c) X(x0 + i, r),
// computed as out^T = X^T V: M = right-hand sides, N = columns, K = rows.
// Wave w takes groups 16 w .. 16 w + 15 of the segment; lane n + 16 k loads
// row 4 s + k of group 16 w + n with one 16-byte load, whose element j is
// B[k][n] of MFMA j (column W (16 w + n) + j).  A[m][k] = X(row 4 s + k,
// rhs m) comes from the x slice in LDS (64 rows x mm_stride<NTM>()).
// Accumulators: JB elements x up to 4 rhs tiles; formats with W > JB make
// W / JB passes over the segment.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// XT = X^T, row-major n x R: the right-hand sides of one index are
// contiguous, so the passes below stage x and z values with the rhs index
// fastest and every read is a contiguous run of R values.  One workgroup
// transposes 64 indices through LDS (reads of X coalesced along the index,
// writes of XT coalesced along the rhs).
// ---------------------------------------------------------------------------

template <typename XT>
__device__ inline void xtrans_tile(const XT* __restrict__ X, int64_t ldx, int64_t n, int R, double* __restrict__ XTb,
                                   int64_t i0, double* tile, int tid) {
    const int TS = ((R + 15) & ~15) + 1;                 // [64][TS]: index, rhs (LDS of mm_lds_bytes(Rp))
    for (int idx = tid; idx < 64 * R; idx += kThr) {
        const int i = idx & 63, r = idx >> 6;
        tile[i * TS + r] = (i0 + i < n) ? double(X[int64_t(r) * ldx + i0 + i]) : 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < 64 * R; idx += kThr) {
        const int i = idx / R, r = idx - i * R;
        if (i0 + i < n) XTb[(i0 + i) * R + r] = tile[i * TS + r];
    }
}

// LDS of one matrix-core workgroup: 64 rows (V pass) or 64 columns (R pass)
// of Rp = 16 NTM right-hand sides, with the row stride 16 NTM + 1 (odd, so
// that neighbouring rows start in different banks; a compile-time constant,
// so that every operand read is one base register plus an immediate
// offset).  Sized by the tile count, not by the largest one: a workgroup of
// one or two tiles takes a quarter or half of the LDS of four, and more of
// them share a CU (32 right-hand sides, mixed plan: 2.05 -> 1.80 ms).
template <int NTM> constexpr int mm_stride() { return 16 * NTM + 1; }
// Waves per SIMD the R-pass kernel is compiled for: 5 for the FP64-only
// kernel with one tile, whose unconditional staging loads took it from 88
// to 92 VGPRs and from 5 to 4 waves (16 right-hand sides 1.40 -> 1.51 ms);
// at 5 it fits in 96 without spilling.  Bounds that would restore the
// four-tile kernels' occupancy, or the batch's one-tile kernel's, spill.
template <int NTM, int FAM> constexpr int mm_min_waves() {
    return (NTM == 1 && FAM == kFamF64) ? 5 : 1;
}
__host__ __device__ inline size_t mm_lds_bytes(int Rp) { return size_t(64) * (Rp + 1) * sizeof(double); }

template <int NTM, typename T, bool kBatchedStage>
__device__ inline void vseg_mm_wg(const DevPlan& P, const VItem& sg, const double* __restrict__ XTb, int R,
                                  double* __restrict__ Pb, double* xsraw, int tid, int wave, int lane) {
    using M = Mm<T>;
    using A = typename M::A;
    using E = Elem<T>;
    constexpr int W = M::W;
    constexpr int JB = mm_jb(W);
    constexpr int NB = MmSteps<W>::value;
    const int m = sg.m;                                  // up to kUnitRows, staged 64 rows at a time
    const int Rp = (R + 15) & ~15;
    const int NT = Rp / 16;
    constexpr int XS = mm_stride<NTM>();
    A* xs = reinterpret_cast<A*>(xsraw);                 // [64][XS]: row, rhs
    constexpr int RS = 16 * NTM;                         // staged rhs slots (constant divisor)
    // rows i0 .. i0 + 63 of the segment's x slice (rhs fastest: contiguous
    // reads of XT).  kBatchedStage: four values per thread at a time, their
    // loads unconditional (XT entry 0 for padding, masked afterwards) and
    // all issued before the LDS stores; loaded one at a time under a
    // divergent condition, each value waited for the one before (64
    // right-hand sides, FP64 plan 3.15 -> 2.96 ms).  The batch's V-pass
    // kernel keeps the one-at-a-time form: there the batched form took the
    // FP64 four-tile kernel from 104 to 248 VGPRs, with spills.
    static_assert((64 * RS) % (4 * kThr) == 0, "whole rounds of four values per thread");
    auto stage = [&](int i0) {
        const int mb = min(64, m - i0);
        if constexpr (kBatchedStage) {
#pragma unroll 1
            for (int q0 = 0; q0 < 64 * RS; q0 += 4 * kThr) {
                A v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int idx = q0 + q * kThr + tid;
                    const int i = idx / RS, r = idx - i * RS;
                    const bool ok = i < mb && r < R;
                    const A a = A(XTb[ok ? (sg.x0 + i0 + i) * R + r : 0]);
                    v[q] = ok ? a : A(0);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int idx = q0 + q * kThr + tid;
                    const int i = idx / RS, r = idx - i * RS;
                    if (r < Rp) xs[i * XS + r] = v[q];
                }
            }
        } else {
            for (int idx = tid; idx < 64 * RS; idx += kThr) {
                const int i = idx / RS, r = idx - i * RS;
                if (r < Rp) xs[i * XS + r] = (i < mb && r < R) ? A(XTb[(sg.x0 + i0 + i) * R + r]) : A(0);
            }
        }
    };
    const bool single = m <= 64;                         // one block: staged once for every pass
    if (single) stage(0);
    __syncthreads();
    // waves past the segment's groups keep taking part in the staging barriers
    const bool wactive = 16 * wave < sg.ng;
    const int n = lane & 15, kk = lane >> 4;
    const int gl = 16 * wave + n;
    const bool gon = gl < sg.ng;
    const int64_t ldb = int64_t(sg.ng) * 16;                // bytes per row of the segment
    const bytep base = elem_addr<T>(P.arena[M::F], sg.data_off) + int64_t(gon ? gl : 0) * 16;
    const int64_t vc0 = sg.vc0;
    for (int jp = 0; jp < W; jp += JB) {
        typename M::V4 acc[JB][NTM];
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int nt = 0; nt < NTM; ++nt) acc[j][nt] = typename M::V4{0, 0, 0, 0};
        for (int i0 = 0; i0 < m; i0 += 64) {
            if (!single) {
                __syncthreads();                         // the previous block's reads are done
                stage(i0);
                __syncthreads();
            }
            const int mb = min(64, m - i0);
            if (!wactive) continue;
            // NB steps of 4 rows: their loads are all in flight before the MFMAs
            for (int s0 = 0; 4 * s0 < mb; s0 += NB) {
                u32x4 raw[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int row = 4 * (s0 + b) + kk;
                    raw[b] = (gon && row < mb) ? raw16(base + int64_t(i0 + row) * ldb) : u32x4{0u, 0u, 0u, 0u};
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (4 * (s0 + b) >= mb) break;       // uniform
                    const int row = 4 * (s0 + b) + kk;
                    A v[W];
                    E::decode(raw[b], v);
#pragma unroll
                    for (int nt = 0; nt < NTM; ++nt) {
                        if (nt < NT) {
                            const A a = xs[row * XS + nt * 16 + n];
#pragma unroll
                            for (int j = 0; j < JB; ++j) acc[j][nt] = M::mma(a, v[jp + j], acc[j][nt]);
                        }
                    }
                }
            }
        }
        // D[mr][n] of MFMA j: rhs nt * 16 + mr, column W gl + jp + j
        if (gon) {
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                const int64_t vc = vc0 + int64_t(gl) * W + jp + j;
                const int32_t d = P.vdest[vc];
                if (d < 0) continue;
                const double sc = E::kScaled ? double(P.vscale[vc]) : 1.0;
#pragma unroll
                for (int nt = 0; nt < NTM; ++nt) {
                    if (nt < NT) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int r = nt * 16 + M::drow(kk, e);
                            if (r < R) Pb[int64_t(d) * R + r] = double(acc[j][nt][e]) * sc;
                        }
                    }
                }
            }
        }
    }
}

// NTM: tiles of 16 right-hand sides the accumulators are sized for
// (compile time); the call handles ceil(R / 16) <= NTM of them
template <int NTM, int FAM = kFamAll, bool kBatchedStage = true>
__device__ inline void vseg_mm_dispatch(const DevPlan& P, const VItem& sg, const double* __restrict__ XTb,
                                        int R, double* __restrict__ Pb, double* xs, int tid, int wave, int lane) {
    HMVM_FMT_SWITCH_FAM(FAM, sg.fmt, vseg_mm_wg<NTM, T, kBatchedStage>(P, sg, XTb, R, Pb, xs, tid, wave, lane))
}

// ---------------------------------------------------------------------------
// partial sums for R right-hand sides (entry-major, rhs innermost)
// ---------------------------------------------------------------------------

// t(e, r .. r + NP - 1) = alpha S_e sum of the entry's partials, for NP = 1
// or 2 adjacent right-hand sides (NP = 2: R even, one 16-byte load per
// partial; P and t are 16-byte aligned and o R + r is even).  Each segment
// of kSegLen partials is summed in order from zero, then the segment sums in
// order: a fixed order, the same for every schedule, batch and NI.
//
// A thread takes up to NI items g0, g0 + stride, ... (those below g1): the
// entry tables and the first partial of all of them are loaded together,
// then each item's remaining partials in rounds of B loads.
template <int NP>
struct PsumVec {
    double v[NP];
};

template <int NP>
__device__ inline PsumVec<NP> psum_load(const double* p) {
    PsumVec<NP> r;
    if constexpr (NP == 2) {
        const double2 d = *reinterpret_cast<const double2*>(p);
        r.v[0] = d.x;
        r.v[1] = d.y;
    } else {
        r.v[0] = *p;
    }
    return r;
}

template <int NP, int NI, int B>
__device__ inline void psum_mm_items(const DevPlan& D, const double* __restrict__ Pb, double* __restrict__ t, int R,
                                     double alpha, int64_t g0, int64_t stride, int64_t g1) {
    const int h = R / NP;                                // items per entry
    int64_t e[NI];
    int r[NI];
    int32_t o[NI], k[NI], cnt[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        const int64_t g = g0 + j * stride;
        const bool ok = g < g1;
        e[j] = ok ? g / h : 0;
        r[j] = ok ? NP * int(g - e[j] * h) : 0;
        o[j] = D.pt_off[e[j]];
        k[j] = D.tstride[e[j]];
        cnt[j] = ok ? D.pt_cnt[e[j]] : 0;
    }
    PsumVec<NP> first[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j)
        first[j] = (cnt[j] > 0) ? psum_load<NP>(Pb + int64_t(o[j]) * R + r[j]) : PsumVec<NP>{};
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        if (g0 + j * stride >= g1) break;
        const double* src = Pb + int64_t(o[j]) * R + r[j];
        const int64_t step = int64_t(k[j]) * R;
        double a[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) a[q] = 0;
        for (int32_t r0 = 0; r0 < cnt[j]; r0 += kSegLen) {
            const int32_t n = min(kSegLen, cnt[j] - r0);
            double s[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) s[q] = 0;
            int32_t b = 0;
            if (r0 == 0) {
#pragma unroll
                for (int q = 0; q < NP; ++q) s[q] += first[j].v[q];
                b = 1;
            }
            for (; b < n; b += B) {
                PsumVec<NP> v[B];
#pragma unroll
                for (int u = 0; u < B; ++u)
                    v[u] = (b + u < n) ? psum_load<NP>(src + (int64_t(r0) + b + u) * step) : PsumVec<NP>{};
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (b + u < n) {
#pragma unroll
                        for (int q = 0; q < NP; ++q) s[q] += v[u].v[q];
                    }
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) a[q] += s[q];
        }
        const double f = alpha * D.tmul[e[j]];
        if constexpr (NP == 2)
            reinterpret_cast<double2*>(t + e[j] * R + r[j])[0] = double2{f * a[0], f * a[1]};
        else
            t[e[j] * R + r[j]] = f * a[0];
    }
}

// The sums of one plan take one item per thread, partials in rounds of 8
// (pairs) or of a whole segment (single right-hand sides).  The batch's sum
// units take up to four items per thread, partials in rounds of 4: its
// entries mostly have one or two partials, and with one item per thread each
// thread waited on two dependent load round trips per item (64 matrices,
// step 12.5 -> 12.1 ms).  In a plan of many right-hand sides, four items per
// thread made the sums slower (16 right-hand sides, FP64 1.40 -> 1.43 ms,
// mixed 1.06 -> 1.10 ms): fewer waves share the entries' longer tails.
template <int NP>
constexpr int kPsumRound = NP == 2 ? 8 : kSegLen;
constexpr int kBsumItems = 4;
constexpr int kBsumRound = 4;


// ---------------------------------------------------------------------------
// R pass: one workgroup per row panel, wave w owns rows 16w .. 16w + 15
// ---------------------------------------------------------------------------

// R pass of one chunk: Y(i, r) += sum_c R(i, c) Z(c, r) with M = rows (16
// per wave), N = right-hand sides, K = columns.  Z is staged 64 columns at
// a time in LDS (64 x mm_stride<NTM>()); lane i + 16 k loads row 16 w + i of group
// g0 + k with one 16-byte load, whose element j is A[i][k] of MFMA j
// (column W (g0 + k) + j), and B[k][n] = Z(W (g0 + k) + j, rhs n).
template <int NTM, typename T, int FAM>
__device__ void rchunk_mm(const DevPlan& P, const RChunk& Ch, int m, double alpha, const double* __restrict__ XTb,
                          const double* __restrict__ t, int R, int Rp, void* zraw,
                          typename Mm<T>::V4 (&acc)[NTM], int tid, int wave, int lane) {
    using M = Mm<T>;
    using A = typename M::A;
    using E = Elem<T>;
    constexpr int W = M::W;
    constexpr bool kUnroll = (FAM == kFamF64);           // the FP64-only kernel
    // groups per staged block: a multiple of the 4 groups of one MFMA step,
    // so that no step reads past the staged columns; the LDS of 64 FP64
    // columns holds 128 FP32 ones (formats decoded to FP32)
    constexpr int kMaxCols = 64 * int(sizeof(double) / sizeof(A));
    constexpr int GB = (kMaxCols / W) & ~3;
    constexpr int SC = GB * W;                           // staged columns per block
    static_assert(GB >= 4, "a staged block holds at least one MFMA step");
    constexpr int NB = MmSteps<W>::value;
    constexpr int RS = 16 * NTM;                         // staged rhs slots (constant divisor)
    A* zs = static_cast<A*>(zraw);                       // [64][ZS]: column, rhs
    constexpr int ZS = mm_stride<NTM>();
    const int NT = Rp / 16;
    const int ncol = Ch.ngroups * W;
    const int row = 16 * wave + (lane & 15);
    const int kk = lane >> 4, n = lane & 15;
    const bytep base = elem_addr<T>(P.arena[M::F], Ch.data_off) + int64_t(row < m ? row : 0) * 16;
    // The z values of a staged block (rhs fastest: contiguous reads of XT
    // and t).  Every load of a round is unconditional, at a safe index
    // (the block's first column, XT row 0) for padding, and masked
    // afterwards: a load under a divergent condition ended its block with a
    // wait, which had serialised the rounds load by load.  When RS divides kThr, a thread keeps one rhs index r and
    // stages columns c0 + j kThr / RS in rounds of ZB: a round issues its
    // column-source loads, then its value loads, so that a block waits for
    // a few load round trips rather than two per value.  Otherwise (RS = 48)
    // a plain loop, which needs fewer registers there.
    constexpr bool kFixedR = (kThr % RS) == 0;
    constexpr int NI = (SC * RS + kThr - 1) / kThr;
    // values per round: up to 16 (one or two rounds per block) in the
    // FP64-only kernel, which has the registers for them (batch step 13.0 ->
    // 12.5 ms); 4 in the kernels of the other families, whose registers
    // are those of their widest decode path and which lose occupancy with
    // more (32 right-hand sides, mixed plan 1.81 -> 2.11 ms at 16)
    constexpr int ZB = NI < (kUnroll ? 16 : 4) ? NI : (kUnroll ? 16 : 4);
    const int c0 = tid / RS, r0 = tid % RS;
    for (int gb0 = 0; gb0 < Ch.ngroups; gb0 += GB) {
        const int64_t cb = int64_t(Ch.col0) + gb0 * W;
        const int lim = min(SC, ncol - gb0 * W);         // columns of this block that exist
        __syncthreads();                                 // the MFMAs of the block before are done with zs
        // The other families, and the FP64-only kernel with two tiles, for
        // an even number of right-hand sides: a thread stages two
        // neighbouring right-hand sides of a column with
        // one 16-byte load (t and XT rows are 16-byte aligned then), so a
        // block takes half the rounds: 8 values per round for the hardware
        // formats, 4 for the packed ones, whose decode paths leave fewer
        // registers.  One load from the selected row (x or t) keeps the
        // registers of the scalar rounds (32 right-hand sides: mixed 1.81
        // -> 1.70 ms, block floating point 2.18 -> 1.93 ms, adaptive
        // floating point 2.30 -> 2.06 ms; FP64 1.95 -> 1.85 ms).  The
        // FP64-only kernels with one or four tiles keep their scalar rounds:
        // the pairs cost them a wave per SIMD (batch step 12.5 -> 12.7 ms).
        constexpr bool kPairs = !kUnroll || NTM == 2;
        constexpr int RS2 = RS / 2;                      // threads per column
        constexpr int NI2 = (SC * RS2 + kThr - 1) / kThr;
        constexpr int ZB2 = NI2 < (FAM == kFamHw ? 8 : 4) ? NI2 : (FAM == kFamHw ? 8 : 4);
        if (kPairs && (kThr % RS2) == 0 && (R % 2) == 0) {
            const int c2 = tid / RS2, r2 = 2 * (tid % RS2);
            auto round2 = [&](int j0) {
                int32_t src[ZB2];
                float sc[ZB2];
#pragma unroll
                for (int u = 0; u < ZB2; ++u) {
                    const int c = c2 + (j0 + u) * (kThr / RS2);
                    const bool ok = j0 + u < NI2 && c < lim;
                    const int64_t ci = cb + (ok ? c : 0);
                    const int32_t v = P.colsrc[ci];
                    src[u] = ok ? v : kPadCol;
                    if constexpr (E::kScaled) {
                        const float w = P.zscale[ci];
                        sc[u] = ok ? w : 0.f;
                    }
                }
                // all value loads of the round, then all stores
                double2 zv[ZB2];
#pragma unroll
                for (int u = 0; u < ZB2; ++u) {
                    const int32_t s = src[u];
                    const bool on = s != kPadCol && r2 < R;
                    // one load from a selected row: x (times alpha) or t;
                    // XT row 0 for padding
                    const double* src2 =
                        !on ? XTb : (s >= 0) ? XTb + int64_t(s) * R + r2 : t + int64_t(~s) * R + r2;
                    zv[u] = *reinterpret_cast<const double2*>(src2);
                }
#pragma unroll
                for (int u = 0; u < ZB2; ++u) {
                    const int c = c2 + (j0 + u) * (kThr / RS2);
                    const int32_t s = src[u];
                    const bool on = s != kPadCol && r2 < R;
                    double2 z = zv[u];
                    if (s >= 0) { z.x *= alpha; z.y *= alpha; }
                    if constexpr (E::kScaled) { z.x *= double(sc[u]); z.y *= double(sc[u]); }
                    if (!on) z = double2{0.0, 0.0};
                    if (j0 + u < NI2 && c < SC) {
                        zs[c * ZS + r2] = A(z.x);
                        zs[c * ZS + r2 + 1] = A(z.y);
                    }
                }
            };
#pragma unroll 1
            for (int j0 = 0; j0 < NI2; j0 += ZB2) round2(j0);
        } else if constexpr (kFixedR) {
            // one round: ZB column-source loads, then ZB value loads
            auto round = [&](int j0) {
                int32_t src[ZB];
                float sc[ZB];
#pragma unroll
                for (int u = 0; u < ZB; ++u) {
                    const int c = c0 + (j0 + u) * (kThr / RS);
                    const bool ok = j0 + u < NI && c < lim;
                    const int64_t ci = cb + (ok ? c : 0);
                    const int32_t v = P.colsrc[ci];
                    src[u] = ok ? v : kPadCol;
                    if constexpr (E::kScaled) {
                        const float w = P.zscale[ci];
                        sc[u] = ok ? w : 0.f;
                    }
                }
                // all value loads of the round, then all stores
                double zv[ZB];
#pragma unroll
                for (int u = 0; u < ZB; ++u) {
                    const int32_t s = src[u];
                    const bool on = s != kPadCol && r0 < R;
                    const double* p1 = !on ? XTb : (s >= 0) ? XTb + int64_t(s) * R + r0 : t + int64_t(~s) * R + r0;
                    zv[u] = *p1;
                }
#pragma unroll
                for (int u = 0; u < ZB; ++u) {
                    const int c = c0 + (j0 + u) * (kThr / RS);
                    const int32_t s = src[u];
                    const bool on = s != kPadCol && r0 < R;
                    const double v = zv[u];
                    double z = (s >= 0) ? alpha * v : v;
                    if constexpr (E::kScaled) z *= double(sc[u]);
                    if (!on) z = 0;
                    if (j0 + u < NI && c < SC) zs[c * ZS + r0] = A(z);
                }
            };
            // unrolled up to two rounds in the FP64-only kernel; elsewhere
            // in a loop, which keeps the registers of one round (the
            // kernel's registers are those of its largest format path)
            if constexpr (kUnroll && NI <= 2 * ZB) {
#pragma unroll
                for (int j0 = 0; j0 < NI; j0 += ZB) round(j0);
            } else {
#pragma unroll 1
                for (int j0 = 0; j0 < NI; j0 += ZB) round(j0);
            }
        } else {
            for (int idx = tid; idx < SC * RS; idx += kThr) {
                const int c = idx / RS, r = idx - c * RS;
                if (r >= Rp) continue;
                double z = 0;
                if (c < lim && r < R) {
                    const int32_t s = P.colsrc[cb + c];
                    z = (s >= 0) ? alpha * XTb[int64_t(s) * R + r] : (s == kPadCol ? 0.0 : t[int64_t(~s) * R + r]);
                    if constexpr (E::kScaled) z *= double(P.zscale[cb + c]);
                }
                zs[c * ZS + r] = A(z);
            }
        }
        __syncthreads();
        const int gend = min(gb0 + GB, int(Ch.ngroups));
        // NB steps of 4 groups: their loads are all in flight before the MFMAs
        for (int g0 = gb0; g0 < gend; g0 += 4 * NB) {
            u32x4 raw[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int g = g0 + 4 * b + kk;
                raw[b] = (row < m && g < gend) ? raw16(base + int64_t(g) * m * 16) : u32x4{0u, 0u, 0u, 0u};
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (g0 + 4 * b >= gend) break;           // uniform
                const int g = g0 + 4 * b + kk;
                A v[W];
                E::decode(raw[b], v);
                const A* zg = zs + ((g - gb0) * W) * ZS + n;
#pragma unroll
                for (int nt = 0; nt < NTM; ++nt) {
                    if (nt < NT) {
#pragma unroll
                        for (int j = 0; j < W; ++j) acc[nt] = M::mma(v[j], zg[j * ZS + nt * 16], acc[nt]);
                    }
                }
            }
        }
    }
}

// the accumulators of a format: FP64 tiles for FP64, FP32 tiles otherwise
template <typename T> struct RAcc {
    template <int N> __device__ static inline auto& pick(d4 (&)[N], f4 (&a32)[N]) { return a32; }
};
template <> struct RAcc<double> {
    template <int N> __device__ static inline auto& pick(d4 (&a64)[N], f4 (&)[N]) { return a64; }
};

// whole row panel, Y(:, 0..R) += R_p Z; all kThr threads call it
template <int NTM, typename YT, int FAM = kFamAll>
__device__ inline void rpanel_mm(const DevPlan& P, const RowPanel& Rw, double alpha, const double* __restrict__ XTb,
                                 const double* __restrict__ t, YT* __restrict__ Y, int64_t ldy, int R,
                                 double* zbuf, int tid, int wave, int lane) {
    const int Rp = (R + 15) & ~15;
    d4 acc64[NTM];
    f4 acc32[NTM];
#pragma unroll
    for (int i = 0; i < NTM; ++i) { acc64[i] = d4{0, 0, 0, 0}; acc32[i] = f4{0, 0, 0, 0}; }
    for (int ci = 0; ci < Rw.nchunks; ++ci) {
        const RChunk Ch = P.rchunks[Rw.chunk0 + ci];
        HMVM_FMT_SWITCH_FAM(FAM, Ch.fmt, rchunk_mm<NTM, T, FAM>(P, Ch, Rw.m, alpha, XTb, t, R, Rp, zbuf,
                                                            RAcc<T>::pick(acc64, acc32), tid, wave, lane))
    }
    // bring the FP64 tiles into the FP32 element order (row 4 kk + e), then
    // one writer per output element: row 16 wave + 4 (lane / 16) + e
    const int kk = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < NTM; ++nt) {
        if (nt >= Rp / 16) break;
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
