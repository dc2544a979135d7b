// mrhs.hip -- forward H-MVM for several right-hand sides on the matrix cores.
//
// Y += a A X with X, Y n x R column-major.  Same plan layout and phases as
// the single right-hand-side path (V pass -> partial sums -> R pass), with
// the inner products done as 16 x 16 x 4 MFMA tiles:
//   FP64 factors        v_mfma_f64_16x16x4f64
//   FP32/BF16/FP8       decoded to FP32, v_mfma_f32_16x16x4f32
// Right-hand sides are processed in tiles of 16 (R padded with zeros).
//
// Lane layout of the 16x16x4 tiles (checked on the device):
//   A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
//   f32: D[i][j] in lane j + 16 (i / 4), element i % 4
//   f64: D[i][j] in lane j + 16 (i % 4), element i / 4
//
// V pass:  out(c, r) = sum_i V(i, c) X(x0 + i, r)   M = columns, K = rows
// R pass:  Y(i, r)  += sum_c R(i, c) Z(c, r)        M = rows (one 16-row tile
//          per wave), K = columns staged in LDS blocks of 64 columns.
// Factor operands are read with one 16-byte load per lane whose W elements
// feed W consecutive MFMAs (the K order inside a tile is free as long as A
// and B agree).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"
#include "mrhs.hpp"
#include "mrhs_dev.hpp"

namespace hmvm {
namespace {

// XT = X^T (n x R row-major), 64 indices per workgroup
template <typename XT>
__global__ __launch_bounds__(kThr) void k_xtrans(const XT* __restrict__ X, int64_t ldx, int64_t n, int R,
                                                 double* __restrict__ XTb) {
    extern __shared__ double tile[];
    xtrans_tile(X, ldx, n, R, XTb, int64_t(blockIdx.x) * 64, tile, threadIdx.x);
}

// one workgroup per V segment (x slices of 64 x Rp in dynamic LDS)
template <int NTM, int FAM>
__global__ __launch_bounds__(kThr) void k_vfwd_mm(DevPlan P, const double* __restrict__ XTb, int R,
                                                  double* __restrict__ Pb) {
    extern __shared__ double xs[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const VItem sg = P.fitems[blockIdx.x];
    // the batched x-slice staging except for the hardware formats with four
    // tiles, whose kernel it takes from 3 to 2 waves per SIMD (64
    // right-hand sides, mixed plan 3.07 -> 3.23 ms)
    vseg_mm_dispatch<NTM, FAM, !(NTM == 4 && FAM == kFamHw)>(P, sg, XTb, R, Pb, xs, tid, wave, lane);
}

// one item (entry, NP right-hand sides) per thread
template <int NP>
__global__ __launch_bounds__(256) void k_psum_mm(DevPlan D, const double* __restrict__ Pb, double* __restrict__ t,
                                                 int R, double alpha) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    psum_mm_items<NP, 1, kPsumRound<NP>>(D, Pb, t, R, alpha, g, 0, D.ntent * (R / NP));
}

template <int NTM, typename YT, int FAM>
__global__ __launch_bounds__(kThr, (mm_min_waves<NTM, FAM>())) void k_ufwd_mm(DevPlan P, double alpha,
                                                                          const double* __restrict__ XTb,
                                                  const double* __restrict__ t, YT* __restrict__ Y, int64_t ldy,
                                                  int R) {
    extern __shared__ double zbuf[];               // 64 x mm_stride<NTM>(), sized at launch
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // runs of 32 consecutive row panels per XCD: the panels of one block row
    // gather the same t and XT rows, which then come from that XCD's L2
    const RowPanel Rw = P.rp[xcd_item(blockIdx.x, P.nrp, 32)];
    rpanel_mm<NTM, YT, FAM>(P, Rw, alpha, XTb, t, Y, ldy, R, zbuf, tid, wave, lane);
}

} // namespace

int launch_apply_mrhs(const DevPlan& D, int fam, double alpha, const void* x, int64_t ldx, void* y, int64_t ldy,
                      int R, int vdt, double* ws, hipStream_t st) {
    if (R < 2 || R > kMaxRp) return 1;
    if (fam != kFamHw && fam != kFamBfp && fam != kFamF64 && fam != kFamAfl) return 1;
    const int64_t pad = 32;
    // workspace: partials | t | XT (ncols x R)
    double* part = ws;
    double* t = part + ((D.psize * R + pad - 1) / pad) * pad;
    double* xt = t + ((D.ntent * R + pad - 1) / pad) * pad + pad;
    auto blocks = [](int64_t n, int per) { return unsigned((n + per - 1) / per); };
    const int Rp = (R + 15) & ~15;
    const size_t lds = mm_lds_bytes(Rp);
    if (D.ncols > 0) {
        if (vdt == 0)
            k_xtrans<double><<<blocks(D.ncols, 64), kThr, lds, st>>>(static_cast<const double*>(x), ldx, D.ncols, R, xt);
        else
            k_xtrans<float><<<blocks(D.ncols, 64), kThr, lds, st>>>(static_cast<const float*>(x), ldx, D.ncols, R, xt);
    }
    // accumulators sized for exactly the ceil(R / 16) tiles of this call
    // one instantiation per format family: a plan's family carries only its
    // own formats' decode paths (and registers)
    auto vpass = [&](auto ntm) {
        constexpr int NTM = decltype(ntm)::value;
        if (fam == kFamHw) k_vfwd_mm<NTM, kFamHw><<<unsigned(D.nfitems), kThr, lds, st>>>(D, xt, R, part);
        else if (fam == kFamF64) k_vfwd_mm<NTM, kFamF64><<<unsigned(D.nfitems), kThr, lds, st>>>(D, xt, R, part);
        else if (fam == kFamAfl) k_vfwd_mm<NTM, kFamAfl><<<unsigned(D.nfitems), kThr, lds, st>>>(D, xt, R, part);
        else k_vfwd_mm<NTM, kFamBfp><<<unsigned(D.nfitems), kThr, lds, st>>>(D, xt, R, part);
    };
    auto rpass = [&](auto ntm) {
        constexpr int NTM = decltype(ntm)::value;
        auto go = [&](auto ytag, auto famtag) {
            using YT = decltype(ytag);
            constexpr int FAM = decltype(famtag)::value;
            k_ufwd_mm<NTM, YT, FAM><<<unsigned(D.nrp), kThr, lds, st>>>(D, alpha, xt, t, static_cast<YT*>(y), ldy, R);
        };
        auto by_fam = [&](auto ytag) {
            if (fam == kFamHw) go(ytag, std::integral_constant<int, kFamHw>());
            else if (fam == kFamF64) go(ytag, std::integral_constant<int, kFamF64>());
            else if (fam == kFamAfl) go(ytag, std::integral_constant<int, kFamAfl>());
            else go(ytag, std::integral_constant<int, kFamBfp>());
        };
        if (vdt == 0) by_fam(double(0));
        else by_fam(float(0));
    };
    auto by_tiles = [&](auto pass) {
        switch (Rp / 16) {
        case 1: pass(std::integral_constant<int, 1>()); break;
        case 2: pass(std::integral_constant<int, 2>()); break;
        case 3: pass(std::integral_constant<int, 3>()); break;
        default: pass(std::integral_constant<int, 4>()); break;
        }
    };
    if (D.nfitems > 0) by_tiles(vpass);
    if (D.ntent > 0) {
        // pairs of right-hand sides when R is even and the workspace 16-byte aligned
        if (R % 2 == 0 && (reinterpret_cast<uintptr_t>(part) & 15) == 0)
            k_psum_mm<2><<<blocks(D.ntent * (R / 2), 256), 256, 0, st>>>(D, part, t, R, alpha);
        else
            k_psum_mm<1><<<blocks(D.ntent * R, 256), 256, 0, st>>>(D, part, t, R, alpha);
    }
    if (D.nrp > 0) by_tiles(rpass);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int64_t mrhs_ws_doubles(const DevPlan& D, int R) {
    const int64_t pad = 32;
    return ((D.psize * R + pad - 1) / pad) * pad + ((D.ntent * R + pad - 1) / pad) * pad + pad +
           ((D.ncols * R + pad - 1) / pad) * pad;
}

} // namespace hmvm
