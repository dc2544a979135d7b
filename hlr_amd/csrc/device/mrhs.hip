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
//          per wave), K = columns staged in LDS blocks of kKB columns.

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "mrhs.hpp"
#include "mrhs_dev.hpp"

namespace hmvm {
namespace {

// one workgroup per V segment (x slice 64 x Rp in dynamic LDS)
template <typename XT>
__global__ __launch_bounds__(kThr) void k_vfwd_mm(DevPlan P, const XT* __restrict__ X, int64_t ldx, int R,
                                                  double* __restrict__ Pb) {
    extern __shared__ double xs[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const VItem sg = P.vitems[blockIdx.x];
    vseg_mm_dispatch(P, sg, X, ldx, R, Pb, xs, tid, wave, lane);
}

__global__ __launch_bounds__(256) void k_psum1_mm(DevPlan D, const double* __restrict__ Pb, double* __restrict__ tmp,
                                                  int R) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g < D.nps * R) psum1_item(D, Pb, tmp, R, g);
}

__global__ __launch_bounds__(256) void k_psum2_mm(DevPlan D, const double* __restrict__ tmp, double* __restrict__ t,
                                                  int R, double alpha) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g < D.ntent * R) psum2_item(D, tmp, t, R, alpha, g);
}

template <typename XT, typename YT>
__global__ __launch_bounds__(kThr) void k_ufwd_mm(DevPlan P, double alpha, const XT* __restrict__ X, int64_t ldx,
                                                  const double* __restrict__ t, YT* __restrict__ Y, int64_t ldy,
                                                  int R) {
    extern __shared__ double zbuf[];               // kKB x Rp, sized at launch
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const RowPanel Rw = P.rp[blockIdx.x];
    rpanel_mm(P, Rw, alpha, X, ldx, t, Y, ldy, R, zbuf, tid, wave, lane);
}

} // namespace

int launch_apply_mrhs(const DevPlan& D, int64_t max_ws_part, double alpha, const void* x, int64_t ldx, void* y,
                      int64_t ldy, int R, int vdt, double* ws, hipStream_t st) {
    if (R < 2 || R > kMaxRp) return 1;
    (void)max_ws_part;
    const int64_t pad = 32;
    double* part = ws;
    double* tmp = part + ((D.psize * R + pad - 1) / pad) * pad;
    double* t = tmp + ((D.nps * R + pad - 1) / pad) * pad;
    auto blocks = [](int64_t n, int per) { return unsigned((n + per - 1) / per); };
    const int Rp = (R + 15) & ~15;
    const size_t lds_v = size_t(64) * Rp * sizeof(double);
    const size_t lds_r = size_t(kKB) * Rp * sizeof(double);
    if (D.nvitems > 0) {
        if (vdt == 0)
            k_vfwd_mm<double><<<unsigned(D.nvitems), kThr, lds_v, st>>>(D, static_cast<const double*>(x), ldx, R, part);
        else
            k_vfwd_mm<float><<<unsigned(D.nvitems), kThr, lds_v, st>>>(D, static_cast<const float*>(x), ldx, R, part);
    }
    if (D.ntent > 0) {
        if (D.nps > 0) k_psum1_mm<<<blocks(D.nps * R, 256), 256, 0, st>>>(D, part, tmp, R);
        k_psum2_mm<<<blocks(D.ntent * R, 256), 256, 0, st>>>(D, tmp, t, R, alpha);
    }
    if (D.nrp > 0) {
        if (vdt == 0)
            k_ufwd_mm<double, double><<<unsigned(D.nrp), kThr, lds_r, st>>>(D, alpha, static_cast<const double*>(x), ldx,
                                                                           t, static_cast<double*>(y), ldy, R);
        else
            k_ufwd_mm<float, float><<<unsigned(D.nrp), kThr, lds_r, st>>>(D, alpha, static_cast<const float*>(x), ldx, t,
                                                                         static_cast<float*>(y), ldy, R);
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int64_t mrhs_ws_doubles(const DevPlan& D, int R) {
    const int64_t pad = 32;
    return ((D.psize * R + pad - 1) / pad) * pad + ((D.nps * R + pad - 1) / pad) * pad +
           ((D.ntent * R + pad - 1) / pad) * pad + pad;
}

} // namespace hmvm
