// mrhs.hpp -- forward product for several right-hand sides (matrix cores).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace hmvm {

constexpr int kMrhsMax = 64;   // right-hand sides per launch sequence

// Y += alpha A X for 2 <= R <= kMrhsMax right-hand sides; vdt 0: FP64
// vectors, 1: FP32; fam: the plan's format family (kFamHw, kFamBfp, kFamAfl or kFamF64).
// ws holds mrhs_ws_doubles(D, R) doubles (partials, t and the
// transposed copy of X).  Returns 0 on success.
int launch_apply_mrhs(const DevPlan& D, int fam, double alpha, const void* x, int64_t ldx, void* y, int64_t ldy,
                      int R, int vdt, double* ws, hipStream_t st);

int64_t mrhs_ws_doubles(const DevPlan& D, int R);

} // namespace hmvm
