// vecops.hip -- spectral norm of a plan by power iteration, on the device.
//
// Restates hlr::norm::detail::spectral (arith/detail/norm.hh:448-527), the
// generic algorithm HLR runs through its arithmetic plugin: x random with
// |x| = 1 (std::default_random_engine, uniform in [-1, 1], as the
// reference); per step t = A x, y = A^H t (squared) or y = A x, lambda =
// sqrt(|<x, y>|) or |y|; stop on relative change < tol, on absolute change
// < min(10 eps, tol) after 5 steps, or when lambda or |y| drop below eps^2;
// x = y / |y|.  The products are hmvm_apply calls; the vector operations are
// the small kernels below.  Reductions use a fixed grid and a fixed
// summation order, so the result is reproducible run to run.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <random>
#include <vector>

#include "hmvm.h"
#include "internal.hpp"

namespace {

constexpr int kRedBlocks = 256;
constexpr int kRedThreads = 256;

// part[b] = sum over the block's grid-stride elements of a_i * b_i
__global__ __launch_bounds__(kRedThreads) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
                                                          int64_t n, double* __restrict__ part) {
    __shared__ double s[kRedThreads];
    double acc = 0;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += int64_t(kRedBlocks) * kRedThreads)
        acc += a[i] * b[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// out = sum of the kRedBlocks partials, in a fixed order
__global__ __launch_bounds__(kRedThreads) void k_dot_final(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double s[kRedThreads];
    s[threadIdx.x] = (int(threadIdx.x) < kRedBlocks) ? part[threadIdx.x] : 0.0;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// y = a * x  (x and y may alias)
__global__ __launch_bounds__(256) void k_scale_copy(double a, const double* x, double* y, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ __launch_bounds__(256) void k_fill(double* y, int64_t n, double v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

struct Guard {
    int prev = -1;
    bool ok = true;
    explicit Guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) { ok = false; return; }
        if (prev != dev && hipSetDevice(dev) != hipSuccess) ok = false;
    }
    ~Guard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

struct DevBuf {
    double* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

} // namespace

extern "C" int hmvm_spectral_norm(const hmvm_plan* plan, double atol, int64_t amax_it, int squared, void* stream,
                                  double* norm, int64_t* iterations) {
    if (!plan || !norm) return HMVM_ERR_INVALID_ARG;
    hmvm_plan_info info{};
    int rc = hmvm_plan_get_info(plan, &info);
    if (rc != HMVM_OK) return rc;
    const int dev = hmvm::plan_device(plan);
    if (dev < 0) return HMVM_ERR_INVALID_ARG;
    Guard g(dev);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nr = info.nrows, nc = info.ncols;
    // x = rand with |x| = 1 (host: the reference's generator and distribution)
    std::vector<double> xh(size_t(std::max<int64_t>(nc, 1)));
    {
        auto generator = std::default_random_engine();
        auto uniform_distr = std::uniform_real_distribution<double>(-1.0, 1.0);
        for (int64_t i = 0; i < nc; ++i) xh[size_t(i)] = uniform_distr(generator);
        double s = 0;
        for (int64_t i = 0; i < nc; ++i) s += xh[size_t(i)] * xh[size_t(i)];
        s = std::sqrt(s);
        if (s > 0)
            for (int64_t i = 0; i < nc; ++i) xh[size_t(i)] /= s;
    }
    DevBuf x, y, t, part;
    if (hipMalloc(&x.p, sizeof(double) * size_t(std::max<int64_t>(nc, 1))) != hipSuccess ||
        hipMalloc(&y.p, sizeof(double) * size_t(std::max<int64_t>(nc, 1))) != hipSuccess ||
        hipMalloc(&t.p, sizeof(double) * size_t(std::max<int64_t>(nr, 1))) != hipSuccess ||
        hipMalloc(&part.p, sizeof(double) * (kRedBlocks + 2)) != hipSuccess)
        return HMVM_ERR_ALLOC;
    if (hipMemcpyAsync(x.p, xh.data(), sizeof(double) * size_t(nc), hipMemcpyHostToDevice, st) != hipSuccess)
        return HMVM_ERR_DEVICE;
    auto blocks = [](int64_t n) { return unsigned((n + 255) / 256); };
    // <a, b> for vectors of length n, synchronised and returned on the host
    auto dot = [&](const double* a, const double* b, int64_t n, double* out) -> int {
        k_dot_part<<<kRedBlocks, kRedThreads, 0, st>>>(a, b, n, part.p);
        k_dot_final<<<1, kRedThreads, 0, st>>>(part.p, part.p + kRedBlocks);
        if (hipMemcpyAsync(out, part.p + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return HMVM_ERR_DEVICE;
        return HMVM_OK;
    };
    const double eps = std::numeric_limits<double>::epsilon();
    const int64_t max_it = (amax_it == 0) ? std::max<int64_t>(5, std::min(nr, nc) / 10) : amax_it;
    const double tol = (atol == 0) ? std::sqrt(eps) : atol;
    const double abs_tol = std::min(10.0 * eps, tol);
    const double zero = eps * eps;
    double lambda = 1.0;
    int64_t it = 0;
    for (int64_t i = 0; i < max_it; ++i) {
        it = i + 1;
        double lambda_new = 0, norm_y = 0, xy = 0, yy = 0;
        if (squared) {
            if (nr > 0) k_fill<<<blocks(nr), 256, 0, st>>>(t.p, nr, 0.0);
            if ((rc = hmvm_apply(plan, 1.0, HMVM_OP_NORMAL, x.p, nc, t.p, nr, 1, HMVM_F64, st)) != HMVM_OK) return rc;
            if (nc > 0) k_fill<<<blocks(nc), 256, 0, st>>>(y.p, nc, 0.0);
            if ((rc = hmvm_apply(plan, 1.0, HMVM_OP_ADJOINT, t.p, nr, y.p, nc, 1, HMVM_F64, st)) != HMVM_OK) return rc;
            if ((rc = dot(x.p, y.p, nc, &xy)) != HMVM_OK) return rc;
            if ((rc = dot(y.p, y.p, nc, &yy)) != HMVM_OK) return rc;
            lambda_new = std::sqrt(std::fabs(xy));
            norm_y = std::sqrt(yy);
        } else {
            // y = A x needs a square operator, as in the reference
            if (nr != nc) return HMVM_ERR_INVALID_ARG;
            if (nc > 0) k_fill<<<blocks(nc), 256, 0, st>>>(y.p, nc, 0.0);
            if ((rc = hmvm_apply(plan, 1.0, HMVM_OP_NORMAL, x.p, nc, y.p, nr, 1, HMVM_F64, st)) != HMVM_OK) return rc;
            if ((rc = dot(y.p, y.p, nc, &yy)) != HMVM_OK) return rc;
            norm_y = lambda_new = std::sqrt(yy);
        }
        if (std::fabs((lambda_new - lambda) / lambda) < tol) { lambda = lambda_new; break; }
        if (i > 5 && std::fabs(lambda_new - lambda) < abs_tol) { lambda = lambda_new; break; }
        if (lambda_new < zero) { lambda = lambda_new; break; }
        lambda = lambda_new;
        if (norm_y <= zero) break;
        if (nc > 0) k_scale_copy<<<blocks(nc), 256, 0, st>>>(1.0 / norm_y, y.p, x.p, nc);
    }
    if (hipStreamSynchronize(st) != hipSuccess) return HMVM_ERR_DEVICE;
    *norm = lambda;
    if (iterations) *iterations = it;
    return HMVM_OK;
}
