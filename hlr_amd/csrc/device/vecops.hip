// vecops.hip -- consumers of the product on the device: the spectral norm by
// power iteration, the 2-norm inversion error |I - A W|, and the linear
// iteration x <- x - W (A x - b).
//
// hmvm_spectral_norm restates hlr::norm::detail::spectral
// (arith/detail/norm.hh:448-527), the generic algorithm HLR runs through its
// arithmetic plugin: x random with |x| = 1 (std::default_random_engine,
// uniform in [-1, 1], as the reference); per step t = A x, y = A^H t
// (squared) or y = A x, lambda = sqrt(|<x, y>|) or |y|; stop on relative
// change < tol, on absolute change < min(10 eps, tol) after 5 steps, or when
// lambda or |y| drop below eps^2; x = y / |y|.
//
// hmvm_inv_error_2 runs the same iteration on the operator I - A W, as
// hlr::norm::inv_error_2 does on the sum operator it builds
// (arith/norm.hh:115-137), applying A and W one after the other instead of
// forming their product.
//
// hmvm_linear_iteration restates linear_iteration (programs/iterrefine.hh:
// 20-80): c = W r, x -= c, e = sol - x, r = A x - b, until |e| (or |r|/|b|
// without a known solution) drops to the tolerance or the step limit is hit.
//
// Products are hmvm_apply calls; the vector operations are the small kernels
// below.  Reductions use a fixed grid and a fixed summation order, so every
// result is reproducible run to run.

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
constexpr int kSlots = 4;   // reduction results kept on the device after the partials

// part[b] = sum over the block's grid-stride elements of a_i * b_i, or of
// (a_i - b_i)^2 when kDiff
template <bool kDiff>
__global__ __launch_bounds__(kRedThreads) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
                                                          int64_t n, double* __restrict__ part) {
    __shared__ double s[kRedThreads];
    double acc = 0;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += int64_t(kRedBlocks) * kRedThreads) {
        if constexpr (kDiff) {
            const double d = a[i] - b[i];
            acc += d * d;
        } else {
            acc += a[i] * b[i];
        }
    }
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

// y += a * x
__global__ __launch_bounds__(256) void k_axpy(double a, const double* __restrict__ x, double* __restrict__ y,
                                              int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
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
    int alloc(int64_t n) {
        return hipMalloc(&p, sizeof(double) * size_t(std::max<int64_t>(n, 1))) == hipSuccess ? HMVM_OK
                                                                                             : HMVM_ERR_ALLOC;
    }
};

unsigned blocks_for(int64_t n) { return unsigned((n + 255) / 256); }

// Vector operations on one stream, with a partials buffer and kSlots
// reduction results behind it on the device.
struct Vec {
    hipStream_t st;
    double* part;   // kRedBlocks partials, then kSlots results

    void fill(double* y, int64_t n, double v) const {
        if (n > 0) k_fill<<<blocks_for(n), 256, 0, st>>>(y, n, v);
    }
    void scale_copy(double a, const double* x, double* y, int64_t n) const {
        if (n > 0) k_scale_copy<<<blocks_for(n), 256, 0, st>>>(a, x, y, n);
    }
    void axpy(double a, const double* x, double* y, int64_t n) const {
        if (n > 0) k_axpy<<<blocks_for(n), 256, 0, st>>>(a, x, y, n);
    }
    // slot s = <a, b> (or |a - b|^2 when diff), left on the device
    void dot(const double* a, const double* b, int64_t n, int s, bool diff = false) const {
        if (diff)
            k_dot_part<true><<<kRedBlocks, kRedThreads, 0, st>>>(a, b, n, part);
        else
            k_dot_part<false><<<kRedBlocks, kRedThreads, 0, st>>>(a, b, n, part);
        k_dot_final<<<1, kRedThreads, 0, st>>>(part, part + kRedBlocks + s);
    }
    // copy slots [0, count) to the host and wait for them
    int read(double* out, int count) const {
        if (hipMemcpyAsync(out, part + kRedBlocks, sizeof(double) * size_t(count), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return HMVM_ERR_DEVICE;
        return HMVM_OK;
    }
};

// The power iteration of arith/detail/norm.hh:448-527 over an operator given
// by two callbacks: fwd(x, t) sets t = B x (nr entries), adj(t, y) sets
// y = B^H t (nc entries); the unsquared iteration (square B) uses fwd only.
template <class Fwd, class Adj>
int power_iteration(const Vec& V, int64_t nr, int64_t nc, double atol, int64_t amax_it, bool squared, Fwd&& fwd,
                    Adj&& adj, double* norm, int64_t* iterations) {
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
    DevBuf x, y, t;
    int rc;
    if ((rc = x.alloc(nc)) != HMVM_OK || (rc = y.alloc(nc)) != HMVM_OK || (rc = t.alloc(nr)) != HMVM_OK) return rc;
    if (hipMemcpyAsync(x.p, xh.data(), sizeof(double) * size_t(nc), hipMemcpyHostToDevice, V.st) != hipSuccess)
        return HMVM_ERR_DEVICE;
    const double eps = std::numeric_limits<double>::epsilon();
    const int64_t max_it = (amax_it == 0) ? std::max<int64_t>(5, std::min(nr, nc) / 10) : amax_it;
    const double tol = (atol == 0) ? std::sqrt(eps) : atol;
    const double abs_tol = std::min(10.0 * eps, tol);
    const double zero = eps * eps;
    double lambda = 1.0;
    int64_t it = 0;
    for (int64_t i = 0; i < max_it; ++i) {
        it = i + 1;
        double lambda_new = 0, norm_y = 0;
        double s[2] = {0, 0};
        if (squared) {
            if ((rc = fwd(x.p, t.p)) != HMVM_OK) return rc;
            if ((rc = adj(t.p, y.p)) != HMVM_OK) return rc;
            V.dot(x.p, y.p, nc, 0);
            V.dot(y.p, y.p, nc, 1);
            if ((rc = V.read(s, 2)) != HMVM_OK) return rc;
            lambda_new = std::sqrt(std::fabs(s[0]));
            norm_y = std::sqrt(s[1]);
        } else {
            // y = B x needs a square operator, as in the reference
            if (nr != nc) return HMVM_ERR_INVALID_ARG;
            if ((rc = fwd(x.p, y.p)) != HMVM_OK) return rc;
            V.dot(y.p, y.p, nc, 1);
            if ((rc = V.read(s, 2)) != HMVM_OK) return rc;
            norm_y = lambda_new = std::sqrt(s[1]);
        }
        if (std::fabs((lambda_new - lambda) / lambda) < tol) { lambda = lambda_new; break; }
        if (i > 5 && std::fabs(lambda_new - lambda) < abs_tol) { lambda = lambda_new; break; }
        if (lambda_new < zero) { lambda = lambda_new; break; }
        lambda = lambda_new;
        if (norm_y <= zero) break;
        V.scale_copy(1.0 / norm_y, y.p, x.p, nc);
    }
    if (hipStreamSynchronize(V.st) != hipSuccess) return HMVM_ERR_DEVICE;
    *norm = lambda;
    if (iterations) *iterations = it;
    return HMVM_OK;
}

// device and sizes of a plan
int plan_dims(const hmvm_plan* p, int* dev, int64_t* nr, int64_t* nc) {
    hmvm_plan_info info{};
    const int rc = hmvm_plan_get_info(p, &info);
    if (rc != HMVM_OK) return rc;
    *dev = hmvm::plan_device(p);
    if (*dev < 0) return HMVM_ERR_INVALID_ARG;
    *nr = info.nrows;
    *nc = info.ncols;
    return HMVM_OK;
}

} // namespace

extern "C" int hmvm_spectral_norm(const hmvm_plan* plan, double atol, int64_t amax_it, int squared, void* stream,
                                  double* norm, int64_t* iterations) {
    if (!plan || !norm || amax_it < 0) return HMVM_ERR_INVALID_ARG;
    int dev;
    int64_t nr, nc;
    int rc = plan_dims(plan, &dev, &nr, &nc);
    if (rc != HMVM_OK) return rc;
    Guard g(dev);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevBuf part;
    if ((rc = part.alloc(kRedBlocks + kSlots)) != HMVM_OK) return rc;
    const Vec V{st, part.p};
    // t = A x ; y = A^H t (both cleared first: hmvm_apply accumulates)
    auto fwd = [&](const double* x, double* t) {
        V.fill(t, nr, 0.0);
        return hmvm_apply(plan, 1.0, HMVM_OP_NORMAL, x, nc, t, nr, 1, HMVM_F64, st);
    };
    auto adj = [&](const double* t, double* y) {
        V.fill(y, nc, 0.0);
        return hmvm_apply(plan, 1.0, HMVM_OP_ADJOINT, t, nr, y, nc, 1, HMVM_F64, st);
    };
    return power_iteration(V, nr, nc, atol, amax_it, squared != 0, fwd, adj, norm, iterations);
}

extern "C" int hmvm_inv_error_2(const hmvm_plan* A, const hmvm_plan* W, double atol, int64_t amax_it, void* stream,
                                double* error, int64_t* iterations) {
    if (!A || !W || !error || amax_it < 0) return HMVM_ERR_INVALID_ARG;
    int dev, devw;
    int64_t n, nc, wr, wc;
    int rc;
    if ((rc = plan_dims(A, &dev, &n, &nc)) != HMVM_OK || (rc = plan_dims(W, &devw, &wr, &wc)) != HMVM_OK) return rc;
    if (n != nc || wr != n || wc != n) return HMVM_ERR_INVALID_ARG;
    if (devw != dev) return HMVM_ERR_INVALID_ARG;
    Guard g(dev);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevBuf part, tmp;
    if ((rc = part.alloc(kRedBlocks + kSlots)) != HMVM_OK || (rc = tmp.alloc(n)) != HMVM_OK) return rc;
    const Vec V{st, part.p};
    // t = (I - A W) x = x - A (W x)
    auto fwd = [&](const double* x, double* t) {
        V.fill(tmp.p, n, 0.0);
        if ((rc = hmvm_apply(W, 1.0, HMVM_OP_NORMAL, x, n, tmp.p, n, 1, HMVM_F64, st)) != HMVM_OK) return rc;
        V.scale_copy(1.0, x, t, n);
        return hmvm_apply(A, -1.0, HMVM_OP_NORMAL, tmp.p, n, t, n, 1, HMVM_F64, st);
    };
    // y = (I - A W)^H t = t - W^H (A^H t)
    auto adj = [&](const double* t, double* y) {
        V.fill(tmp.p, n, 0.0);
        if ((rc = hmvm_apply(A, 1.0, HMVM_OP_ADJOINT, t, n, tmp.p, n, 1, HMVM_F64, st)) != HMVM_OK) return rc;
        V.scale_copy(1.0, t, y, n);
        return hmvm_apply(W, -1.0, HMVM_OP_ADJOINT, tmp.p, n, y, n, 1, HMVM_F64, st);
    };
    return power_iteration(V, n, n, atol, amax_it, true, fwd, adj, error, iterations);
}

extern "C" int hmvm_linear_iteration(const hmvm_plan* A, const hmvm_plan* W, const double* b, double* x,
                                     const double* sol, double tol, int64_t max_steps, void* stream, int64_t* steps,
                                     double* history) {
    if (!A || !W || !b || !x || tol < 0 || max_steps < 0) return HMVM_ERR_INVALID_ARG;
    int dev, devw;
    int64_t n, nc, wr, wc;
    int rc;
    if ((rc = plan_dims(A, &dev, &n, &nc)) != HMVM_OK || (rc = plan_dims(W, &devw, &wr, &wc)) != HMVM_OK) return rc;
    if (n != nc || wr != n || wc != n) return HMVM_ERR_INVALID_ARG;
    if (devw != dev) return HMVM_ERR_INVALID_ARG;
    Guard g(dev);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double stop = (tol == 0) ? 1e-8 : tol;
    const int64_t nsteps = (max_steps == 0) ? 100 : max_steps;
    DevBuf part, r, c;
    if ((rc = part.alloc(kRedBlocks + kSlots)) != HMVM_OK || (rc = r.alloc(n)) != HMVM_OK ||
        (rc = c.alloc(n)) != HMVM_OK)
        return rc;
    const Vec V{st, part.p};
    // slots: 0 = |r|^2, 1 = |c|^2, 2 = |e|^2, 3 = |b|^2
    // r = A x - b
    auto residual = [&]() {
        V.scale_copy(-1.0, b, r.p, n);
        return hmvm_apply(A, 1.0, HMVM_OP_NORMAL, x, n, r.p, n, 1, HMVM_F64, st);
    };
    if ((rc = residual()) != HMVM_OK) return rc;
    V.dot(r.p, r.p, n, 0);
    V.dot(b, b, n, 3);
    double s[kSlots] = {0, 0, 0, 0};
    if ((rc = V.read(s, kSlots)) != HMVM_OK) return rc;
    const double norm_b = std::sqrt(s[3]);
    double norm_r = std::sqrt(s[0]);
    int64_t step = 0;
    // without a known solution the iteration stops on the relative residual
    bool done = !sol && norm_r <= stop * norm_b;
    while (!done && step < nsteps) {
        // c = W r ; x -= c
        V.fill(c.p, n, 0.0);
        if ((rc = hmvm_apply(W, 1.0, HMVM_OP_NORMAL, r.p, n, c.p, n, 1, HMVM_F64, st)) != HMVM_OK) return rc;
        V.axpy(-1.0, c.p, x, n);
        V.dot(c.p, c.p, n, 1);
        if (sol) V.dot(sol, x, n, 2, true);
        // r = A x - b
        if ((rc = residual()) != HMVM_OK) return rc;
        V.dot(r.p, r.p, n, 0);
        if ((rc = V.read(s, 3)) != HMVM_OK) return rc;
        const double norm_c = std::sqrt(s[1]);
        const double norm_e = sol ? std::sqrt(s[2]) : std::numeric_limits<double>::quiet_NaN();
        if (history) {
            history[3 * step + 0] = norm_r;
            history[3 * step + 1] = norm_c;
            history[3 * step + 2] = norm_e;
        }
        ++step;
        norm_r = std::sqrt(s[0]);
        done = sol ? (norm_e <= stop) : (norm_r <= stop * norm_b);
    }
    if (hipStreamSynchronize(st) != hipSuccess) return HMVM_ERR_DEVICE;
    if (steps) *steps = step;
    return HMVM_OK;
}
