// hmvm.hip -- kernels and C-ABI of the MI355X H-/TLR-matrix-vector product.
//
// Forward y += a A x (arith/mulvec.hh:27-64 semantics, leaves as in
// matrix/lrmatrix.hh:625-679 and matrix/dense_matrix.hh:459-481):
//   k_vfwd    P[e, q] = sum_{i in q} V(i, e) x_i     one wave per <= 64 column
//                                                    groups of a column panel,
//                                                    lane = column group
//   k_tsum    t_e = a S_e sum_q P[e, q]              one thread per t entry
//   k_ufwd    y_p += sum_c R_p(:, c) z_c             one workgroup per row
//             z_c = a x_j (dense) | t_e (low-rank)   panel, lane = row
// Transposed y += a A^T x (arith/blas.hh:910-948):
//   k_rtrn    Q[c] = sum_{i in p} R_p(i, c) x_i      wave per chunk, lane =
//                                                    row, butterfly reduction
//   k_tsum    s_e = a S_e sum_p Q[e, p]
//   k_vtrn    y_q += V_q s + a sum Q_dense           workgroup per column panel
// No atomics: every output element has exactly one writer and a fixed
// summation order, so results are bitwise reproducible run to run.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "hmvm.h"
#include "kernels.hpp"
#include "plan_host.hpp"

using namespace hmvm;

namespace {

constexpr int kWaves = 4;            // waves per workgroup
constexpr int kThreads = 64 * kWaves;

// ---------------------------------------------------------------------------
// forward: V pass
// ---------------------------------------------------------------------------

template <typename T, typename XT>
__global__ __launch_bounds__(kThreads) void k_vfwd(DevPlan P, const VItem* __restrict__ items, int64_t nitems,
                                                   const XT* __restrict__ x, double* __restrict__ Pb) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    __shared__ A xs[kWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t item = int64_t(blockIdx.x) * kWaves + wave;
    if (item >= nitems) return;
    const VItem it = items[item];
    const ColPanel C = P.cp[it.q];
    const int m = C.m;
    const int f = it.fmt;
    if (lane < m) xs[wave][lane] = A(x[C.col0 + lane]);
    wave_sync();
    const int64_t kp = C.kpad[f];
    int gp = 1, lg = 0;
    while (gp < it.ng) { gp <<= 1; ++lg; }
    const int g = lane & (gp - 1), ph = lane >> lg, nph = 64 >> lg;
    const bool active = g < it.ng;
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    if (active) {
        const T* base = static_cast<const T*>(P.arena[f]) + C.off[f] + int64_t(it.g0 + g) * W;
#pragma unroll 4
        for (int i = ph; i < m; i += nph) {
            A v[W];
            E::load(base + int64_t(i) * kp, v);
            const A xi = xs[wave][i];
#pragma unroll
            for (int w = 0; w < W; ++w) acc[w] += v[w] * xi;
        }
    }
    for (int s = gp; s < 64; s <<= 1) {
#pragma unroll
        for (int w = 0; w < W; ++w) acc[w] += __shfl_xor(acc[w], s, 64);
    }
    if (active && ph == 0) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int64_t vc = C.vcol0[f] + int64_t(it.g0 + g) * W + w;
            const int32_t d = P.vdest[vc];
            if (d >= 0) Pb[d] = double(acc[w]) * (P.vscale ? double(P.vscale[vc]) : 1.0);
        }
    }
}

// t_e = alpha * S_e * sum_r P[off_e + r]
__global__ __launch_bounds__(256) void k_tsum(const int32_t* __restrict__ off, const int32_t* __restrict__ cnt,
                                              const double* __restrict__ tmul, const double* __restrict__ Pb,
                                              double* __restrict__ t, int64_t ne, double alpha) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int32_t o = off[e], c = cnt[e];
    double s = 0;
    for (int32_t r = 0; r < c; ++r) s += Pb[o + r];
    t[e] = alpha * tmul[e] * s;
}

// ---------------------------------------------------------------------------
// forward: R pass (dense rows and U rows of a row panel)
// ---------------------------------------------------------------------------

template <typename T, typename XT>
__device__ inline double chunk_fwd(const DevPlan& P, const RChunk& C, int m, double alpha,
                                   const XT* __restrict__ x, const double* __restrict__ t, void* zbuf, int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    A* zs = static_cast<A*>(zbuf);
    const int ncol = C.ngroups * W;
    wave_sync();
    for (int j = lane; j < ncol; j += 64) {
        const int32_t s = P.colsrc[C.col0 + j];
        double z = (s >= 0) ? alpha * double(x[s]) : (s == kPadCol ? 0.0 : t[~s]);
        if (P.zscale) z *= double(P.zscale[C.col0 + j]);
        zs[j] = A(z);
    }
    wave_sync();
    A acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = A(0);
    if (lane < m) {
        const T* base = static_cast<const T*>(P.arena[C.fmt]) + C.data_off + int64_t(lane) * W;
        const int64_t gs = int64_t(m) * W;
#pragma unroll 8
        for (int g = 0; g < C.ngroups; ++g) {
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

template <typename XT, typename YT>
__global__ __launch_bounds__(kThreads) void k_ufwd(DevPlan P, double alpha, const XT* __restrict__ x,
                                                   const double* __restrict__ t, YT* __restrict__ y) {
    __shared__ __align__(16) float zbuf[kWaves][1024];
    __shared__ double red[kWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const RowPanel R = P.rp[blockIdx.x];
    double acc = 0;
    for (int ci = wave; ci < R.nchunks; ci += kWaves) {
        const RChunk C = P.rchunks[R.chunk0 + ci];
        switch (C.fmt) {
        case F64: acc += chunk_fwd<double>(P, C, R.m, alpha, x, t, zbuf[wave], lane); break;
        case F32: acc += chunk_fwd<float>(P, C, R.m, alpha, x, t, zbuf[wave], lane); break;
        case BF16: acc += chunk_fwd<uint16_t>(P, C, R.m, alpha, x, t, zbuf[wave], lane); break;
        default: acc += chunk_fwd<uint8_t>(P, C, R.m, alpha, x, t, zbuf[wave], lane); break;
        }
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && lane < R.m) {
        const double s = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
        y[R.row0 + lane] = YT(double(y[R.row0 + lane]) + s);
    }
}

// ---------------------------------------------------------------------------
// transposed: R pass with column reduction
// ---------------------------------------------------------------------------

template <typename T, typename XT>
__global__ __launch_bounds__(kThreads) void k_rtrn(DevPlan P, const int32_t* __restrict__ ids, int64_t n,
                                                   const XT* __restrict__ x, double* __restrict__ Q) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    constexpr int SG = 64 / W;   // groups per 64 columns
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t idx = int64_t(blockIdx.x) * kWaves + wave;
    if (idx >= n) return;
    const RChunk C = P.rchunks[ids[idx]];
    const RowPanel R = P.rp[C.panel];
    const int m = R.m;
    const A xi = (lane < m) ? A(x[R.row0 + lane]) : A(0);
    const T* base = static_cast<const T*>(P.arena[C.fmt]) + C.data_off + int64_t(lane) * W;
    const int64_t gs = int64_t(m) * W;
    for (int g0 = 0; g0 < C.ngroups; g0 += SG) {
        A v[64];
#pragma unroll
        for (int gg = 0; gg < SG; ++gg) {
            const int g = g0 + gg;
            if (g < C.ngroups && lane < m) {
                A u[W];
                E::load(base + g * gs, u);
#pragma unroll
                for (int w = 0; w < W; ++w) v[gg * W + w] = u[w] * xi;
            } else {
#pragma unroll
                for (int w = 0; w < W; ++w) v[gg * W + w] = A(0);
            }
        }
        const A s = butterfly64(v, lane);
        const int col = g0 * W + lane;
        if (col < C.ngroups * W) {
            const int64_t gc = C.col0 + col;
            const int32_t d = P.rdest[gc];
            if (d >= 0) Q[d] = double(s) * (P.zscale ? double(P.zscale[gc]) : 1.0);
        }
    }
}

// ---------------------------------------------------------------------------
// transposed: V pass, y_q += V_q s  (+ dense partials)
// ---------------------------------------------------------------------------

template <typename T>
__device__ inline void chunk_vtrn(const DevPlan& P, const ColPanel& C, int f, int g0, int ng,
                                  const double* __restrict__ s, double (&acc)[64], int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    const int g = g0 + lane;
    if (g >= ng) return;
    double sv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int64_t vc = C.vcol0[f] + int64_t(g) * W + w;
        const int32_t e = P.vsrc[vc];
        sv[w] = (e >= 0) ? s[e] * (P.vscale ? double(P.vscale[vc]) : 1.0) : 0.0;
    }
    const int64_t kp = C.kpad[f];
    const T* base = static_cast<const T*>(P.arena[f]) + C.off[f] + int64_t(g) * W;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        if (i < C.m) {
            A v[W];
            E::load(base + int64_t(i) * kp, v);
            double d = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) d += double(v[w]) * sv[w];
            acc[i] += d;
        }
    }
}

template <typename YT>
__global__ __launch_bounds__(kThreads) void k_vtrn(DevPlan P, double alpha, const double* __restrict__ s,
                                                   const double* __restrict__ Q, YT* __restrict__ y) {
    __shared__ double red[kWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const ColPanel C = P.cp[blockIdx.x];
    double acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0;
    int ck = 0;
    for (int f = 0; f < NFMT; ++f) {
        const int W = fmt_width(f);
        const int ng = C.kpad[f] / W;
        for (int g0 = 0; g0 < ng; g0 += 64, ++ck) {
            if ((ck % kWaves) != wave) continue;
            switch (f) {
            case F64: chunk_vtrn<double>(P, C, f, g0, ng, s, acc, lane); break;
            case F32: chunk_vtrn<float>(P, C, f, g0, ng, s, acc, lane); break;
            case BF16: chunk_vtrn<uint16_t>(P, C, f, g0, ng, s, acc, lane); break;
            default: chunk_vtrn<uint8_t>(P, C, f, g0, ng, s, acc, lane); break;
            }
        }
    }
    const double r = butterfly64(acc, lane);
    red[wave][lane] = r;
    __syncthreads();
    if (wave == 0 && lane < C.m) {
        const int64_t j = C.col0 + lane;
        double d = 0;
        const int32_t o = P.qd_off[j], c = P.qd_cnt[j];
        for (int32_t k = 0; k < c; ++k) d += Q[o + k];
        const double v = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane] + alpha * d;
        y[j] = YT(double(y[j]) + v);
    }
}

// ---------------------------------------------------------------------------
// adjoint: factor gradients in arena layout (FP64 values)
// ---------------------------------------------------------------------------

// dR(i, c) = ybar_i * z_c with z = x_j (dense) | t_e (low-rank)
template <typename T, typename XT>
__global__ __launch_bounds__(kThreads) void k_grad_r(DevPlan P, const int32_t* __restrict__ ids, int64_t n,
                                                     const XT* __restrict__ x, const XT* __restrict__ ybar,
                                                     const double* __restrict__ t, double* __restrict__ gR) {
    constexpr int W = Elem<T>::W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t idx = int64_t(blockIdx.x) * kWaves + wave;
    if (idx >= n) return;
    const RChunk C = P.rchunks[ids[idx]];
    const RowPanel R = P.rp[C.panel];
    const int m = R.m;
    if (lane >= m) return;
    const double yb = double(ybar[R.row0 + lane]);
    for (int g = 0; g < C.ngroups; ++g) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int64_t gc = C.col0 + int64_t(g) * W + w;
            const int32_t s = P.colsrc[gc];
            const double z = (s >= 0) ? double(x[s]) : (s == kPadCol ? 0.0 : t[~s]);
            gR[C.data_off + (int64_t(g) * m + lane) * W + w] = yb * z;
        }
    }
}

// dV(i, c) = x_i * s_e
template <typename T, typename XT>
__global__ __launch_bounds__(kThreads) void k_grad_v(DevPlan P, const VItem* __restrict__ items, int64_t nitems,
                                                     const XT* __restrict__ x, const double* __restrict__ s,
                                                     double* __restrict__ gV) {
    constexpr int W = Elem<T>::W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t item = int64_t(blockIdx.x) * kWaves + wave;
    if (item >= nitems) return;
    const VItem it = items[item];
    const ColPanel C = P.cp[it.q];
    const int f = it.fmt;
    const int64_t kp = C.kpad[f];
    if (lane >= it.ng) return;
    const int g = it.g0 + lane;
    double sv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int32_t e = P.vsrc[C.vcol0[f] + int64_t(g) * W + w];
        sv[w] = (e >= 0) ? s[e] : 0.0;
    }
    for (int i = 0; i < C.m; ++i) {
        const double xi = double(x[C.col0 + i]);
#pragma unroll
        for (int w = 0; w < W; ++w) gV[C.off[f] + int64_t(i) * kp + int64_t(g) * W + w] = xi * sv[w];
    }
}

} // namespace

// ===========================================================================
// plan object
// ===========================================================================

struct hmvm_plan {
    HostLayout L;
    int device = -1;
    DevPlan dp{};
    std::vector<void*> allocs;
    void* arena_dev[NFMT] = {nullptr, nullptr, nullptr, nullptr};
    // work lists on the device, grouped by format
    VItem* vitems_dev = nullptr;
    int64_t vitem_start[NFMT + 1] = {0, 0, 0, 0, 0};
    int32_t* rchunk_ids_dev = nullptr;
    int64_t rchunk_start[NFMT + 1] = {0, 0, 0, 0, 0};
};

struct hmvm_batch {
    int unused;
};

namespace {

struct DeviceGuard {
    int prev = -1;
    bool ok = true;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) { ok = false; return; }
        if (prev != dev && hipSetDevice(dev) != hipSuccess) ok = false;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename V>
int upload(hmvm_plan* p, const std::vector<V>& h, const V** out) {
    *out = nullptr;
    if (h.empty()) return HMVM_OK;
    void* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(V)) != hipSuccess) return HMVM_ERR_ALLOC;
    p->allocs.push_back(d);
    if (hipMemcpy(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice) != hipSuccess) return HMVM_ERR_DEVICE;
    *out = static_cast<const V*>(d);
    return HMVM_OK;
}

int upload_plan(hmvm_plan* p) {
    HostLayout& L = p->L;
    DevPlan& D = p->dp;
    int rc;
#define UP(field, vec) if ((rc = upload(p, vec, &D.field)) != HMVM_OK) return rc
    UP(rp, L.rp);
    UP(rchunks, L.rchunks);
    UP(cp, L.cp);
    UP(colsrc, L.colsrc);
    UP(zscale, L.zscale);
    UP(rdest, L.rdest);
    UP(vdest, L.vdest);
    UP(vsrc, L.vsrc);
    UP(vscale, L.vscale);
    UP(tmul, L.tmul);
    UP(pt_off, L.pt_off);
    UP(pt_cnt, L.pt_cnt);
    UP(qt_off, L.qt_off);
    UP(qt_cnt, L.qt_cnt);
    UP(qd_off, L.qd_off);
    UP(qd_cnt, L.qd_cnt);
#undef UP
    // work lists grouped by format
    std::vector<VItem> vi;
    for (int f = 0; f < NFMT; ++f) {
        p->vitem_start[f] = int64_t(vi.size());
        for (const VItem& it : L.vitems) if (it.fmt == f) vi.push_back(it);
    }
    p->vitem_start[NFMT] = int64_t(vi.size());
    const VItem* vd = nullptr;
    if ((rc = upload(p, vi, &vd)) != HMVM_OK) return rc;
    p->vitems_dev = const_cast<VItem*>(vd);
    D.vitems = vd;
    std::vector<int32_t> ids;
    for (int f = 0; f < NFMT; ++f) {
        p->rchunk_start[f] = int64_t(ids.size());
        for (size_t c = 0; c < L.rchunks.size(); ++c) if (L.rchunks[c].fmt == f) ids.push_back(int32_t(c));
    }
    p->rchunk_start[NFMT] = int64_t(ids.size());
    const int32_t* idd = nullptr;
    if ((rc = upload(p, ids, &idd)) != HMVM_OK) return rc;
    p->rchunk_ids_dev = const_cast<int32_t*>(idd);
    // arenas
    const void* a[NFMT] = {nullptr, nullptr, nullptr, nullptr};
    if ((rc = upload(p, L.a64, reinterpret_cast<const double**>(&a[0]))) != HMVM_OK) return rc;
    if ((rc = upload(p, L.a32, reinterpret_cast<const float**>(&a[1]))) != HMVM_OK) return rc;
    if ((rc = upload(p, L.a16, reinterpret_cast<const uint16_t**>(&a[2]))) != HMVM_OK) return rc;
    if ((rc = upload(p, L.a8, reinterpret_cast<const uint8_t**>(&a[3]))) != HMVM_OK) return rc;
    for (int f = 0; f < NFMT; ++f) { D.arena[f] = a[f]; p->arena_dev[f] = const_cast<void*>(a[f]); }
    D.nrp = int64_t(L.rp.size());
    D.ncp = int64_t(L.cp.size());
    D.nvitems = int64_t(vi.size());
    D.nrchunks = int64_t(L.rchunks.size());
    D.ntent = L.ntent;
    D.psize = L.psize;
    D.qsize = L.qsize;
    D.nrows = L.nrows;
    D.ncols = L.ncols;
    // the host copies of the arenas are no longer needed on the device path
    std::vector<double>().swap(L.a64);
    std::vector<float>().swap(L.a32);
    std::vector<uint16_t>().swap(L.a16);
    std::vector<uint8_t>().swap(L.a8);
    return HMVM_OK;
}

inline unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

int64_t ws_doubles(const hmvm_plan* p, int op) {
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const HostLayout& L = p->L;
    const int64_t a = trans ? L.qsize : L.psize;
    return ((a + 31) / 32) * 32 + ((L.ntent + 31) / 32) * 32 + 32;
}

template <typename XT, typename YT>
int launch_apply(const hmvm_plan* p, double alpha, int op, const XT* x, YT* y, double* ws, hipStream_t st,
                 hipEvent_t* ev = nullptr) {
    const DevPlan& D = p->dp;
    const HostLayout& L = p->L;
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t abuf = trans ? L.qsize : L.psize;
    double* part = ws;
    double* t = ws + ((abuf + 31) / 32) * 32;
    if (ev) (void)hipEventRecord(ev[0], st);
    if (!trans) {
        for (int f = 0; f < NFMT; ++f) {
            const int64_t n = p->vitem_start[f + 1] - p->vitem_start[f];
            if (n == 0) continue;
            const VItem* it = p->vitems_dev + p->vitem_start[f];
            const unsigned nb = blocks_for(n, kWaves);
            switch (f) {
            case F64: k_vfwd<double, XT><<<nb, kThreads, 0, st>>>(D, it, n, x, part); break;
            case F32: k_vfwd<float, XT><<<nb, kThreads, 0, st>>>(D, it, n, x, part); break;
            case BF16: k_vfwd<uint16_t, XT><<<nb, kThreads, 0, st>>>(D, it, n, x, part); break;
            default: k_vfwd<uint8_t, XT><<<nb, kThreads, 0, st>>>(D, it, n, x, part); break;
            }
        }
        if (ev) (void)hipEventRecord(ev[1], st);
        if (L.ntent > 0)
            k_tsum<<<blocks_for(L.ntent, 256), 256, 0, st>>>(D.pt_off, D.pt_cnt, D.tmul, part, t, L.ntent, alpha);
        if (ev) (void)hipEventRecord(ev[2], st);
        if (D.nrp > 0) k_ufwd<XT, YT><<<unsigned(D.nrp), kThreads, 0, st>>>(D, alpha, x, t, y);
        if (ev) (void)hipEventRecord(ev[3], st);
    } else {
        for (int f = 0; f < NFMT; ++f) {
            const int64_t n = p->rchunk_start[f + 1] - p->rchunk_start[f];
            if (n == 0) continue;
            const int32_t* ids = p->rchunk_ids_dev + p->rchunk_start[f];
            const unsigned nb = blocks_for(n, kWaves);
            switch (f) {
            case F64: k_rtrn<double, XT><<<nb, kThreads, 0, st>>>(D, ids, n, x, part); break;
            case F32: k_rtrn<float, XT><<<nb, kThreads, 0, st>>>(D, ids, n, x, part); break;
            case BF16: k_rtrn<uint16_t, XT><<<nb, kThreads, 0, st>>>(D, ids, n, x, part); break;
            default: k_rtrn<uint8_t, XT><<<nb, kThreads, 0, st>>>(D, ids, n, x, part); break;
            }
        }
        if (ev) (void)hipEventRecord(ev[1], st);
        if (L.ntent > 0)
            k_tsum<<<blocks_for(L.ntent, 256), 256, 0, st>>>(D.qt_off, D.qt_cnt, D.tmul, part, t, L.ntent, alpha);
        if (ev) (void)hipEventRecord(ev[2], st);
        if (D.ncp > 0) k_vtrn<YT><<<unsigned(D.ncp), kThreads, 0, st>>>(D, alpha, t, part, y);
        if (ev) (void)hipEventRecord(ev[3], st);
    }
    return hipGetLastError() == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

int apply_impl(const hmvm_plan* p, double alpha, int op, const void* x, int64_t ldx, void* y, int64_t ldy,
               int nrhs, int vdt, void* ws, int64_t ws_bytes, hipStream_t st) {
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t nin = trans ? p->L.nrows : p->L.ncols;
    const int64_t nout = trans ? p->L.ncols : p->L.nrows;
    if (nrhs > 1 && (ldx < nin || ldy < nout)) return HMVM_ERR_LAYOUT;
    if (ws_bytes < ws_doubles(p, op) * int64_t(sizeof(double))) return HMVM_ERR_INVALID_ARG;
    for (int r = 0; r < nrhs; ++r) {
        int rc;
        if (vdt == HMVM_F64)
            rc = launch_apply(p, alpha, op, static_cast<const double*>(x) + r * ldx, static_cast<double*>(y) + r * ldy,
                              static_cast<double*>(ws), st);
        else
            rc = launch_apply(p, alpha, op, static_cast<const float*>(x) + r * ldx, static_cast<float*>(y) + r * ldy,
                              static_cast<double*>(ws), st);
        if (rc != HMVM_OK) return rc;
    }
    return HMVM_OK;
}

int check_apply_args(const hmvm_plan* p, int op, const void* x, const void* y, int nrhs, int vdt) {
    if (!p) return HMVM_ERR_INVALID_ARG;
    if (op < HMVM_OP_NORMAL || op > HMVM_OP_CONJUGATE) return HMVM_ERR_INVALID_ARG;
    if (nrhs < 1 || (vdt != HMVM_F64 && vdt != HMVM_F32)) return HMVM_ERR_INVALID_ARG;
    if (p->device < 0) return HMVM_ERR_INVALID_ARG;                   // host-only plan
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t nin = trans ? p->L.nrows : p->L.ncols;
    const int64_t nout = trans ? p->L.ncols : p->L.nrows;
    if ((nin > 0 && !x) || (nout > 0 && !y)) return HMVM_ERR_INVALID_ARG;
    const size_t al = (vdt == HMVM_F64) ? 8 : 4;
    if ((reinterpret_cast<uintptr_t>(x) % al) || (reinterpret_cast<uintptr_t>(y) % al)) return HMVM_ERR_LAYOUT;
    return HMVM_OK;
}

} // namespace

// ===========================================================================
// C-ABI
// ===========================================================================

extern "C" {

const char* hmvm_error_string(int code) {
    switch (code) {
    case HMVM_OK: return "success";
    case HMVM_ERR_INVALID_ARG: return "invalid argument";
    case HMVM_ERR_BAD_BLOCK: return "invalid block descriptor";
    case HMVM_ERR_OVERLAP: return "leaf blocks overlap";
    case HMVM_ERR_NONFINITE: return "non-finite value in a factor";
    case HMVM_ERR_ALLOC: return "allocation failed";
    case HMVM_ERR_DEVICE: return "HIP runtime error";
    case HMVM_ERR_UNSUPPORTED: return "unsupported combination";
    case HMVM_ERR_LAYOUT: return "bad leading dimension or alignment";
    case HMVM_ERR_KERNEL: return "kernel reported a failed check";
    default: return "unknown error";
    }
}

int hmvm_plan_create(const hmvm_block* blocks, int64_t nblocks, int64_t nrows, int64_t ncols, const double* arena,
                     int64_t arena_len, const hmvm_plan_opts* opts, int device, hmvm_plan** out) {
    if (!out) return HMVM_ERR_INVALID_ARG;
    *out = nullptr;
    hmvm_plan_opts o;
    if (opts) o = *opts;
    else { std::memset(&o, 0, sizeof(o)); o.precision = HMVM_PREC_FP64; o.check_finite = 1; }
    std::unique_ptr<hmvm_plan> p;
    try {
        p.reset(new hmvm_plan);
        const int rc = build_layout(blocks, nblocks, nrows, ncols, arena, arena_len, o, p->L);
        if (rc != HMVM_OK) return rc;
    } catch (const std::bad_alloc&) {
        return HMVM_ERR_ALLOC;
    } catch (...) {
        return HMVM_ERR_INVALID_ARG;
    }
    p->device = device;
    if (device >= 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) return HMVM_ERR_DEVICE;
        DeviceGuard g(device);
        if (!g.ok) return HMVM_ERR_DEVICE;
        const int rc = upload_plan(p.get());
        if (rc != HMVM_OK) {
            for (void* a : p->allocs) (void)hipFree(a);
            return rc;
        }
    }
    *out = p.release();
    return HMVM_OK;
}

int hmvm_plan_destroy(hmvm_plan* p) {
    if (!p) return HMVM_OK;
    if (p->device >= 0) {
        DeviceGuard g(p->device);
        (void)hipDeviceSynchronize();
        for (void* a : p->allocs) (void)hipFree(a);
    }
    delete p;
    return HMVM_OK;
}

int hmvm_plan_get_info(const hmvm_plan* p, hmvm_plan_info* info) {
    if (!p || !info) return HMVM_ERR_INVALID_ARG;
    const HostLayout& L = p->L;
    std::memset(info, 0, sizeof(*info));
    info->nrows = L.nrows;
    info->ncols = L.ncols;
    info->nblocks = int64_t(L.blocks.size());
    info->ndense = L.ndense;
    info->nlowrank = L.nlowrank;
    info->nrow_panels = int64_t(L.rp.size());
    info->ncol_panels = int64_t(L.cp.size());
    int64_t fb = 0;
    for (int f = 0; f < NFMT; ++f) {
        info->bytes_by_format[f] = L.bytes_by_format[f];
        info->cols_by_format[f] = L.cols_by_format[f];
        fb += L.bytes_by_format[f];
    }
    info->factor_bytes = fb;
    info->scale_bytes = int64_t(L.zscale.size() + L.vscale.size()) * 4;
    info->flops = L.flops;
    info->datasize = L.datasize + 2.0 * 8.0 * double(L.nrows + L.ncols);
    info->one_level = L.one_level ? 1 : 0;
    info->index_bytes = int64_t(L.colsrc.size() + L.vdest.size()) * 4;
    info->bytes_rpass = L.bytes_rpass;
    info->bytes_vpass = L.bytes_vpass;
    return HMVM_OK;
}

int hmvm_workspace_size(const hmvm_plan* p, int op, int nrhs, int64_t* bytes) {
    if (!p || !bytes || nrhs < 1) return HMVM_ERR_INVALID_ARG;
    (void)nrhs;
    *bytes = ws_doubles(p, op) * int64_t(sizeof(double));
    return HMVM_OK;
}

int hmvm_apply_ws(const hmvm_plan* p, double alpha, int op, const void* x, int64_t ldx, void* y, int64_t ldy,
                  int nrhs, int vdt, void* ws, int64_t ws_bytes, void* stream) {
    int rc = check_apply_args(p, op, x, y, nrhs, vdt);
    if (rc != HMVM_OK) return rc;
    if (alpha == 0.0) return HMVM_OK;
    if (!ws) return HMVM_ERR_INVALID_ARG;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    return apply_impl(p, alpha, op, x, ldx, y, ldy, nrhs, vdt, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

int hmvm_apply(const hmvm_plan* p, double alpha, int op, const void* x, int64_t ldx, void* y, int64_t ldy, int nrhs,
               int vdt, void* stream) {
    int rc = check_apply_args(p, op, x, y, nrhs, vdt);
    if (rc != HMVM_OK) return rc;
    if (alpha == 0.0) return HMVM_OK;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t bytes = ws_doubles(p, op) * int64_t(sizeof(double));
    void* ws = nullptr;
    if (hipMallocAsync(&ws, size_t(bytes), st) != hipSuccess) return HMVM_ERR_ALLOC;
    rc = apply_impl(p, alpha, op, x, ldx, y, ldy, nrhs, vdt, ws, bytes, st);
    if (hipFreeAsync(ws, st) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    return rc;
}

int hmvm_apply_timed(const hmvm_plan* p, double alpha, int op, const void* x, void* y, int vdt, void* stream,
                     float ms[3]) {
    int rc = check_apply_args(p, op, x, y, 1, vdt);
    if (rc != HMVM_OK) return rc;
    if (!ms) return HMVM_ERR_INVALID_ARG;
    ms[0] = ms[1] = ms[2] = 0.0f;
    if (alpha == 0.0) return HMVM_OK;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t bytes = ws_doubles(p, op) * int64_t(sizeof(double));
    void* ws = nullptr;
    if (hipMallocAsync(&ws, size_t(bytes), st) != hipSuccess) return HMVM_ERR_ALLOC;
    hipEvent_t ev[4];
    for (int i = 0; i < 4; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) return HMVM_ERR_DEVICE;
    if (vdt == HMVM_F64)
        rc = launch_apply(p, alpha, op, static_cast<const double*>(x), static_cast<double*>(y),
                          static_cast<double*>(ws), st, ev);
    else
        rc = launch_apply(p, alpha, op, static_cast<const float*>(x), static_cast<float*>(y),
                          static_cast<double*>(ws), st, ev);
    if (hipFreeAsync(ws, st) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    if (hipEventSynchronize(ev[3]) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    for (int i = 0; i < 3 && rc == HMVM_OK; ++i)
        if (hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]) != hipSuccess) rc = HMVM_ERR_DEVICE;
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
    return rc;
}

int hmvm_tlr_apply(const hmvm_plan* p, double alpha, int op, const void* x, void* y, int vdt, void* stream) {
    if (!p) return HMVM_ERR_INVALID_ARG;
    if (!p->L.one_level) return HMVM_ERR_INVALID_ARG;
    return hmvm_apply(p, alpha, op, x, 0, y, 0, 1, vdt, stream);
}

int hmvm_grad(const hmvm_plan* p, const void* x, const void* ybar, void* gx, void* grad_arena[4], int vdt,
              void* stream) {
    int rc = check_apply_args(p, HMVM_OP_NORMAL, x, ybar, 1, vdt);
    if (rc != HMVM_OK) return rc;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const HostLayout& L = p->L;
    const DevPlan& D = p->dp;
    const int64_t np = ((L.psize + 31) / 32) * 32, nq = ((L.qsize + 31) / 32) * 32, ne = ((L.ntent + 31) / 32) * 32 + 32;
    const int64_t bytes = (np + nq + 2 * ne) * int64_t(sizeof(double));
    void* ws = nullptr;
    if (hipMallocAsync(&ws, size_t(bytes), st) != hipSuccess) return HMVM_ERR_ALLOC;
    double* Pb = static_cast<double*>(ws);
    double* Qb = Pb + np;
    double* t = Qb + nq;
    double* s = t + ne;
    auto run = [&](auto xt_tag) -> int {
        using XT = decltype(xt_tag);
        const XT* xx = static_cast<const XT*>(x);
        const XT* yb = static_cast<const XT*>(ybar);
        // t = S V^T x, s = S U^T ybar (alpha = 1)
        for (int f = 0; f < NFMT; ++f) {
            const int64_t n = p->vitem_start[f + 1] - p->vitem_start[f];
            if (n == 0) continue;
            const VItem* it = p->vitems_dev + p->vitem_start[f];
            const unsigned nb = blocks_for(n, kWaves);
            switch (f) {
            case F64: k_vfwd<double, XT><<<nb, kThreads, 0, st>>>(D, it, n, xx, Pb); break;
            case F32: k_vfwd<float, XT><<<nb, kThreads, 0, st>>>(D, it, n, xx, Pb); break;
            case BF16: k_vfwd<uint16_t, XT><<<nb, kThreads, 0, st>>>(D, it, n, xx, Pb); break;
            default: k_vfwd<uint8_t, XT><<<nb, kThreads, 0, st>>>(D, it, n, xx, Pb); break;
            }
        }
        for (int f = 0; f < NFMT; ++f) {
            const int64_t n = p->rchunk_start[f + 1] - p->rchunk_start[f];
            if (n == 0) continue;
            const int32_t* ids = p->rchunk_ids_dev + p->rchunk_start[f];
            const unsigned nb = blocks_for(n, kWaves);
            switch (f) {
            case F64: k_rtrn<double, XT><<<nb, kThreads, 0, st>>>(D, ids, n, yb, Qb); break;
            case F32: k_rtrn<float, XT><<<nb, kThreads, 0, st>>>(D, ids, n, yb, Qb); break;
            case BF16: k_rtrn<uint16_t, XT><<<nb, kThreads, 0, st>>>(D, ids, n, yb, Qb); break;
            default: k_rtrn<uint8_t, XT><<<nb, kThreads, 0, st>>>(D, ids, n, yb, Qb); break;
            }
        }
        if (L.ntent > 0) {
            k_tsum<<<blocks_for(L.ntent, 256), 256, 0, st>>>(D.pt_off, D.pt_cnt, D.tmul, Pb, t, L.ntent, 1.0);
            k_tsum<<<blocks_for(L.ntent, 256), 256, 0, st>>>(D.qt_off, D.qt_cnt, D.tmul, Qb, s, L.ntent, 1.0);
        }
        if (gx && D.ncp > 0) k_vtrn<XT><<<unsigned(D.ncp), kThreads, 0, st>>>(D, 1.0, s, Qb, static_cast<XT*>(gx));
        if (grad_arena) {
            for (int f = 0; f < NFMT; ++f) {
                if (!grad_arena[f]) continue;
                double* gA = static_cast<double*>(grad_arena[f]);
                const int64_t nr = p->rchunk_start[f + 1] - p->rchunk_start[f];
                if (nr > 0) {
                    const int32_t* ids = p->rchunk_ids_dev + p->rchunk_start[f];
                    const unsigned nb = blocks_for(nr, kWaves);
                    switch (f) {
                    case F64: k_grad_r<double, XT><<<nb, kThreads, 0, st>>>(D, ids, nr, xx, yb, t, gA); break;
                    case F32: k_grad_r<float, XT><<<nb, kThreads, 0, st>>>(D, ids, nr, xx, yb, t, gA); break;
                    case BF16: k_grad_r<uint16_t, XT><<<nb, kThreads, 0, st>>>(D, ids, nr, xx, yb, t, gA); break;
                    default: k_grad_r<uint8_t, XT><<<nb, kThreads, 0, st>>>(D, ids, nr, xx, yb, t, gA); break;
                    }
                }
                const int64_t nv = p->vitem_start[f + 1] - p->vitem_start[f];
                if (nv > 0) {
                    const VItem* it = p->vitems_dev + p->vitem_start[f];
                    const unsigned nb = blocks_for(nv, kWaves);
                    switch (f) {
                    case F64: k_grad_v<double, XT><<<nb, kThreads, 0, st>>>(D, it, nv, xx, s, gA); break;
                    case F32: k_grad_v<float, XT><<<nb, kThreads, 0, st>>>(D, it, nv, xx, s, gA); break;
                    case BF16: k_grad_v<uint16_t, XT><<<nb, kThreads, 0, st>>>(D, it, nv, xx, s, gA); break;
                    default: k_grad_v<uint8_t, XT><<<nb, kThreads, 0, st>>>(D, it, nv, xx, s, gA); break;
                    }
                }
            }
        }
        return hipGetLastError() == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
    };
    rc = (vdt == HMVM_F64) ? run(double(0)) : run(float(0));
    if (hipFreeAsync(ws, st) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    return rc;
}

int hmvm_plan_arena_elems(const hmvm_plan* p, int fmt, int64_t* elems) {
    if (!p || !elems || fmt < 0 || fmt >= NFMT) return HMVM_ERR_INVALID_ARG;
    *elems = p->device >= 0 ? p->L.bytes_by_format[fmt] / fmt_bytes(fmt) : p->L.arena_elems(fmt);
    return HMVM_OK;
}

int hmvm_plan_get_arena(const hmvm_plan* p, int fmt, void* dst, void* stream) {
    if (!p || !dst || fmt < 0 || fmt >= NFMT) return HMVM_ERR_INVALID_ARG;
    const int64_t bytes = p->L.bytes_by_format[fmt];
    if (bytes == 0) return HMVM_OK;
    if (p->device < 0) {
        const HostLayout& L = p->L;
        const void* src = fmt == F64 ? (const void*)L.a64.data() : fmt == F32 ? (const void*)L.a32.data()
                        : fmt == BF16 ? (const void*)L.a16.data() : (const void*)L.a8.data();
        std::memcpy(dst, src, size_t(bytes));
        return HMVM_OK;
    }
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    return hipMemcpyAsync(dst, p->arena_dev[fmt], size_t(bytes), hipMemcpyDefault, static_cast<hipStream_t>(stream)) ==
                   hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

int hmvm_plan_set_arena(hmvm_plan* p, int fmt, const void* src, void* stream) {
    if (!p || !src || fmt < 0 || fmt >= NFMT) return HMVM_ERR_INVALID_ARG;
    if (p->device < 0) return HMVM_ERR_INVALID_ARG;
    const int64_t bytes = p->L.bytes_by_format[fmt];
    if (bytes == 0) return HMVM_OK;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    return hipMemcpyAsync(p->arena_dev[fmt], src, size_t(bytes), hipMemcpyDefault, static_cast<hipStream_t>(stream)) ==
                   hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

int hmvm_plan_unpack_block(const hmvm_plan* p, int64_t block, const void* const host_arenas[4], int as_f64,
                           double* a, double* b) {
    if (!p || !host_arenas) return HMVM_ERR_INVALID_ARG;
    return unpack_block(p->L, block, host_arenas, as_f64 != 0, a, b);
}

int hmvm_batch_create(const hmvm_plan* const*, int, const int*, int, hmvm_batch** out) {
    if (out) *out = nullptr;
    return HMVM_ERR_UNSUPPORTED;
}
int hmvm_batch_destroy(hmvm_batch* b) { delete b; return HMVM_OK; }
int hmvm_batch_apply(hmvm_batch*, const double*, const void* const*, const int64_t*, void* const*, const int64_t*,
                     void*) { return HMVM_ERR_UNSUPPORTED; }
int hmvm_batch_capture(hmvm_batch*, const double*, const void* const*, const int64_t*, void* const*, const int64_t*,
                       void*) { return HMVM_ERR_UNSUPPORTED; }
int hmvm_batch_replay(hmvm_batch*, void*) { return HMVM_ERR_UNSUPPORTED; }
int hmvm_batch_tune(hmvm_batch*, const void* const*, const int64_t*, void* const*, const int64_t*, void*, int*) {
    return HMVM_ERR_UNSUPPORTED;
}

} // extern "C"
