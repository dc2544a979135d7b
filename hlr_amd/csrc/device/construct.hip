// construct.hip -- low-rank leaves built on the device: adaptive cross
// approximation (ACA) with maximal-pivot search, then SVD truncation to a
// relative accuracy, one workgroup per leaf.
//
// Reference behaviour (steps restated, not code):
//   ACA          approx/aca.hh:117-223 (pivot search), :325-434 (iteration,
//                stopping on |u_k| |v_k| < eps |M_k|, norm update),
//                bem/aca.hh:59-78 (ACA, then truncation)
//   truncation   approx/svd.hh:90-209 (QR of both factors, SVD of
//                R_U R_V^T, rank from the singular values, factors
//                Q_U U_s and Q_V V_s; a rank as large as the block
//                re-factors the dense product, :118-131)
// The host generator (hlrgen.cpp, aca_block / truncate_lr) runs the same
// steps in the same order; the device keeps every column update, pivot rule
// and stopping test, with block-wide reductions in place of sequential sums.
//
// Work per leaf (m x n, ACA rank r <= 128): ACA O((m + n) r^2), Householder
// QR of both factors O((m + n) r^2), one-sided Jacobi SVD of R_U R_V^T on
// one wave, new factors O((m + n) r k).  Factors live in a per-leaf
// workspace in HBM; leaves are processed largest first, in chunks whose
// workspace fits a fixed budget.  Every leaf first runs with room for ACA
// rank 64; the few whose ACA needs more run again with room for 128.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <type_traits>
#include <vector>

#include "hmvm.h"

namespace {

constexpr int kCapSmall = 64, kCapLarge = 128;   // ACA ranks a leaf has room for (two passes)
constexpr int kT = 256;             // threads per leaf
constexpr int kWv = kT / 64;        // waves per leaf
constexpr int kVec = 16;            // dot products reduced together
constexpr int64_t kChunkBytes = int64_t(32) << 30;  // workspace budget per launch (at most half the free HBM)

struct Guard {
    int prev = -1;
    bool ok = true;
    explicit Guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) { ok = false; return; }
        if (prev != dev && hipSetDevice(dev) != hipSuccess) ok = false;
    }
    ~Guard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

struct KFn {
    int32_t kind;
    int64_t n;
    double h, ell, sigma;
    const double* X;                // n x 3, internal order
    const int64_t* perm;            // internal -> external
};

// a(i, j) of internal indices i != j (the leaves hold no diagonal entry);
// the same formulas as the host generator, without contraction, so that
// the evaluations round alike
#pragma clang fp contract(off)
__device__ double kfn_eval(const KFn& K, int64_t i, int64_t j) {
    switch (K.kind) {
    case HMVM_KFN_LOG:
        return log(fabs(K.X[3 * i] - K.X[3 * j]));
    case HMVM_KFN_LOG_GALERKIN: {
        const int64_t a = K.perm[i], b = K.perm[j];
        const double h = K.h;
        const double d = h * (fabs(double(a - b)) - 1.0);
        const double t1 = d + h, t2 = d + 2 * h;
        double v = -1.5 * h * h + 0.5 * t2 * t2 * log(t2) - t1 * t1 * log(t1);
        if (fabs(d) > 1e-8) v += 0.5 * d * d * log(d);
        return -v;
    }
    case HMVM_KFN_MATERN: {
        const double dx = K.X[3 * i] - K.X[3 * j], dy = K.X[3 * i + 1] - K.X[3 * j + 1];
        const double dz = K.X[3 * i + 2] - K.X[3 * j + 2];
        const double s = sqrt(3.0) * sqrt(dx * dx + dy * dy + dz * dz) / K.ell;
        return K.sigma * K.sigma * (1 + s) * exp(-s);
    }
    default: {
        const double dx = K.X[3 * i] - K.X[3 * j], dy = K.X[3 * i + 1] - K.X[3 * j + 1];
        const double dz = K.X[3 * i + 2] - K.X[3 * j + 2];
        return 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    }
    }
}


struct Leaf {
    int64_t row0, col0;
    int32_t m, n;
    int64_t ws;                     // first double of the leaf's workspace
};

// workspace of a leaf with room for ACA rank CAP (doubles): U, HU, QU
// (m x CAP each), V, HV, QV (n x CAP each), Vj (CAP x CAP), S (CAP), used
// columns (n bytes)
inline int64_t ws_doubles(int64_t m, int64_t n, int cap) {
    return 3 * (m + n) * cap + int64_t(cap) * cap + cap + (n + 7) / 8 + 32;
}

struct Ws {
    double *U, *HU, *QU, *V, *HV, *QV, *Vj, *S;
    unsigned char* used;
};

template <int CAP>
__device__ inline Ws ws_of(double* base, const Leaf& L) {
    Ws w;
    const int64_t m = L.m, n = L.n;
    double* p = base + L.ws;
    w.U = p; p += m * CAP;
    w.HU = p; p += m * CAP;
    w.QU = p; p += m * CAP;
    w.V = p; p += n * CAP;
    w.HV = p; p += n * CAP;
    w.QV = p; p += n * CAP;
    w.Vj = p; p += CAP * CAP;
    w.S = p; p += CAP;
    w.used = reinterpret_cast<unsigned char*>(p);
    return w;
}

// ---- block-wide reductions (fixed order: lanes by a shuffle tree, then
//      the waves in order), every thread returns the result ----

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
    return v;
}

__device__ inline double block_sum(double v, double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int w = 0; w < kWv; ++w) s += red[w];
    return s;
}

// the largest a (first index on ties; a < 0: none), every thread returns it
struct AMax {
    double a;                       // |value| compared
    int64_t i;
};

__device__ inline bool amax_better(const AMax& x, const AMax& y) {
    return x.a > y.a || (x.a == y.a && x.i >= 0 && (y.i < 0 || x.i < y.i));
}

__device__ inline AMax block_amax(AMax v, double* reda, int64_t* redi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        AMax o{__shfl_xor(v.a, s, 64), __shfl_xor(v.i, s, 64)};
        if (amax_better(o, v)) v = o;
    }
    __syncthreads();
    if (lane == 0) { reda[wave] = v.a; redi[wave] = v.i; }
    __syncthreads();
    AMax r{reda[0], redi[0]};
#pragma unroll
    for (int w = 1; w < kWv; ++w) {
        const AMax o{reda[w], redi[w]};
        if (amax_better(o, r)) r = o;
    }
    return r;
}

// out[c] = sum_i A[c * lda + i] b[i] over i in [i0, m) for c < nc (A column-major)
__device__ void block_dots(const double* A, int64_t lda, int nc, const double* b, int64_t i0, int64_t m,
                           double* out, double* red) {
    for (int c0 = 0; c0 < nc; c0 += kVec) {
        double acc[kVec];
#pragma unroll
        for (int u = 0; u < kVec; ++u) acc[u] = 0;
        for (int64_t i = i0 + threadIdx.x; i < m; i += kT) {
            const double bi = b[i];
#pragma unroll
            for (int u = 0; u < kVec; ++u)
                if (c0 + u < nc) acc[u] += A[int64_t(c0 + u) * lda + i] * bi;
        }
#pragma unroll
        for (int u = 0; u < kVec; ++u) {
            if (c0 + u >= nc) break;                  // uniform
            const double s = block_sum(acc[u], red);
            if (threadIdx.x == 0) out[c0 + u] = s;
        }
    }
    __syncthreads();
}

// ---- ACA with maximal-pivot search (approx/aca.hh:117-223, 325-434) ----
// Returns the rank (U: m x k, V: n x k, column-major, ld m / n), or -1 when
// CAP steps did not converge.
template <int CAP>
__device__ int aca_leaf(const KFn& K, const Leaf& L, const Ws& w, double eps, double* red, double* reda,
                        int64_t* redi, double* sv) {
    const int64_t m = L.m, n = L.n;
    const int tid = threadIdx.x;
    const double zero_val = DBL_EPSILON * DBL_EPSILON;
    const int64_t max_rank = min(m, n);
    for (int64_t j = tid; j < (n + 7) / 8; j += kT) w.used[j] = 0;
    __syncthreads();
    int64_t next_col = 0;
    int k = 0;
    double normM = 0;
    for (int64_t it = 0; it < max_rank; ++it) {
        if (k == CAP) return -1;
        const int64_t pc = next_col;
        if (tid == 0) w.used[pc >> 3] |= (unsigned char)(1u << (pc & 7));
        for (int l = tid; l < k; l += kT) sv[l] = w.V[int64_t(l) * n + pc];
        __syncthreads();
        // column pc of the residual, its largest entry
        double* col = w.U + int64_t(k) * m;
        AMax best{-1.0, -1};
        for (int64_t i = tid; i < m; i += kT) {
            double c = kfn_eval(K, L.row0 + i, L.col0 + pc);
            for (int l = 0; l < k; ++l) c -= sv[l] * w.U[int64_t(l) * m + i];
            col[i] = c;
            if (fabs(c) > best.a) best = AMax{fabs(c), i};
        }
        best = block_amax(best, reda, redi);
        const int64_t pr = best.i;                      // -1 only for a column of NaNs
        if (pr < 0) break;
        const double maxval = col[pr];
        if (fabs(maxval) <= zero_val) break;
        for (int64_t i = tid; i < m; i += kT) col[i] /= maxval;
        for (int l = tid; l < k; l += kT) sv[l] = w.U[int64_t(l) * m + pr];
        __syncthreads();
        // row pr of the residual; next column: its largest entry among the
        // unused columns (the first unused column when that row is zero)
        double* row = w.V + int64_t(k) * n;
        AMax nb{0.0, -1}, first{0.0, -1};
        for (int64_t j = tid; j < n; j += kT) {
            double r = kfn_eval(K, L.row0 + pr, L.col0 + j);
            for (int l = 0; l < k; ++l) r -= sv[l] * w.V[int64_t(l) * n + j];
            row[j] = r;
            const bool unused = !((w.used[j >> 3] >> (j & 7)) & 1u);
            if (unused && fabs(r) > nb.a) nb = AMax{fabs(r), j};
            if (unused && first.i < 0) first = AMax{1.0, j};
        }
        nb = block_amax(nb, reda, redi);
        first = block_amax(first, reda, redi);
        next_col = (nb.i >= 0) ? nb.i : first.i;
        // |u_k| |v_k| against eps |M_k| (norm of the approximation so far)
        double sq = 0;
        for (int64_t i = tid; i < m; i += kT) sq += col[i] * col[i];
        const double nu = sqrt(block_sum(sq, red));
        sq = 0;
        for (int64_t j = tid; j < n; j += kT) sq += row[j] * row[j];
        const double nv = sqrt(block_sum(sq, red));
        const double norm_i = nu * nv;
        const bool last = (norm_i < eps * normM);
        if (!last) {
            double upd = norm_i * norm_i;
            if (k > 0) {
                block_dots(w.U, m, k, col, 0, m, sv, red);
                block_dots(w.V, n, k, row, 0, n, sv + CAP, red);
                for (int l = 0; l < k; ++l) upd += 2 * sv[l] * sv[CAP + l];
            }
            normM = sqrt(normM * normM + fabs(upd));
        }
        ++k;
        __syncthreads();
        if (last || next_col < 0) break;
    }
    __syncthreads();
    return k;
}

// ---- Householder QR of A (m x r, column-major, in place) ----
// On return the upper triangle of A holds R, H holds the reflectors
// (column j: rows j..m-1, unit 2-norm, or zero) and Q = H_0 ... H_{r-1}
// [I_r; 0] (m x r).
__device__ void house_qr(double* A, double* H, double* Q, int64_t m, int r, double* red, double* sv) {
    const int tid = threadIdx.x;
    for (int j = 0; j < r; ++j) {
        double* aj = A + int64_t(j) * m;
        double* v = H + int64_t(j) * m;
        double sq = 0;
        for (int64_t i = j + tid; i < m; i += kT) sq += aj[i] * aj[i];
        const double nx = sqrt(block_sum(sq, red));
        if (nx == 0) {
            for (int64_t i = tid; i < m; i += kT) v[i] = 0;
            __syncthreads();
            continue;
        }
        const double alpha = (aj[j] >= 0) ? -nx : nx;
        sq = 0;
        for (int64_t i = tid; i < m; i += kT) {
            const double vi = (i < j) ? 0.0 : (i == j ? aj[i] - alpha : aj[i]);
            v[i] = vi;
            sq += vi * vi;
        }
        const double nv = sqrt(block_sum(sq, red));
        for (int64_t i = tid; i < m; i += kT) v[i] = (nv == 0) ? 0.0 : v[i] / nv;
        __syncthreads();
        if (nv == 0) continue;
        // columns j..r-1: a_c -= 2 (v . a_c) v
        block_dots(A + int64_t(j) * m, m, r - j, v, j, m, sv, red);
        for (int64_t i = j + tid; i < m; i += kT) {
            const double vi = v[i];
            for (int c = j; c < r; ++c) A[int64_t(c) * m + i] -= 2 * sv[c - j] * vi;
        }
        __syncthreads();
    }
    for (int c = 0; c < r; ++c)
        for (int64_t i = tid; i < m; i += kT) Q[int64_t(c) * m + i] = (i == c) ? 1.0 : 0.0;
    __syncthreads();
    for (int j = r - 1; j >= 0; --j) {
        const double* v = H + int64_t(j) * m;
        block_dots(Q, m, r, v, j, m, sv, red);
        for (int64_t i = j + tid; i < m; i += kT) {
            const double vi = v[i];
            for (int c = 0; c < r; ++c) Q[int64_t(c) * m + i] -= 2 * sv[c] * vi;
        }
        __syncthreads();
    }
}

// ---- one-sided Jacobi SVD (Hestenes) of the r x r matrix W (LDS, column
//      major, ld CAP), on wave 0: lane l holds rows l, l + 64, ...; Vj
//      (global, ld CAP) gets the right vectors.  Pairs in cyclic order,
//      sweeps until no rotation. ----
template <int CAP>
__device__ void jacobi_wave(double* W, double* Vj, int r) {
    constexpr int RPL = CAP / 64;                   // rows per lane
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < r; ++c)
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int i = lane + 64 * t;
            if (i < r) Vj[c * CAP + i] = (i == c) ? 1.0 : 0.0;
        }
    const double tol = 1e-15;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < r - 1; ++p) {
            for (int q = p + 1; q < r; ++q) {
                double wp[RPL], wq[RPL], a = 0, b = 0, g = 0;
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    const int i = lane + 64 * t;
                    wp[t] = (i < r) ? W[p * CAP + i] : 0.0;
                    wq[t] = (i < r) ? W[q * CAP + i] : 0.0;
                    a += wp[t] * wp[t];
                    b += wq[t] * wq[t];
                    g += wp[t] * wq[t];
                }
                a = wave_sum(a);
                b = wave_sum(b);
                g = wave_sum(g);
                if (g == 0 || fabs(g) <= tol * sqrt(a * b)) continue;
                rotated = true;
                const double zeta = (b - a) / (2 * g);
                const double t0 = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                const double c = 1 / sqrt(1 + t0 * t0);
                const double s = c * t0;
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    const int i = lane + 64 * t;
                    if (i < r) {
                        W[p * CAP + i] = c * wp[t] - s * wq[t];
                        W[q * CAP + i] = s * wp[t] + c * wq[t];
                        const double vp = Vj[p * CAP + i], vq = Vj[q * CAP + i];
                        Vj[p * CAP + i] = c * vp - s * vq;
                        Vj[q * CAP + i] = s * vp + c * vq;
                    }
                }
            }
        }
        if (!rotated) break;
    }
}

template <int CAP>
__global__ __launch_bounds__(kT) void k_lowrank_leaf(KFn K, const Leaf* __restrict__ leaves, double* ws, double eps,
                                                     int truncate, int32_t* __restrict__ rank_out) {
    __shared__ double red[kWv], reda[kWv];
    __shared__ int64_t redi[kWv];
    __shared__ double sv[2 * CAP];
    __shared__ double W[CAP * CAP];                 // R_U R_V^T, then its left vectors
    __shared__ double sval[CAP];
    __shared__ int order[CAP];
    const Leaf L = leaves[blockIdx.x];
    const Ws w = ws_of<CAP>(ws, L);
    const int tid = threadIdx.x;
    const int64_t m = L.m, n = L.n;
    int r = aca_leaf<CAP>(K, L, w, eps, red, reda, redi, sv);
    if (r < 0 || !truncate || r == 0) {
        if (tid == 0) rank_out[blockIdx.x] = r;
        return;
    }
    // a rank as large as the block: re-factor the dense product
    // (approx/svd.hh:118-131).  min(m, n) = r <= CAP, so the product fits
    // the Q workspace of the longer side.
    if (r >= min(m, n)) {
        if (n <= m) {
            for (int64_t j = 0; j < n; ++j)
                for (int64_t i = tid; i < m; i += kT) {
                    double s = 0;
                    for (int l = 0; l < r; ++l) s += w.U[int64_t(l) * m + i] * w.V[int64_t(l) * n + j];
                    w.QU[j * m + i] = s;
                }
            __syncthreads();
            for (int64_t j = 0; j < n; ++j)
                for (int64_t i = tid; i < m; i += kT) w.U[j * m + i] = w.QU[j * m + i];
            for (int64_t c = 0; c < n; ++c)
                for (int64_t i = tid; i < n; i += kT) w.V[c * n + i] = (i == c) ? 1.0 : 0.0;
            r = int(n);
        } else {
            for (int64_t i = 0; i < m; ++i)
                for (int64_t j = tid; j < n; j += kT) {
                    double s = 0;
                    for (int l = 0; l < r; ++l) s += w.U[int64_t(l) * m + i] * w.V[int64_t(l) * n + j];
                    w.QV[i * n + j] = s;            // V = M^T (n x m)
                }
            __syncthreads();
            for (int64_t i = 0; i < m; ++i)
                for (int64_t j = tid; j < n; j += kT) w.V[i * n + j] = w.QV[i * n + j];
            for (int64_t c = 0; c < m; ++c)
                for (int64_t i = tid; i < m; i += kT) w.U[c * m + i] = (i == c) ? 1.0 : 0.0;
            r = int(m);
        }
        __syncthreads();
    }
    house_qr(w.U, w.HU, w.QU, m, r, red, sv);
    house_qr(w.V, w.HV, w.QV, n, r, red, sv);
    // W = R_U R_V^T (upper triangles of U and V)
    for (int e = tid; e < r * r; e += kT) {
        const int i = e % r, j = e / r;
        double s = 0;
        for (int l = max(i, j); l < r; ++l) s += w.U[int64_t(l) * m + i] * w.V[int64_t(l) * n + j];
        W[j * CAP + i] = s;
    }
    __syncthreads();
    if (tid < 64) jacobi_wave<CAP>(W, w.Vj, r);
    __syncthreads();
    // singular values (column norms), descending (stable)
    if (tid < 64) {
        for (int c = 0; c < r; ++c) {
            double x2 = 0;
#pragma unroll
            for (int t = 0; t < CAP / 64; ++t) {
                const int i = tid + 64 * t;
                const double x = (i < r) ? W[c * CAP + i] : 0.0;
                x2 += x * x;
            }
            const double s = sqrt(wave_sum(x2));
            if (tid == 0) sval[c] = s;
        }
    }
    __syncthreads();
    if (tid < r) {
        int pos = 0;
        for (int c = 0; c < r; ++c)
            pos += (sval[c] > sval[tid] || (sval[c] == sval[tid] && c < tid)) ? 1 : 0;
        order[pos] = tid;
    }
    __syncthreads();
    const double s0 = sval[order[0]];
    int k = 0;
    if (s0 != 0)
        while (k < r && sval[order[k]] > eps * s0) ++k;
    // new factors: U_o(:, c) = Q_U U_s(:, c), V_o(:, c) = Q_V V_s(:, c)
    for (int64_t i = tid; i < m; i += kT)
        for (int c = 0; c < k; ++c) {
            const int o = order[c];
            const double so = sval[o];
            double s = 0;
            for (int l = 0; l < r; ++l) s += w.QU[int64_t(l) * m + i] * (W[o * CAP + l] / so);
            w.U[int64_t(c) * m + i] = s;
        }
    for (int64_t i = tid; i < n; i += kT)
        for (int c = 0; c < k; ++c) {
            const int o = order[c];
            double s = 0;
            for (int l = 0; l < r; ++l) s += w.QV[int64_t(l) * n + i] * w.Vj[o * CAP + l];
            w.V[int64_t(c) * n + i] = s;
        }
    if (tid < k) w.S[tid] = sval[order[tid]];
    if (tid == 0) rank_out[blockIdx.x] = k;
}

// pack each leaf's U (m x k), V (n x k) and S (k; ones without truncation)
// at its offset of the output buffer: U, then V, then S
template <int CAP>
__global__ __launch_bounds__(kT) void k_lowrank_pack(const Leaf* __restrict__ leaves, const double* __restrict__ ws,
                                                     const int32_t* __restrict__ rank, const int64_t* __restrict__ off,
                                                     int truncate, double* __restrict__ out) {
    const Leaf L = leaves[blockIdx.x];
    const int k = rank[blockIdx.x];
    if (k <= 0) return;
    const double* U = ws + L.ws;
    const double* V = U + 3 * int64_t(L.m) * CAP;
    const double* S = V + 3 * int64_t(L.n) * CAP + CAP * CAP;
    double* o = out + off[blockIdx.x];
    const int64_t nu = int64_t(L.m) * k, nv = int64_t(L.n) * k;
    for (int64_t e = threadIdx.x; e < nu; e += kT) o[e] = U[e];
    for (int64_t e = threadIdx.x; e < nv; e += kT) o[nu + e] = V[e];
    for (int e = threadIdx.x; e < k; e += kT) o[nu + nv + e] = truncate ? S[e] : 1.0;
}

} // namespace

struct hmvm_lowrank {
    int64_t nleaves = 0;
    std::vector<hmvm_lr_leaf> leaves;
    std::vector<int32_t> rank;
    std::vector<int64_t> off;       // per leaf: first value of U, V, S in `data`
    std::unique_ptr<double[]> data; // packed factors (host)
    int64_t ndata = 0;
    int truncate = 1;
    double seconds = 0;             // whole build (device events)
    double kernel_seconds = 0;      // the chunks' kernels
    double copy_seconds = 0;        // factors to the host
};

extern "C" int hmvm_lowrank_build(const hmvm_kernel_fn* fn, const hmvm_lr_leaf* leaves, int64_t nleaves, double eps,
                                  int truncate, int device, hmvm_lowrank** out) {
    if (!fn || !out || nleaves < 0 || (nleaves > 0 && !leaves) || !(eps > 0)) return HMVM_ERR_INVALID_ARG;
    if (fn->kind < HMVM_KFN_LOG || fn->kind > HMVM_KFN_INVR || fn->n < 1 || !fn->coords) return HMVM_ERR_INVALID_ARG;
    if (fn->kind == HMVM_KFN_LOG_GALERKIN && !fn->perm) return HMVM_ERR_INVALID_ARG;
    *out = nullptr;
    for (int64_t i = 0; i < nleaves; ++i) {
        const hmvm_lr_leaf& l = leaves[i];
        if (l.nrows < 1 || l.ncols < 1 || l.row0 < 0 || l.col0 < 0 || l.row0 > fn->n - l.nrows ||
            l.col0 > fn->n - l.ncols)
            return HMVM_ERR_BAD_BLOCK;
        // no diagonal entry: the row and column ranges are disjoint
        if (l.row0 < l.col0 + l.ncols && l.col0 < l.row0 + l.nrows) return HMVM_ERR_BAD_BLOCK;
    }
    std::unique_ptr<hmvm_lowrank> R;
    try {
        R.reset(new hmvm_lowrank);
        R->nleaves = nleaves;
        R->leaves.assign(leaves, leaves + nleaves);
        R->rank.assign(nleaves, 0);
        R->off.assign(nleaves, 0);
    } catch (...) { return HMVM_ERR_ALLOC; }
    R->truncate = truncate ? 1 : 0;
    Guard g(device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    // e0 / e1: the whole build; k0 / k1: each chunk's leaf kernel
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    for (hipEvent_t& e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            for (hipEvent_t& f : ev)
                if (f) (void)hipEventDestroy(f);
            return HMVM_ERR_DEVICE;
        }
    hipEvent_t e0 = ev[0], e1 = ev[1], k0 = ev[2], k1 = ev[3];
    (void)hipEventRecord(e0, nullptr);
    // problem data on the device
    const int64_t np = fn->n;
    double* dX = nullptr;
    int64_t* dperm = nullptr;
    int rc = HMVM_OK;
    if (hipMalloc(&dX, size_t(np) * 3 * sizeof(double)) != hipSuccess) rc = HMVM_ERR_ALLOC;
    if (rc == HMVM_OK && fn->perm && hipMalloc(&dperm, size_t(np) * sizeof(int64_t)) != hipSuccess) rc = HMVM_ERR_ALLOC;
    if (rc == HMVM_OK && hipMemcpy(dX, fn->coords, size_t(np) * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = HMVM_ERR_DEVICE;
    if (rc == HMVM_OK && dperm && hipMemcpy(dperm, fn->perm, size_t(np) * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess)
        rc = HMVM_ERR_DEVICE;
    KFn K{fn->kind, np, fn->h, fn->ell, fn->sigma, dX, dperm};
    // largest leaves first; chunks of at most `budget` bytes of workspace
    std::vector<int64_t> ord(nleaves);
    std::iota(ord.begin(), ord.end(), int64_t(0));
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        return int64_t(leaves[a].nrows) + leaves[a].ncols > int64_t(leaves[b].nrows) + leaves[b].ncols;
    });
    size_t free_b = 0, total_b = 0;
    if (rc == HMVM_OK && hipMemGetInfo(&free_b, &total_b) != hipSuccess) rc = HMVM_ERR_DEVICE;
    const int64_t budget = std::max<int64_t>(int64_t(64) << 20, std::min<int64_t>(kChunkBytes, int64_t(free_b / 2)));
    // packed factors of each chunk stay on the device until every pass is
    // done; then one host buffer of the exact size takes them
    std::vector<std::pair<double*, int64_t>> outs;  // device buffer, values
    int64_t total = 0;
    double kernel_ms = 0;
    double* dws = nullptr;
    Leaf* dleaf = nullptr;
    int32_t* drank = nullptr;
    int64_t* doff = nullptr;
    int64_t ws_cap = 0, leaf_cap = 0;
    auto grow = [&](auto*& p, int64_t& cap, int64_t need) {
        if (need <= cap) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, size_t(need) * sizeof(*p)) != hipSuccess) return false;
        cap = need;
        return true;
    };
    int64_t leaf_cap_r = 0, leaf_cap_o = 0;
    // one pass over the leaves `todo` (indices into `leaves`, largest first)
    // with room for ACA rank CAP; leaves whose ACA needs more go to `again`
    std::vector<int64_t> again;
    auto pass = [&](auto capc, const std::vector<int64_t>& todo) {
        constexpr int CAP = decltype(capc)::value;
        const int64_t nt = int64_t(todo.size());
        for (int64_t c0 = 0; rc == HMVM_OK && c0 < nt;) {
            // the chunk: todo[c0 .. c1)
            std::vector<Leaf> chunk;
            int64_t wsz = 0, c1 = c0;
            while (c1 < nt) {
                const hmvm_lr_leaf& l = leaves[todo[c1]];
                const int64_t need = ws_doubles(l.nrows, l.ncols, CAP);
                if (!chunk.empty() && (wsz + need) * 8 > budget) break;
                chunk.push_back(Leaf{l.row0, l.col0, l.nrows, l.ncols, wsz});
                wsz += (need + 31) / 32 * 32;
                ++c1;
            }
            const int64_t nc = int64_t(chunk.size());
            if (!grow(dws, ws_cap, wsz) || !grow(dleaf, leaf_cap, nc) || !grow(drank, leaf_cap_r, nc) ||
                !grow(doff, leaf_cap_o, nc)) { rc = HMVM_ERR_ALLOC; break; }
            if (hipMemcpy(dleaf, chunk.data(), size_t(nc) * sizeof(Leaf), hipMemcpyHostToDevice) != hipSuccess) {
                rc = HMVM_ERR_DEVICE; break;
            }
            (void)hipEventRecord(k0, nullptr);
            k_lowrank_leaf<CAP><<<unsigned(nc), kT, 0, nullptr>>>(K, dleaf, dws, eps, R->truncate, drank);
            if (hipGetLastError() != hipSuccess) { rc = HMVM_ERR_KERNEL; break; }
            (void)hipEventRecord(k1, nullptr);
            std::vector<int32_t> rk(nc);
            if (hipMemcpy(rk.data(), drank, size_t(nc) * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
                rc = HMVM_ERR_DEVICE; break;
            }
            std::vector<int64_t> off(nc);
            int64_t osz = 0;
            for (int64_t i = 0; i < nc; ++i) {
                off[i] = osz;
                if (rk[i] > 0) osz += (int64_t(chunk[i].m) + chunk[i].n + 1) * rk[i];
            }
            {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, k0, k1);
                kernel_ms += ms;
            }
            if (osz > 0) {
                double* dout = nullptr;
                if (hipMalloc(&dout, size_t(osz) * sizeof(double)) != hipSuccess) { rc = HMVM_ERR_ALLOC; break; }
                outs.push_back({dout, osz});
                if (hipMemcpy(doff, off.data(), size_t(nc) * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
                    rc = HMVM_ERR_DEVICE; break;
                }
                k_lowrank_pack<CAP><<<unsigned(nc), kT, 0, nullptr>>>(dleaf, dws, drank, doff, R->truncate, dout);
                if (hipGetLastError() != hipSuccess) { rc = HMVM_ERR_KERNEL; break; }
            }
            for (int64_t i = 0; i < nc; ++i) {
                if (rk[i] < 0) { again.push_back(todo[c0 + i]); continue; }
                R->rank[todo[c0 + i]] = rk[i];
                R->off[todo[c0 + i]] = total + off[i];
            }
            total += osz;
            c0 = c1;
        }
    };
    pass(std::integral_constant<int, kCapSmall>(), ord);
    if (rc == HMVM_OK && !again.empty()) {
        std::vector<int64_t> todo;
        todo.swap(again);
        pass(std::integral_constant<int, kCapLarge>(), todo);
        if (rc == HMVM_OK && !again.empty()) rc = HMVM_ERR_UNSUPPORTED;     // ACA rank beyond kCapLarge
    }
    // the workspace goes before the host copies
    for (void* p : {static_cast<void*>(dws), static_cast<void*>(dleaf), static_cast<void*>(drank),
                    static_cast<void*>(doff)})
        if (p) (void)hipFree(p);
    dws = nullptr; dleaf = nullptr; drank = nullptr; doff = nullptr;
    if (rc == HMVM_OK && total > 0) {
        hipEvent_t c0, c1;
        (void)hipEventCreate(&c0);
        (void)hipEventCreate(&c1);
        (void)hipEventRecord(c0, nullptr);
        try { R->data.reset(new double[size_t(total)]); } catch (...) { rc = HMVM_ERR_ALLOC; }
        if (rc == HMVM_OK) {
            // fault the pages in on all host threads: the copies below then
            // stream into resident memory instead of faulting page by page
            double* d = R->data.get();
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < total; i += 512) d[i] = 0.0;
        }
        int64_t pos = 0;
        for (auto& o : outs) {
            if (rc != HMVM_OK) break;
            if (hipMemcpy(R->data.get() + pos, o.first, size_t(o.second) * sizeof(double), hipMemcpyDeviceToHost) !=
                hipSuccess) rc = HMVM_ERR_DEVICE;
            pos += o.second;
        }
        (void)hipEventRecord(c1, nullptr);
        (void)hipEventSynchronize(c1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c0, c1);
        R->copy_seconds = ms * 1e-3;
        (void)hipEventDestroy(c0);
        (void)hipEventDestroy(c1);
    }
    R->ndata = total;
    R->kernel_seconds = kernel_ms * 1e-3;
    for (auto& o : outs) (void)hipFree(o.first);
    (void)hipEventDestroy(k0);
    (void)hipEventDestroy(k1);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    R->seconds = ms * 1e-3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (void* p : {static_cast<void*>(dX), static_cast<void*>(dperm)})
        if (p) (void)hipFree(p);
    if (rc != HMVM_OK) return rc;
    *out = R.release();
    return HMVM_OK;
}

extern "C" int hmvm_lowrank_ranks(const hmvm_lowrank* lr, int32_t* ranks) {
    if (!lr || (!ranks && lr->nleaves > 0)) return HMVM_ERR_INVALID_ARG;
    std::copy(lr->rank.begin(), lr->rank.end(), ranks);
    return HMVM_OK;
}

extern "C" int hmvm_lowrank_copy(const hmvm_lowrank* lr, const int64_t* off_a, const int64_t* off_b,
                                 const int64_t* off_s, int fold, double* arena, int64_t arena_len) {
    if (!lr || (lr->nleaves > 0 && (!off_a || !off_b || !arena))) return HMVM_ERR_INVALID_ARG;
    auto inside = [&](int64_t o, int64_t len) { return o >= 0 && len >= 0 && len <= arena_len && o <= arena_len - len; };
    for (int64_t i = 0; i < lr->nleaves; ++i) {
        const int64_t k = lr->rank[i], m = lr->leaves[i].nrows, n = lr->leaves[i].ncols;
        if (k == 0) continue;
        if (!inside(off_a[i], m * k) || !inside(off_b[i], n * k)) return HMVM_ERR_INVALID_ARG;
        if (off_s && off_s[i] >= 0 && !inside(off_s[i], k)) return HMVM_ERR_INVALID_ARG;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < lr->nleaves; ++i) {
        const int64_t k = lr->rank[i], m = lr->leaves[i].nrows, n = lr->leaves[i].ncols;
        if (k == 0) continue;
        const double* src = lr->data.get() + lr->off[i];
        const double* S = src + (m + n) * k;
        std::copy(src, src + m * k, arena + off_a[i]);
        std::copy(src + m * k, src + (m + n) * k, arena + off_b[i]);
        if (off_s && off_s[i] >= 0) std::copy(S, S + k, arena + off_s[i]);
        if (fold) {
            // S into the factor with fewer rows (approx/svd.hh:206-207)
            double* F = (m < n) ? arena + off_a[i] : arena + off_b[i];
            const int64_t rows = (m < n) ? m : n;
            for (int64_t l = 0; l < k; ++l)
                for (int64_t r = 0; r < rows; ++r) F[l * rows + r] *= S[l];
        }
    }
    return HMVM_OK;
}

extern "C" double hmvm_lowrank_seconds(const hmvm_lowrank* lr) { return lr ? lr->seconds : 0.0; }

extern "C" int hmvm_lowrank_timing(const hmvm_lowrank* lr, double* kernel_seconds, double* copy_seconds) {
    if (!lr || !kernel_seconds || !copy_seconds) return HMVM_ERR_INVALID_ARG;
    *kernel_seconds = lr->kernel_seconds;
    *copy_seconds = lr->copy_seconds;
    return HMVM_OK;
}

extern "C" void hmvm_lowrank_free(hmvm_lowrank* lr) { delete lr; }
