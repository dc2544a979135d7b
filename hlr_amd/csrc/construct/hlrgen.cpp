// hlrgen.cpp -- H-/TLR-matrix generator (host C++17, OpenMP).
//
// Builds the matrices the product runs on.  Construction is not part of the
// device path; it is written for throughput on a many-core host so that the
// n = 262144 matrices are generated in seconds to a minute.
//
// Reference behaviour restated here (semantics, not code):
//   points      src/apps/radial.cc:28-73 (seeded std::mt19937_64{1}),
//               src/apps/log_kernel.cc:94-113 (grid cell centres)
//   kernels     src/apps/log_kernel.cc:37-68 (Galerkin log entries),
//               matrix/radial.hh:295-340 (Matern covariance)
//   clustering  src/cluster/h.cc:22-29 (cardinality-balanced BSP, adaptive
//               split axis), src/cluster/tlr.cc:22-29 (one-level tiling)
//   admissible  programs/common/common.hh:279-290 (std. geometric, eta)
//   build       seq/matrix.hh:51-176 (leaf-wise build), bem/aca.hh:59-78
//   ACA         approx/aca.hh:117-223 (max pivot), :325-434 (iteration)
//   truncation  approx/svd.hh:90-209 (QR of both factors, SVD of R_U R_V^T)

#include "hlrgen.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <numeric>
#include <random>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// dense linear algebra helpers (column-major)
// ---------------------------------------------------------------------------

inline double dot(int64_t n, const double* a, const double* b) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

inline double nrm2(int64_t n, const double* a) {
    // scaled to avoid overflow/underflow of the squares
    double scale = 0, ssq = 1;
    for (int64_t i = 0; i < n; ++i) {
        const double v = std::fabs(a[i]);
        if (v == 0) continue;
        if (scale < v) { ssq = 1 + ssq * (scale / v) * (scale / v); scale = v; }
        else           { ssq += (v / scale) * (v / scale); }
    }
    return scale * std::sqrt(ssq);
}

// Householder QR, A (m x k), m >= k.  Q (m x k) explicit, R (k x k) upper.
void householder_qr(int64_t m, int64_t k, const double* A, double* Q, double* R) {
    std::vector<double> W(A, A + m * k);
    std::vector<double> vs(m * k, 0.0);     // reflector j stored in column j (rows j..m-1)
    for (int64_t j = 0; j < k; ++j) {
        double* wj = &W[j * m];
        const double nx = nrm2(m - j, wj + j);
        double* v = &vs[j * m];
        if (nx == 0) continue;
        const double alpha = (wj[j] >= 0) ? -nx : nx;
        for (int64_t i = j; i < m; ++i) v[i] = wj[i];
        v[j] -= alpha;
        const double nv = nrm2(m - j, v + j);
        if (nv == 0) { std::fill(v + j, v + m, 0.0); continue; }
        for (int64_t i = j; i < m; ++i) v[i] /= nv;
        for (int64_t c = j; c < k; ++c) {
            double* wc = &W[c * m];
            const double s = 2 * dot(m - j, v + j, wc + j);
            for (int64_t i = j; i < m; ++i) wc[i] -= s * v[i];
        }
    }
    for (int64_t c = 0; c < k; ++c)
        for (int64_t r = 0; r < k; ++r)
            R[c * k + r] = (r <= c) ? W[c * m + r] : 0.0;
    // Q = H_0 ... H_{k-1} [I_k; 0]
    std::fill(Q, Q + m * k, 0.0);
    for (int64_t c = 0; c < k; ++c) Q[c * m + c] = 1.0;
    for (int64_t j = k - 1; j >= 0; --j) {
        const double* v = &vs[j * m];
        bool nz = false;
        for (int64_t i = j; i < m && !nz; ++i) nz = (v[i] != 0);
        if (!nz) continue;
        for (int64_t c = 0; c < k; ++c) {
            double* qc = Q + c * m;
            const double s = 2 * dot(m - j, v + j, qc + j);
            if (s != 0)
                for (int64_t i = j; i < m; ++i) qc[i] -= s * v[i];
        }
    }
}

// One-sided Jacobi SVD (Hestenes) of a k x k matrix: A = U diag(S) V^T.
void jacobi_svd(int64_t k, const double* A, double* U, double* S, double* V) {
    std::vector<double> W(A, A + k * k);
    std::vector<double> Vw(k * k, 0.0);
    for (int64_t i = 0; i < k; ++i) Vw[i * k + i] = 1.0;
    const double tol = 1e-15;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int64_t p = 0; p < k - 1; ++p) {
            for (int64_t q = p + 1; q < k; ++q) {
                double* wp = &W[p * k];
                double* wq = &W[q * k];
                const double a = dot(k, wp, wp);
                const double b = dot(k, wq, wq);
                const double g = dot(k, wp, wq);
                if (g == 0 || std::fabs(g) <= tol * std::sqrt(a * b)) continue;
                rotated = true;
                const double zeta = (b - a) / (2 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double c = 1 / std::sqrt(1 + t * t);
                const double s = c * t;
                for (int64_t i = 0; i < k; ++i) {
                    const double x = wp[i], y = wq[i];
                    wp[i] = c * x - s * y;
                    wq[i] = s * x + c * y;
                }
                double* vp = &Vw[p * k];
                double* vq = &Vw[q * k];
                for (int64_t i = 0; i < k; ++i) {
                    const double x = vp[i], y = vq[i];
                    vp[i] = c * x - s * y;
                    vq[i] = s * x + c * y;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<double> sv(k);
    for (int64_t i = 0; i < k; ++i) sv[i] = nrm2(k, &W[i * k]);
    std::vector<int64_t> order(k);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return sv[x] > sv[y]; });
    for (int64_t o = 0; o < k; ++o) {
        const int64_t i = order[o];
        S[o] = sv[i];
        for (int64_t r = 0; r < k; ++r) {
            U[o * k + r] = (sv[i] > 0) ? W[i * k + r] / sv[i] : 0.0;
            V[o * k + r] = Vw[i * k + r];
        }
    }
}

// Relative-precision rank: keep sigma_i > eps * sigma_0 (fixed_prec(eps) in
// the spectral norm; the exact HLIBpro trunc_rank rule is not available).
int64_t trunc_rank(int64_t k, const double* S, double eps) {
    if (k == 0 || S[0] == 0) return 0;
    int64_t r = 0;
    while (r < k && S[r] > eps * S[0]) ++r;
    return r;
}

// Truncate U V^T (U m x r, V n x r) -> Uo (m x k) diag(So) Vo^T (n x k).
// Follows the structure of approx/svd.hh:90-209: QR of both factors, SVD of
// R_U R_V^T, rank from the singular values, new factors Q_U U_s and Q_V V_s.
int64_t truncate_lr(int64_t m, int64_t n, int64_t r, const double* U, const double* V, double eps,
                    std::vector<double>& Uo, std::vector<double>& So, std::vector<double>& Vo) {
    if (r == 0) { Uo.clear(); So.clear(); Vo.clear(); return 0; }
    std::vector<double> Ux, Vx;
    const double* pU = U;
    const double* pV = V;
    if (r >= std::min(m, n)) {
        // rank as large as the block: re-factor the dense product (svd.hh:118-131)
        std::vector<double> M(m * n, 0.0);
        for (int64_t l = 0; l < r; ++l)
            for (int64_t j = 0; j < n; ++j) {
                const double v = V[l * n + j];
                for (int64_t i = 0; i < m; ++i) M[j * m + i] += U[l * m + i] * v;
            }
        if (n <= m) {
            Ux = std::move(M);
            Vx.assign(n * n, 0.0);
            for (int64_t i = 0; i < n; ++i) Vx[i * n + i] = 1.0;
            r = n;
        } else {
            Ux.assign(m * m, 0.0);
            for (int64_t i = 0; i < m; ++i) Ux[i * m + i] = 1.0;
            Vx.assign(n * m, 0.0);       // V = M^T (n x m)
            for (int64_t j = 0; j < n; ++j)
                for (int64_t i = 0; i < m; ++i) Vx[i * n + j] = M[j * m + i];
            r = m;
        }
        pU = Ux.data();
        pV = Vx.data();
    }
    std::vector<double> QU(m * r), RU(r * r), QV(n * r), RV(r * r);
    householder_qr(m, r, pU, QU.data(), RU.data());
    householder_qr(n, r, pV, QV.data(), RV.data());
    // R = RU * RV^T
    std::vector<double> R(r * r, 0.0);
    for (int64_t j = 0; j < r; ++j)
        for (int64_t l = 0; l < r; ++l) {
            const double v = RV[l * r + j];       // RV^T(l, j) = RV(j, l)
            if (v == 0) continue;
            for (int64_t i = 0; i < r; ++i) R[j * r + i] += RU[l * r + i] * v;
        }
    std::vector<double> Us(r * r), Ss(r), Vs(r * r);
    jacobi_svd(r, R.data(), Us.data(), Ss.data(), Vs.data());
    const int64_t k = trunc_rank(r, Ss.data(), eps);
    Uo.assign(m * k, 0.0);
    Vo.assign(n * k, 0.0);
    So.assign(Ss.begin(), Ss.begin() + k);
    for (int64_t c = 0; c < k; ++c) {
        for (int64_t l = 0; l < r; ++l) {
            const double u = Us[c * r + l];
            const double v = Vs[c * r + l];
            if (u != 0) for (int64_t i = 0; i < m; ++i) Uo[c * m + i] += QU[l * m + i] * u;
            if (v != 0) for (int64_t i = 0; i < n; ++i) Vo[c * n + i] += QV[l * n + i] * v;
        }
    }
    return k;
}

// ---------------------------------------------------------------------------
// problem: coordinates and kernel
// ---------------------------------------------------------------------------

struct Problem {
    int kernel = HLRGEN_KERNEL_LOG;
    int64_t n = 0;
    double h = 0;                     // 1D cell width
    double ell = 0.1, sigma = 1.0;
    std::vector<double> X;            // internal order, 3 per point
    std::vector<int64_t> perm;        // internal -> external
    std::vector<double> diag;         // 1/r: diagonal entries (internal order)

    inline double dist(int64_t i, int64_t j) const {
        const double dx = X[3 * i] - X[3 * j];
        const double dy = X[3 * i + 1] - X[3 * j + 1];
        const double dz = X[3 * i + 2] - X[3 * j + 2];
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    }

    // a(i, j) for internal indices
    inline double eval(int64_t i, int64_t j) const {
        switch (kernel) {
        case HLRGEN_KERNEL_LOG:
            if (i == j) return std::log(h) - 1.5;     // (1/h^2) int int log|s-t| over a cell
            return std::log(std::fabs(X[3 * i] - X[3 * j]));
        case HLRGEN_KERNEL_LOG_GALERKIN: {
            // Galerkin entry -int_{cell a} int_{cell b} log|s-t| for piecewise
            // constants on a uniform grid (src/apps/log_kernel.cc:54-68)
            const int64_t a = perm[i], b = perm[j];
            double v;
            if (a == b) {
                v = h * h * std::log(h) - 1.5 * h * h;
            } else {
                const double d = h * (std::fabs(double(a - b)) - 1.0);
                const double t1 = d + h, t2 = d + 2 * h;
                v = -1.5 * h * h + 0.5 * t2 * t2 * std::log(t2) - t1 * t1 * std::log(t1);
                if (std::fabs(d) > 1e-8) v += 0.5 * d * d * std::log(d);
            }
            return -v;
        }
        case HLRGEN_KERNEL_MATERN: {
            // nu = 3/2: sigma^2 (1 + sqrt(3) r / l) exp(-sqrt(3) r / l)
            const double s = std::sqrt(3.0) * dist(i, j) / ell;
            return sigma * sigma * (1 + s) * std::exp(-s);
        }
        case HLRGEN_KERNEL_INVR:
        default:
            if (i == j) return diag[i];
            return 1.0 / dist(i, j);
        }
    }
};

// ---------------------------------------------------------------------------
// cluster tree: cardinality-balanced binary space partitioning
// ---------------------------------------------------------------------------

struct Cluster {
    int64_t first = 0, size = 0;
    double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    int32_t son[2] = {-1, -1};
    int32_t level = 0;
    bool leaf() const { return son[0] < 0; }
};

struct ClusterTree {
    std::vector<Cluster> nodes;
    std::vector<int32_t> leaves;     // in index order
};

void bbox(const Problem& P, const std::vector<int64_t>& idx, int64_t first, int64_t size,
          double* bmin, double* bmax, double pad) {
    for (int d = 0; d < 3; ++d) { bmin[d] = std::numeric_limits<double>::max(); bmax[d] = -bmin[d]; }
    for (int64_t t = first; t < first + size; ++t) {
        const double* x = &P.X[3 * idx[t]];
        for (int d = 0; d < 3; ++d) { bmin[d] = std::min(bmin[d], x[d]); bmax[d] = std::max(bmax[d], x[d]); }
    }
    for (int d = 0; d < 3; ++d) { bmin[d] -= pad; bmax[d] += pad; }
}

// idx holds external point ids; on return idx[t] is the external id of internal t
int32_t build_bsp(const Problem& P, std::vector<int64_t>& idx, int64_t first, int64_t size,
                  int64_t leafsize, int level, double pad, ClusterTree& T) {
    const int32_t id = int32_t(T.nodes.size());
    T.nodes.emplace_back();
    {
        Cluster& c = T.nodes.back();
        c.first = first; c.size = size; c.level = level;
        bbox(P, idx, first, size, c.bmin, c.bmax, pad);
    }
    if (size <= leafsize) { T.leaves.push_back(id); return id; }
    int axis = 0;
    {
        const Cluster& c = T.nodes[id];
        double ext = -1;
        for (int d = 0; d < 3; ++d)
            if (c.bmax[d] - c.bmin[d] > ext) { ext = c.bmax[d] - c.bmin[d]; axis = d; }
    }
    const int64_t half = size / 2;
    auto beg = idx.begin() + first;
    std::nth_element(beg, beg + half, beg + size, [&](int64_t a, int64_t b) {
        const double xa = P.X[3 * a + axis], xb = P.X[3 * b + axis];
        return (xa < xb) || (xa == xb && a < b);
    });
    const int32_t s0 = build_bsp(P, idx, first, half, leafsize, level + 1, pad, T);
    const int32_t s1 = build_bsp(P, idx, first + half, size - half, leafsize, level + 1, pad, T);
    T.nodes[id].son[0] = s0;
    T.nodes[id].son[1] = s1;
    return id;
}

// one-level tiling for TLR: root with all tiles of size ntile as sons
void build_tlr_tiles(const Problem& P, std::vector<int64_t>& idx, int64_t n, int64_t ntile,
                     double pad, ClusterTree& T) {
    // order points by a BSP down to the tile size, then take the BSP leaves
    ClusterTree bsp;
    build_bsp(P, idx, 0, n, ntile, 0, pad, bsp);
    T.nodes.clear(); T.leaves.clear();
    T.nodes.emplace_back();
    T.nodes[0].first = 0; T.nodes[0].size = n;
    bbox(P, idx, 0, n, T.nodes[0].bmin, T.nodes[0].bmax, pad);
    for (int32_t l : bsp.leaves) {
        Cluster c = bsp.nodes[l];
        c.level = 1; c.son[0] = c.son[1] = -1;
        T.leaves.push_back(int32_t(T.nodes.size()));
        T.nodes.push_back(c);
    }
}

double diameter(const Cluster& c) {
    double s = 0;
    for (int d = 0; d < 3; ++d) s += (c.bmax[d] - c.bmin[d]) * (c.bmax[d] - c.bmin[d]);
    return std::sqrt(s);
}

double distance(const Cluster& a, const Cluster& b) {
    double s = 0;
    for (int d = 0; d < 3; ++d) {
        double g = 0;
        if (a.bmax[d] < b.bmin[d]) g = b.bmin[d] - a.bmax[d];
        else if (b.bmax[d] < a.bmin[d]) g = a.bmin[d] - b.bmax[d];
        s += g * g;
    }
    return std::sqrt(s);
}

// standard geometric admissibility: min(diam) <= eta * dist
bool std_admissible(const Cluster& a, const Cluster& b, double eta) {
    const double dist = distance(a, b);
    if (dist <= 0) return false;
    return std::min(diameter(a), diameter(b)) <= eta * dist;
}

struct LeafRef {
    int32_t rc, cc;
    bool lowrank;
};

// leaves in hlr::mul_vec visiting order: row-major over block sons, recursive
void build_blocks(const ClusterTree& T, int32_t r, int32_t c, double eta, std::vector<LeafRef>& out) {
    const Cluster& R = T.nodes[r];
    const Cluster& C = T.nodes[c];
    if (std_admissible(R, C, eta)) { out.push_back({r, c, true}); return; }
    if (R.leaf() || C.leaf()) { out.push_back({r, c, false}); return; }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            build_blocks(T, R.son[i], C.son[j], eta, out);
}

// ---------------------------------------------------------------------------
// ACA with maximal-pivot search (approx/aca.hh:117-223, 325-434)
// ---------------------------------------------------------------------------

struct AcaResult {
    std::vector<double> U, V;   // nr x k, nc x k
    int64_t k = 0;
};

void aca_block(const Problem& P, int64_t r0, int64_t nr, int64_t c0, int64_t nc, double eps,
               int64_t kmax, AcaResult& res) {
    const double zero_val = std::numeric_limits<double>::epsilon() * std::numeric_limits<double>::epsilon();
    const int64_t max_rank = std::min<int64_t>(kmax, std::min(nr, nc));
    std::vector<char> used_cols(nc, 0);
    std::vector<double>& U = res.U;
    std::vector<double>& V = res.V;
    U.clear(); V.clear();
    std::vector<double> col(nr), row(nc);
    int64_t next_col = 0, k = 0;
    double normM = 0;
    for (int64_t it = 0; it < max_rank; ++it) {
        const int64_t pc = next_col;
        used_cols[pc] = 1;
        for (int64_t i = 0; i < nr; ++i) col[i] = P.eval(r0 + i, c0 + pc);
        for (int64_t l = 0; l < k; ++l) {
            const double v = V[l * nc + pc];
            const double* u = &U[l * nr];
            for (int64_t i = 0; i < nr; ++i) col[i] -= v * u[i];
        }
        int64_t pr = 0;
        double amax = -1;
        for (int64_t i = 0; i < nr; ++i)
            if (std::fabs(col[i]) > amax) { amax = std::fabs(col[i]); pr = i; }
        const double maxval = col[pr];
        if (std::fabs(maxval) <= zero_val) break;
        for (int64_t i = 0; i < nr; ++i) col[i] /= maxval;
        for (int64_t j = 0; j < nc; ++j) row[j] = P.eval(r0 + pr, c0 + j);
        for (int64_t l = 0; l < k; ++l) {
            const double u = U[l * nr + pr];
            const double* v = &V[l * nc];
            for (int64_t j = 0; j < nc; ++j) row[j] -= u * v[j];
        }
        // next column: largest entry of the new row among unused columns
        int64_t mj = -1;
        double mv = 0;
        for (int64_t j = 0; j < nc; ++j)
            if (!used_cols[j] && std::fabs(row[j]) > mv) { mv = std::fabs(row[j]); mj = j; }
        if (mj < 0)
            for (int64_t j = 0; j < nc; ++j)
                if (!used_cols[j]) { mj = j; break; }
        next_col = mj;
        const double norm_i = nrm2(nr, col.data()) * nrm2(nc, row.data());
        const bool last = (norm_i < eps * normM);
        if (!last) {
            double upd = norm_i * norm_i;
            for (int64_t l = 0; l < k; ++l)
                upd += 2 * dot(nr, &U[l * nr], col.data()) * dot(nc, &V[l * nc], row.data());
            normM = std::sqrt(normM * normM + std::fabs(upd));
        }
        U.insert(U.end(), col.begin(), col.end());
        V.insert(V.end(), row.begin(), row.end());
        ++k;
        if (last || next_col < 0) break;
    }
    res.k = k;
}

// nearest-neighbour diagonal for 1/r: mean of the nn entries 1/|x_i - x_j|
// over the nn nearest other points, searched in the dense (near-field)
// leaves of the block tree (every point within a leaf diameter is there).
void nn_diagonal(Problem& P, const ClusterTree& T, const std::vector<LeafRef>& leaves, int nn) {
    const int64_t n = P.n;
    P.diag.assign(n, 0.0);
    // near-field column clusters per row leaf cluster
    std::vector<std::vector<int32_t>> near(T.nodes.size());
    for (const auto& L : leaves)
        if (!L.lowrank) near[L.rc].push_back(L.cc);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t li = 0; li < int64_t(T.leaves.size()); ++li) {
        const int32_t rc = T.leaves[li];
        const Cluster& R = T.nodes[rc];
        std::vector<double> best;
        for (int64_t i = R.first; i < R.first + R.size; ++i) {
            best.clear();
            for (int32_t cc : near[rc]) {
                const Cluster& C = T.nodes[cc];
                for (int64_t j = C.first; j < C.first + C.size; ++j) {
                    if (j == i) continue;
                    const double d = P.dist(i, j);
                    if (int(best.size()) < nn) {
                        best.push_back(d);
                        std::push_heap(best.begin(), best.end());
                    } else if (d < best.front()) {
                        std::pop_heap(best.begin(), best.end());
                        best.back() = d;
                        std::push_heap(best.begin(), best.end());
                    }
                }
            }
            double s = 0;
            for (double d : best) s += 1.0 / d;
            P.diag[i] = best.empty() ? 0.0 : s / double(best.size());
        }
    }
}

} // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

struct hlrgen_matrix {
    hlrgen_config cfg{};
    Problem P;
    std::vector<hmvm_block> blocks;
    std::vector<double> arena;
    double seconds = 0;
};

extern "C" void hlrgen_default_config(hlrgen_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->kernel = HLRGEN_KERNEL_LOG;
    c->geometry = HLRGEN_GEOM_LINE;
    c->n = 4096;
    c->seed = 1;
    c->ntile = 64;
    c->format = HLRGEN_FORMAT_H;
    c->eta = 2.0;
    c->eps = 1e-6;
    c->matern_ell = 0.1;
    c->matern_sigma = 1.0;
    c->nn_diag = 8;
    c->lrformat = HLRGEN_LR_UV;
    c->nthreads = 0;
    c->tlr_std_adm = 0;
    c->verbose = 0;
}

static void make_points(const hlrgen_config& cfg, Problem& P) {
    const int64_t n = cfg.n;
    P.X.assign(3 * n, 0.0);
    std::mt19937_64 gen{cfg.seed};
    switch (cfg.geometry) {
    case HLRGEN_GEOM_LINE: {
        std::uniform_real_distribution<double> d{0.0, 1.0};
        for (int64_t i = 0; i < n; ++i) P.X[3 * i] = d(gen);
        break;
    }
    case HLRGEN_GEOM_GRID1D:
        for (int64_t i = 0; i < n; ++i) P.X[3 * i] = P.h * (double(i) + 0.5);
        break;
    case HLRGEN_GEOM_CUBE: {
        std::uniform_real_distribution<double> d{0.0, 1.0};
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) P.X[3 * i + k] = d(gen);
        break;
    }
    case HLRGEN_GEOM_SPHERE:
    default: {
        // normalised Gaussian triples are uniform on the sphere
        std::normal_distribution<double> d{0.0, 1.0};
        for (int64_t i = 0; i < n; ++i) {
            double v[3], s;
            do {
                for (int k = 0; k < 3; ++k) v[k] = d(gen);
                s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            } while (s < 1e-12);
            for (int k = 0; k < 3; ++k) P.X[3 * i + k] = v[k] / s;
        }
        break;
    }
    }
}

extern "C" int hlrgen_build(const hlrgen_config* cfgp, hlrgen_matrix** out) {
    if (!cfgp || !out) return HMVM_ERR_INVALID_ARG;
    const hlrgen_config cfg = *cfgp;
    if (cfg.n < 1 || cfg.ntile < 1 || cfg.eps <= 0 || cfg.eta <= 0) return HMVM_ERR_INVALID_ARG;
    if (cfg.kernel == HLRGEN_KERNEL_INVR && cfg.nn_diag < 1) return HMVM_ERR_INVALID_ARG;
    auto t0 = std::chrono::steady_clock::now();
    if (cfg.nthreads > 0) omp_set_num_threads(cfg.nthreads);
    std::unique_ptr<hlrgen_matrix> M;
    try {
        M.reset(new hlrgen_matrix);
    } catch (...) { return HMVM_ERR_ALLOC; }
    M->cfg = cfg;
    Problem& P = M->P;
    P.kernel = cfg.kernel;
    P.n = cfg.n;
    P.h = 1.0 / double(cfg.n);
    P.ell = cfg.matern_ell;
    P.sigma = cfg.matern_sigma;
    try {
        make_points(cfg, P);
        // clustering on external ids, then permute to internal order
        std::vector<int64_t> idx(cfg.n);
        std::iota(idx.begin(), idx.end(), 0);
        ClusterTree T;
        const double pad = (cfg.geometry == HLRGEN_GEOM_GRID1D) ? 0.5 * P.h : 0.0;
        if (cfg.format == HLRGEN_FORMAT_TLR) build_tlr_tiles(P, idx, cfg.n, cfg.ntile, pad, T);
        else build_bsp(P, idx, 0, cfg.n, cfg.ntile, 0, pad, T);
        std::vector<double> Xi(3 * cfg.n);
        for (int64_t t = 0; t < cfg.n; ++t)
            for (int k = 0; k < 3; ++k) Xi[3 * t + k] = P.X[3 * idx[t] + k];
        P.X.swap(Xi);
        P.perm = idx;
        // block tree
        std::vector<LeafRef> leaves;
        if (cfg.format == HLRGEN_FORMAT_TLR) {
            for (int32_t r : T.leaves)
                for (int32_t c : T.leaves) {
                    bool lr = (r != c);
                    if (cfg.tlr_std_adm) lr = std_admissible(T.nodes[r], T.nodes[c], cfg.eta);
                    leaves.push_back({r, c, lr});
                }
        } else {
            build_blocks(T, 0, 0, cfg.eta, leaves);
        }
        if (cfg.kernel == HLRGEN_KERNEL_INVR) nn_diagonal(P, T, leaves, cfg.nn_diag);

        // leaf build, largest first
        const int64_t nl = int64_t(leaves.size());
        struct LeafData { std::vector<double> a, b, s; int64_t k = 0; };
        std::vector<LeafData> data(nl);
        std::vector<int64_t> order(nl);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
            const auto& A = leaves[x];
            const auto& B = leaves[y];
            const double ca = double(T.nodes[A.rc].size) * T.nodes[A.cc].size * (A.lowrank ? 0.3 : 1.0);
            const double cb = double(T.nodes[B.rc].size) * T.nodes[B.cc].size * (B.lowrank ? 0.3 : 1.0);
            return ca > cb;
        });
        std::atomic<int> failed{0};
#pragma omp parallel
        {
            AcaResult aca;
            std::vector<double> Uo, So, Vo;
#pragma omp for schedule(dynamic, 1)
            for (int64_t oi = 0; oi < nl; ++oi) {
                if (failed.load()) continue;
                try {
                    const int64_t li = order[oi];
                    const LeafRef& L = leaves[li];
                    const Cluster& R = T.nodes[L.rc];
                    const Cluster& C = T.nodes[L.cc];
                    const int64_t m = R.size, nn = C.size;
                    LeafData& d = data[li];
                    if (!L.lowrank) {
                        d.a.resize(m * nn);
                        for (int64_t j = 0; j < nn; ++j)
                            for (int64_t i = 0; i < m; ++i) d.a[j * m + i] = P.eval(R.first + i, C.first + j);
                    } else if (cfg.skip_lowrank) {
                        d.k = 0;
                    } else {
                        aca_block(P, R.first, m, C.first, nn, cfg.eps, std::min(m, nn), aca);
                        const int64_t k = truncate_lr(m, nn, aca.k, aca.U.data(), aca.V.data(), cfg.eps, Uo, So, Vo);
                        d.k = k;
                        if (cfg.lrformat == HLRGEN_LR_UV) {
                            // fold S into the factor with fewer rows (approx/svd.hh:206-207)
                            std::vector<double>& F = (m < nn) ? Uo : Vo;
                            const int64_t rows = (m < nn) ? m : nn;
                            for (int64_t l = 0; l < k; ++l)
                                for (int64_t i = 0; i < rows; ++i) F[l * rows + i] *= So[l];
                        }
                        d.a = Uo; d.b = Vo; d.s = So;
                    }
                } catch (...) { failed = 1; }
            }
        }
        if (failed) return HMVM_ERR_ALLOC;

        // pack: blocks in visiting order, factors 32-byte aligned
        auto pad4 = [](int64_t v) { return (v + 3) & ~int64_t(3); };
        M->blocks.resize(nl);
        int64_t pos = 0;
        for (int64_t li = 0; li < nl; ++li) {
            const LeafRef& L = leaves[li];
            const Cluster& R = T.nodes[L.rc];
            const Cluster& C = T.nodes[L.cc];
            hmvm_block& b = M->blocks[li];
            b.row_ofs = R.first; b.col_ofs = C.first;
            b.nrows = int32_t(R.size); b.ncols = int32_t(C.size);
            b.off_s = -1; b.off_b = -1;
            if (!L.lowrank) {
                b.kind = HMVM_BLOCK_DENSE; b.rank = 0;
                b.off_a = pos; pos += pad4(int64_t(R.size) * C.size);
            } else {
                const int64_t k = data[li].k;
                b.kind = (cfg.lrformat == HLRGEN_LR_USV) ? HMVM_BLOCK_LOWRANK_SV : HMVM_BLOCK_LOWRANK;
                b.rank = int32_t(k);
                b.off_a = pos; pos += pad4(R.size * k);
                b.off_b = pos; pos += pad4(C.size * k);
                b.off_s = pos; pos += pad4(k);
            }
        }
        M->arena.assign(pos, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t li = 0; li < nl; ++li) {
            const hmvm_block& b = M->blocks[li];
            LeafData& d = data[li];
            std::copy(d.a.begin(), d.a.end(), M->arena.begin() + b.off_a);
            if (b.kind != HMVM_BLOCK_DENSE) {
                std::copy(d.b.begin(), d.b.end(), M->arena.begin() + b.off_b);
                std::copy(d.s.begin(), d.s.end(), M->arena.begin() + b.off_s);
            }
            std::vector<double>().swap(d.a);
            std::vector<double>().swap(d.b);
            std::vector<double>().swap(d.s);
        }
    } catch (const std::bad_alloc&) {
        return HMVM_ERR_ALLOC;
    } catch (...) {
        return HMVM_ERR_INVALID_ARG;
    }
    M->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (cfg.verbose)
        std::fprintf(stderr, "hlrgen: n=%lld blocks=%zu arena=%.1f MB in %.2f s\n", (long long)cfg.n,
                     M->blocks.size(), M->arena.size() * 8e-6, M->seconds);
    *out = M.release();
    return HMVM_OK;
}

extern "C" void hlrgen_free(hlrgen_matrix* m) { delete m; }
extern "C" int64_t hlrgen_n(const hlrgen_matrix* m) { return m->P.n; }
extern "C" int64_t hlrgen_nblocks(const hlrgen_matrix* m) { return int64_t(m->blocks.size()); }
extern "C" const hmvm_block* hlrgen_blocks(const hlrgen_matrix* m) { return m->blocks.data(); }
extern "C" int64_t hlrgen_arena_len(const hlrgen_matrix* m) { return int64_t(m->arena.size()); }
extern "C" const double* hlrgen_arena(const hlrgen_matrix* m) { return m->arena.data(); }
extern "C" const int64_t* hlrgen_perm(const hlrgen_matrix* m) { return m->P.perm.data(); }
extern "C" const double* hlrgen_coords(const hlrgen_matrix* m) { return m->P.X.data(); }
extern "C" double hlrgen_build_seconds(const hlrgen_matrix* m) { return m->seconds; }

extern "C" int hlrgen_eval(const hlrgen_matrix* m, int64_t nr, const int64_t* rows, int64_t nc,
                           const int64_t* cols, double* out) {
    if (!m || !rows || !cols || !out) return HMVM_ERR_INVALID_ARG;
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            if (rows[i] < 0 || rows[i] >= m->P.n || cols[j] < 0 || cols[j] >= m->P.n) return HMVM_ERR_INVALID_ARG;
            out[j * nr + i] = m->P.eval(rows[i], cols[j]);
        }
    return HMVM_OK;
}

extern "C" void hlrgen_random_vector(uint64_t seed, int64_t n, double lo, double hi, int gaussian, double* out) {
    std::mt19937_64 gen{seed};
    if (gaussian) {
        std::normal_distribution<double> d{0.0, 1.0};
        for (int64_t i = 0; i < n; ++i) out[i] = d(gen);
    } else {
        std::uniform_real_distribution<double> d{lo, hi};
        for (int64_t i = 0; i < n; ++i) out[i] = d(gen);
    }
}

extern "C" int hlrgen_qr(int64_t m, int64_t k, const double* A, double* Q, double* R) {
    if (m < k || k < 0 || !A || !Q || !R) return HMVM_ERR_INVALID_ARG;
    householder_qr(m, k, A, Q, R);
    return HMVM_OK;
}

extern "C" int hlrgen_svd(int64_t k, const double* A, double* U, double* S, double* V) {
    if (k < 0 || !A || !U || !S || !V) return HMVM_ERR_INVALID_ARG;
    jacobi_svd(k, A, U, S, V);
    return HMVM_OK;
}

extern "C" int hlrgen_truncate(int64_t m, int64_t n, int64_t r, const double* U, const double* V, double eps,
                               double* Uo, double* So, double* Vo, int64_t* k_out) {
    if (!U || !V || !Uo || !So || !Vo || !k_out || r < 0) return HMVM_ERR_INVALID_ARG;
    std::vector<double> u, s, v;
    const int64_t k = truncate_lr(m, n, r, U, V, eps, u, s, v);
    std::copy(u.begin(), u.end(), Uo);
    std::copy(s.begin(), s.end(), So);
    std::copy(v.begin(), v.end(), Vo);
    *k_out = k;
    return HMVM_OK;
}

extern "C" int hlrgen_aca(const hlrgen_matrix* m, int64_t nr, const int64_t* rows, int64_t nc,
                          const int64_t* cols, double eps, int64_t kmax, double* U, double* V, int64_t* k_out) {
    if (!m || !rows || !cols || !U || !V || !k_out || nr < 1 || nc < 1) return HMVM_ERR_INVALID_ARG;
    // rows/cols must be contiguous internal ranges
    for (int64_t i = 1; i < nr; ++i) if (rows[i] != rows[0] + i) return HMVM_ERR_INVALID_ARG;
    for (int64_t j = 1; j < nc; ++j) if (cols[j] != cols[0] + j) return HMVM_ERR_INVALID_ARG;
    AcaResult res;
    aca_block(m->P, rows[0], nr, cols[0], nc, eps, kmax, res);
    std::copy(res.U.begin(), res.U.end(), U);
    std::copy(res.V.begin(), res.V.end(), V);
    *k_out = res.k;
    return HMVM_OK;
}
