// uniform.hip -- uniform-H (shared cluster bases) and H^2 (nested bases)
// products, arith/uniform.hh:114-145 and arith/h2.hh:18-44.
//
// A uniform-H / H^2 matrix is run as a chain of ordinary device plans, each
// a flattened block matrix over the coefficient vectors of the cluster
// bases (entries of the clusters of one tree depth are contiguous):
//   B    xc  = B x          explicit column bases: a leaf (I_k, V_c) per
//                           cluster, rows = its coefficient entries
//   Tc_d xc[d-1] += E^T xc[d]  column tree transfers, deepest depth first
//                           (leaves E_s at rows of son s, columns of its
//                           parent; used transposed)
//   C    yr += alpha C xc   coupling matrices S_b as dense leaves
//   Tr_d yr[d] += E yr[d-1] row tree transfers, root side first
//   D    y  += D yr         explicit row bases: a leaf (U_r, I_k) per cluster
//   N    y  += alpha N x    dense near field
// The transposed product runs the same plans transposed in reverse order.
// Every plan streams its factors once per product with the forward or
// transposed kernels of hmvm.hip; the identity factors cost k^2 values per
// basis.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "hmvm.h"

namespace {

__global__ __launch_bounds__(256) void k_zero(double* y, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

void zero(double* y, int64_t n, hipStream_t st) {
    if (n > 0) k_zero<<<unsigned((n + 255) / 256), 256, 0, st>>>(y, n);
}

struct Level {
    hmvm_plan* p = nullptr;
    int64_t son0 = 0, par0 = 0;    // first coefficient entry of the son and parent depths
};

struct Side {
    std::vector<int> depth;        // per cluster
    std::vector<int64_t> entry;    // first coefficient entry per cluster
    std::vector<int64_t> lvl0;     // first entry per depth (size maxdepth + 2)
    int64_t K = 0;                 // coefficient entries
    int maxdepth = 0;
};

// depths and coefficient numbering (depth 0 first, clusters in order)
int number_side(const hmvm_cluster* cb, int64_t n, Side& s) {
    s.depth.assign(size_t(n), 0);
    for (int64_t i = 0; i < n; ++i) {
        int d = 0;
        int64_t p = cb[i].parent;
        while (p >= 0) {
            if (p >= n || ++d > n) return HMVM_ERR_BAD_BLOCK;     // bad parent or cycle
            p = cb[p].parent;
        }
        s.depth[size_t(i)] = d;
        s.maxdepth = std::max(s.maxdepth, d);
    }
    s.entry.assign(size_t(n), 0);
    s.lvl0.assign(size_t(s.maxdepth) + 2, 0);
    int64_t e = 0;
    for (int d = 0; d <= s.maxdepth; ++d) {
        s.lvl0[size_t(d)] = e;
        for (int64_t i = 0; i < n; ++i)
            if (s.depth[size_t(i)] == d) { s.entry[size_t(i)] = e; e += cb[i].rank; }
    }
    s.lvl0[size_t(s.maxdepth) + 1] = e;
    s.K = e;
    return HMVM_OK;
}

int check_side(const hmvm_cluster* cb, int64_t n, int64_t dim, int64_t arena_len) {
    auto inside = [](int64_t off, int64_t len, int64_t total) {
        return off >= 0 && len >= 0 && len <= total && off <= total - len;
    };
    for (int64_t i = 0; i < n; ++i) {
        const hmvm_cluster& c = cb[i];
        if (c.size < 0 || c.rank < 0 || !inside(c.ofs, c.size, dim)) return HMVM_ERR_BAD_BLOCK;
        if (c.parent < -1 || c.parent >= n) return HMVM_ERR_BAD_BLOCK;
        if (c.rank == 0) continue;
        if (c.rank > c.size) return HMVM_ERR_BAD_BLOCK;
        if (c.off_basis >= 0 && !inside(c.off_basis, int64_t(c.size) * c.rank, arena_len)) return HMVM_ERR_BAD_BLOCK;
        if (c.off_s >= 0 && !inside(c.off_s, c.rank, arena_len)) return HMVM_ERR_BAD_BLOCK;
        if (c.off_transfer >= 0) {
            if (c.parent < 0) return HMVM_ERR_BAD_BLOCK;
            const hmvm_cluster& p = cb[c.parent];
            if (!inside(c.off_transfer, int64_t(c.rank) * p.rank, arena_len)) return HMVM_ERR_BAD_BLOCK;
            if (c.ofs < p.ofs || c.ofs + c.size > p.ofs + p.size) return HMVM_ERR_BAD_BLOCK;
        }
    }
    // a cluster with a basis gets its coefficients from an explicit basis or
    // from at least one son with a transfer
    std::vector<char> fed(size_t(n), 0);
    for (int64_t i = 0; i < n; ++i)
        if (cb[i].rank > 0 && cb[i].off_transfer >= 0) fed[size_t(cb[i].parent)] = 1;
    for (int64_t i = 0; i < n; ++i)
        if (cb[i].rank > 0 && cb[i].off_basis < 0 && !fed[size_t(i)]) return HMVM_ERR_BAD_BLOCK;
    return HMVM_OK;
}

} // namespace

struct hmvm_uplan {
    int device = -1;
    int64_t nrows = 0, ncols = 0;
    Side row, col;
    hmvm_plan* B = nullptr;
    hmvm_plan* D = nullptr;
    hmvm_plan* C = nullptr;
    hmvm_plan* N = nullptr;
    std::vector<Level> tcol;       // index d - 1 for sons at depth d
    std::vector<Level> trow;
    int64_t ws_doubles = 0;        // workspace of the largest sub-plan call
    hmvm_uplan_info info{};

    ~hmvm_uplan() {
        for (hmvm_plan* p : {B, D, C, N}) hmvm_plan_destroy(p);
        for (auto& l : tcol) hmvm_plan_destroy(l.p);
        for (auto& l : trow) hmvm_plan_destroy(l.p);
    }
};

namespace {

// a sub-plan over blocks `bl` (empty: none)
int make_plan(const std::vector<hmvm_block>& bl, int64_t nr, int64_t nc, const std::vector<double>& arena,
              const hmvm_plan_opts& o, int device, hmvm_plan** out, int64_t& ws) {
    *out = nullptr;
    if (bl.empty() || nr == 0 || nc == 0) return HMVM_OK;
    int rc = hmvm_plan_create(bl.data(), int64_t(bl.size()), nr, nc, arena.data(), int64_t(arena.size()), &o, device,
                              out);
    if (rc != HMVM_OK) return rc;
    for (int op : {HMVM_OP_NORMAL, HMVM_OP_TRANSPOSED}) {
        int64_t b = 0;
        if ((rc = hmvm_workspace_size(*out, op, 1, &b)) != HMVM_OK) return rc;
        ws = std::max(ws, (b + 7) / 8);
    }
    return HMVM_OK;
}

hmvm_block leaf(int64_t r0, int64_t c0, int64_t m, int64_t n, int kind, int64_t k, int64_t a, int64_t b, int64_t s) {
    hmvm_block x;
    std::memset(&x, 0, sizeof(x));
    x.row_ofs = r0; x.col_ofs = c0; x.nrows = int32_t(m); x.ncols = int32_t(n);
    x.rank = int32_t(k); x.kind = kind; x.off_a = a; x.off_b = b; x.off_s = s;
    return x;
}

int build(hmvm_uplan& U, const hmvm_cluster* rcb, int64_t nr, const hmvm_cluster* ccb, int64_t nc,
          const hmvm_ublock* blocks, int64_t nb, const double* arena, int64_t alen, const hmvm_plan_opts& o,
          int device) {
    int rc;
    if ((rc = check_side(rcb, nr, U.nrows, alen)) != HMVM_OK) return rc;
    if ((rc = check_side(ccb, nc, U.ncols, alen)) != HMVM_OK) return rc;
    if ((rc = number_side(rcb, nr, U.row)) != HMVM_OK) return rc;
    if ((rc = number_side(ccb, nc, U.col)) != HMVM_OK) return rc;
    auto inside = [](int64_t off, int64_t len, int64_t total) {
        return off >= 0 && len >= 0 && len <= total && off <= total - len;
    };
    // the plans' arena: the caller's values, then one identity per rank
    std::vector<double> ar(arena, arena + alen);
    std::map<int64_t, int64_t> ident;
    auto identity = [&](int64_t k) {
        auto it = ident.find(k);
        if (it != ident.end()) return it->second;
        const int64_t off = int64_t(ar.size());
        ar.resize(size_t(off + k * k), 0.0);
        for (int64_t i = 0; i < k; ++i) ar[size_t(off + i * k + i)] = 1.0;
        ident[k] = off;
        return off;
    };
    hmvm_uplan_info& I = U.info;
    I.nrows = U.nrows; I.ncols = U.ncols; I.nrow_clusters = nr; I.ncol_clusters = nc;
    std::vector<hmvm_block> bB, bD, bC, bN;
    for (int64_t j = 0; j < nc; ++j) {
        const hmvm_cluster& c = ccb[j];
        if (c.rank == 0 || c.off_basis < 0) continue;
        bB.push_back(leaf(U.col.entry[size_t(j)], c.ofs, c.rank, c.size, HMVM_BLOCK_LOWRANK, c.rank,
                          identity(c.rank), c.off_basis, c.off_s));
        I.basis_bytes += 8 * int64_t(c.size) * c.rank;
    }
    for (int64_t i = 0; i < nr; ++i) {
        const hmvm_cluster& c = rcb[i];
        if (c.rank == 0 || c.off_basis < 0) continue;
        bD.push_back(leaf(c.ofs, U.row.entry[size_t(i)], c.size, c.rank, HMVM_BLOCK_LOWRANK, c.rank, c.off_basis,
                          identity(c.rank), c.off_s));
        I.basis_bytes += 8 * int64_t(c.size) * c.rank;
    }
    for (int64_t b = 0; b < nb; ++b) {
        const hmvm_ublock& x = blocks[b];
        if (x.kind == HMVM_BLOCK_DENSE) {
            if (!inside(x.row_ofs, x.nrows, U.nrows) || !inside(x.col_ofs, x.ncols, U.ncols)) return HMVM_ERR_BAD_BLOCK;
            if (!inside(x.off, int64_t(x.nrows) * x.ncols, alen)) return HMVM_ERR_BAD_BLOCK;
            bN.push_back(leaf(x.row_ofs, x.col_ofs, x.nrows, x.ncols, HMVM_BLOCK_DENSE, 0, x.off, -1, -1));
            ++I.ndense;
            I.dense_bytes += 8 * int64_t(x.nrows) * x.ncols;
        } else if (x.kind == HMVM_BLOCK_COUPLING) {
            if (x.rowcb < 0 || x.rowcb >= nr || x.colcb < 0 || x.colcb >= nc) return HMVM_ERR_BAD_BLOCK;
            const hmvm_cluster& r = rcb[x.rowcb];
            const hmvm_cluster& c = ccb[x.colcb];
            if (x.row_ofs != r.ofs || x.nrows != r.size || x.col_ofs != c.ofs || x.ncols != c.size)
                return HMVM_ERR_BAD_BLOCK;
            ++I.ncoupling;
            if (r.rank == 0 || c.rank == 0) continue;
            if (!inside(x.off, int64_t(r.rank) * c.rank, alen)) return HMVM_ERR_BAD_BLOCK;
            bC.push_back(leaf(U.row.entry[size_t(x.rowcb)], U.col.entry[size_t(x.colcb)], r.rank, c.rank,
                              HMVM_BLOCK_DENSE, 0, x.off, -1, -1));
            I.coupling_bytes += 8 * int64_t(r.rank) * c.rank;
        } else {
            return HMVM_ERR_BAD_BLOCK;
        }
    }
    // coupling and transfer blocks: the dense-leaf rule of the options
    hmvm_plan_opts od = o;
    U.ws_doubles = 0;
    if ((rc = make_plan(bB, U.col.K, U.ncols, ar, o, device, &U.B, U.ws_doubles)) != HMVM_OK) return rc;
    if ((rc = make_plan(bD, U.nrows, U.row.K, ar, o, device, &U.D, U.ws_doubles)) != HMVM_OK) return rc;
    if ((rc = make_plan(bC, U.row.K, U.col.K, ar, od, device, &U.C, U.ws_doubles)) != HMVM_OK) return rc;
    if ((rc = make_plan(bN, U.nrows, U.ncols, ar, od, device, &U.N, U.ws_doubles)) != HMVM_OK) return rc;
    // transfers, one plan per son depth: leaf E_s (k_s x k_parent) at rows of
    // the son's entries (within its depth), columns of the parent's
    auto transfers = [&](const hmvm_cluster* cb, int64_t n, const Side& s, std::vector<Level>& out) {
        out.assign(size_t(s.maxdepth), Level{});
        for (int d = 1; d <= s.maxdepth; ++d) {
            std::vector<hmvm_block> bl;
            for (int64_t i = 0; i < n; ++i) {
                const hmvm_cluster& c = cb[i];
                if (s.depth[size_t(i)] != d || c.rank == 0 || c.off_transfer < 0 || cb[c.parent].rank == 0) continue;
                bl.push_back(leaf(s.entry[size_t(i)] - s.lvl0[size_t(d)], s.entry[size_t(c.parent)] - s.lvl0[size_t(d) - 1],
                                  c.rank, cb[c.parent].rank, HMVM_BLOCK_DENSE, 0, c.off_transfer, -1, -1));
                I.transfer_bytes += 8 * int64_t(c.rank) * cb[c.parent].rank;
            }
            Level& L = out[size_t(d) - 1];
            L.son0 = s.lvl0[size_t(d)];
            L.par0 = s.lvl0[size_t(d) - 1];
            const int64_t nson = s.lvl0[size_t(d) + 1] - s.lvl0[size_t(d)];
            const int64_t npar = s.lvl0[size_t(d)] - s.lvl0[size_t(d) - 1];
            int r2 = make_plan(bl, nson, npar, ar, od, device, &L.p, U.ws_doubles);
            if (r2 != HMVM_OK) return r2;
            if (L.p) { I.nested = 1; ++I.levels; }
        }
        return int(HMVM_OK);
    };
    if ((rc = transfers(ccb, nc, U.col, U.tcol)) != HMVM_OK) return rc;
    if ((rc = transfers(rcb, nr, U.row, U.trow)) != HMVM_OK) return rc;
    // flops and data size of arith/uniform.hh / arith/h2.hh: both
    // transforms, couplings, dense leaves; vectors in and out
    double fl = 0;
    for (const auto* cbv : {rcb, ccb}) {
        const int64_t n = (cbv == rcb) ? nr : nc;
        for (int64_t i = 0; i < n; ++i) {
            const hmvm_cluster& c = cbv[i];
            if (c.rank == 0) continue;
            if (c.off_basis >= 0) fl += 2.0 * double(c.size) * c.rank;
            if (c.off_transfer >= 0) fl += 2.0 * double(c.rank) * cbv[c.parent].rank;
        }
    }
    fl += 2.0 * double(I.coupling_bytes / 8) + 2.0 * double(I.dense_bytes / 8);
    I.flops = fl;
    I.datasize = double(I.basis_bytes + I.transfer_bytes + I.coupling_bytes + I.dense_bytes) +
                 8.0 * double(U.nrows + U.ncols);
    return HMVM_OK;
}

inline int sub(const hmvm_plan* p, double a, int op, const double* x, double* y, double* ws, int64_t wsd,
               hipStream_t st) {
    if (!p) return HMVM_OK;
    return hmvm_apply_ws(p, a, op, x, 0, y, 0, 1, HMVM_F64, ws, wsd * 8, st);
}

} // namespace

extern "C" {

int hmvm_uplan_create(const hmvm_cluster* rowcb, int64_t nrowcb, const hmvm_cluster* colcb, int64_t ncolcb,
                      const hmvm_ublock* blocks, int64_t nblocks, int64_t nrows, int64_t ncols, const double* arena,
                      int64_t arena_len, const hmvm_plan_opts* opts, int device, hmvm_uplan** out) {
    if (!out) return HMVM_ERR_INVALID_ARG;
    *out = nullptr;
    if (nrowcb < 0 || ncolcb < 0 || nblocks < 0 || nrows < 0 || ncols < 0 || arena_len < 0) return HMVM_ERR_INVALID_ARG;
    if ((nrowcb > 0 && !rowcb) || (ncolcb > 0 && !colcb) || (nblocks > 0 && !blocks) || (arena_len > 0 && !arena))
        return HMVM_ERR_INVALID_ARG;
    hmvm_plan_opts o;
    if (opts) o = *opts;
    else { std::memset(&o, 0, sizeof(o)); o.precision = HMVM_PREC_FP64; o.check_finite = 1; }
    try {
        std::unique_ptr<hmvm_uplan> U(new hmvm_uplan);
        U->device = device;
        U->nrows = nrows;
        U->ncols = ncols;
        const int rc = build(*U, rowcb, nrowcb, colcb, ncolcb, blocks, nblocks, arena, arena_len, o, device);
        if (rc != HMVM_OK) return rc;
        *out = U.release();
    } catch (const std::bad_alloc&) {
        return HMVM_ERR_ALLOC;
    }
    return HMVM_OK;
}

int hmvm_uplan_destroy(hmvm_uplan* p) {
    delete p;
    return HMVM_OK;
}

int hmvm_uplan_get_info(const hmvm_uplan* p, hmvm_uplan_info* info) {
    if (!p || !info) return HMVM_ERR_INVALID_ARG;
    *info = p->info;
    return HMVM_OK;
}

int hmvm_uplan_apply(const hmvm_uplan* p, double alpha, int op, const double* x, double* y, void* stream) {
    if (!p || op < HMVM_OP_NORMAL || op > HMVM_OP_CONJUGATE) return HMVM_ERR_INVALID_ARG;
    if (p->device < 0) return HMVM_ERR_INVALID_ARG;                   // host-only plan
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t nin = trans ? p->nrows : p->ncols, nout = trans ? p->ncols : p->nrows;
    if ((nin > 0 && !x) || (nout > 0 && !y)) return HMVM_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) % 8) return HMVM_ERR_LAYOUT;
    if (alpha == 0.0) return HMVM_OK;
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) return HMVM_ERR_DEVICE;
    if (prev != p->device && hipSetDevice(p->device) != hipSuccess) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t kr = p->row.K, kc = p->col.K;
    const int64_t wsd = p->ws_doubles + 32;
    void* mem = nullptr;
    int rc = HMVM_OK;
    if (hipMallocAsync(&mem, size_t(kr + kc + wsd + 64) * 8, st) != hipSuccess) rc = HMVM_ERR_ALLOC;
    if (rc == HMVM_OK) {
        double* cr = static_cast<double*>(mem);            // row-tree coefficients
        double* cc = cr + ((kr + 31) / 32) * 32;           // column-tree coefficients
        double* ws = cc + ((kc + 31) / 32) * 32;
        const int T = HMVM_OP_TRANSPOSED, Nn = HMVM_OP_NORMAL;
        auto run = [&](int r) { if (rc == HMVM_OK) rc = r; };
        if (!trans) {
            zero(cc, kc, st);
            run(sub(p->B, 1.0, Nn, x, cc, ws, wsd, st));
            for (int d = int(p->tcol.size()); d >= 1; --d) {
                const Level& L = p->tcol[size_t(d) - 1];
                run(sub(L.p, 1.0, T, cc + L.son0, cc + L.par0, ws, wsd, st));
            }
            zero(cr, kr, st);
            run(sub(p->C, alpha, Nn, cc, cr, ws, wsd, st));
            for (int d = 1; d <= int(p->trow.size()); ++d) {
                const Level& L = p->trow[size_t(d) - 1];
                run(sub(L.p, 1.0, Nn, cr + L.par0, cr + L.son0, ws, wsd, st));
            }
            run(sub(p->D, 1.0, Nn, cr, y, ws, wsd, st));
            run(sub(p->N, alpha, Nn, x, y, ws, wsd, st));
        } else {
            zero(cr, kr, st);
            run(sub(p->D, 1.0, T, x, cr, ws, wsd, st));
            for (int d = int(p->trow.size()); d >= 1; --d) {
                const Level& L = p->trow[size_t(d) - 1];
                run(sub(L.p, 1.0, T, cr + L.son0, cr + L.par0, ws, wsd, st));
            }
            zero(cc, kc, st);
            run(sub(p->C, alpha, T, cr, cc, ws, wsd, st));
            for (int d = 1; d <= int(p->tcol.size()); ++d) {
                const Level& L = p->tcol[size_t(d) - 1];
                run(sub(L.p, 1.0, Nn, cc + L.par0, cc + L.son0, ws, wsd, st));
            }
            run(sub(p->B, 1.0, T, cc, y, ws, wsd, st));
            run(sub(p->N, alpha, T, x, y, ws, wsd, st));
        }
        if (hipGetLastError() != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
        if (hipFreeAsync(mem, st) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    }
    if (prev != p->device) (void)hipSetDevice(prev);
    return rc;
}

} // extern "C"
