// hmvm.hip -- kernels and C-ABI of the MI355X H-/TLR-matrix-vector product.
//
// Forward y += a A x (arith/mulvec.hh:27-64 semantics, leaves as in
// matrix/lrmatrix.hh:625-679 and matrix/dense_matrix.hh:459-481):
//   k_vfwd    P[e, q] = sum_{i in q} V(i, e) x_i     one wave per pack of <= 64
//                                                    column groups (lane = group)
//   k_psum    t_e = a S_e sum_q P[e, q]              lanes per entry, wave tree
//   k_ufwd    y_p += sum_c R_p(:, c) z_c             one workgroup per row
//             z_c = a x_j (dense) | t_e (low-rank)   panel, lane = row
// Transposed y += a A^T x (arith/blas.hh:910-948):
//   k_rtrn_lds Q[c] = sum_{i in p} R_p(i, c) x_i     wave per work item, lane =
//                                                    row, column sums via LDS
//   k_psum    s_e = a S_e sum_p Q[e, p]
//   k_vtrn    y_q += V_q s + a sum Q_dense           workgroup per column panel
// No atomics: every output element has exactly one writer and a fixed
// summation order, so results are bitwise reproducible run to run.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "hmvm.h"
#include "kernels.hpp"
#include "fwd.hpp"
#include "mrhs.hpp"
#include "internal.hpp"
#include "plan_host.hpp"

using namespace hmvm;

namespace {


// ---------------------------------------------------------------------------
// forward: V pass, one wave per pack of consecutive segments (lane = group)
// ---------------------------------------------------------------------------

// at least 4 waves per SIMD (at most 128 VGPRs), the occupancy the pack
// stream was measured at
template <typename XT, int FAM>
__global__ __launch_bounds__(kThreads, 4) void k_vfwd(DevPlan P, const XT* __restrict__ x, double* __restrict__ Pb) {
    __shared__ double xs[kWaves][kPackXs];
    // the wave index as a wave-uniform value: the pack's table is then read
    // with scalar loads
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t pi = int64_t(blockIdx.x) * kWaves + wave;
    if (pi >= P.nvpacks) return;
    const VPack* pk = P.vpacks + pi;
    vpack_stage_x(pk, x, xs[wave], lane);
    wave_sync();
    HMVM_FMT_SWITCH_FAM(FAM, pk->fmt, vpack_fwd<T>(P, pk, Pb, xs[wave], lane))
}

// t_e = alpha * S_e * sum of the entry's partials, in one kernel: the lanes
// of entry e are adjacent in one wave; lane i sums segments i, i + L, ... of
// kSegLen partials (each in order), then a fixed tree over the entry's lanes
// (lane i takes lane i + s when i % 2s == 0) leaves the sum in lane 0.
__global__ __launch_bounds__(256) void k_psum(const SumLane* __restrict__ lanes, const int32_t* __restrict__ lil,
                                              const double* __restrict__ tmul, const double* __restrict__ Pb,
                                              double* __restrict__ t, int64_t nl, double alpha) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;   // nl is a multiple of 64
    SumLane ln{0, 0, 0, -1};
    int32_t il = 0;
    if (g < nl) { ln = lanes[g]; il = lil[g]; }
    const int32_t e = ln.e;
    const int i = il & 255, L = il >> 8;
    // the entry's factor, loaded with the partials (index 0 for idle lanes)
    // rather than after the tree: one load round trip less
    const double tm = tmul[e >= 0 ? e : 0];
    double v = 0;
    if (e >= 0) {
        const int32_t o = ln.off, c = ln.cnt, k = ln.stride;
        for (int32_t r0 = i * kSegLen; r0 < c; r0 += L * kSegLen) {
            double a[kSegLen];
#pragma unroll
            for (int r = 0; r < kSegLen; ++r) a[r] = (r0 + r < c) ? Pb[o + int64_t(r0 + r) * k] : 0.0;
            double s = 0;
#pragma unroll
            for (int r = 0; r < kSegLen; ++r) s += a[r];
            v += s;
        }
    }
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const double w = __shfl_down(v, s, 64);
        if ((i & (2 * s - 1)) == 0 && i + s < L) v += w;
    }
    if (e >= 0 && i == 0) t[e] = alpha * tm * v;
}

// ---------------------------------------------------------------------------
// forward: R pass (dense rows and U rows of a row panel), one workgroup per
// row panel; a chunk's z values are staged in LDS by the whole workgroup and
// each wave streams a contiguous quarter of the chunk's groups
// ---------------------------------------------------------------------------

template <typename XT, typename YT, bool kWhole, int FAM>
__global__ __launch_bounds__(kThreads) void k_ufwd(DevPlan P, double alpha, const XT* __restrict__ x,
                                                   const double* __restrict__ t, YT* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char zdyn[];
    __shared__ double red[kWaves][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const RowPanel R = P.rp[blockIdx.x];
    double acc = 0;
    if (kWhole) {
        panel_stage(P, R, alpha, x, t, zdyn, tid);
        __syncthreads();
        for (int ci = 0; ci < R.nchunks; ++ci) {
            const RChunk C = P.rchunks[R.chunk0 + ci];
            const void* zb = zdyn + C.zoff;
            HMVM_FMT_SWITCH_FAM(FAM, C.fmt, acc += chunk_compute<T>(P, C, R.m, zb, wave, lane))
        }
    } else {
        acc = rpanel_fwd_chunked<XT, FAM>(P, R, alpha, x, t, zdyn, tid, wave, lane);
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && lane < R.m) {
        const double s = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
        y[R.row0 + lane] = YT(double(y[R.row0 + lane]) + s);
    }
}

// ---------------------------------------------------------------------------
// transposed: R pass with column reduction, one wave per chunk
// ---------------------------------------------------------------------------

// A wave takes one work item: at most kRtSpan consecutive groups of a chunk
// (every column belongs to one item, so each Q entry has one writer).
// Transposed R pass through LDS instead of cross-lane reductions: a step
// takes GS groups (GS W <= TC columns); every lane (row i) writes its
// products R(i, c) x_i into the wave's tile [row][col] (row stride TS
// elements, 16-byte aligned), then lane L sums rows L / TC, L / TC + LPC,
// ... of column L % TC in order and log2(LPC) shuffles add the LPC partial
// sums; lane c < GS W writes Q for column c.  FP32 accumulation: TC = 32,
// FP64: TC = 16; either way the wave's tile is 9 KB, which leaves room for
// 4 workgroups per CU.
template <typename A> struct RtTile;
template <> struct RtTile<float> { static constexpr int TC = 32, TS = 36; };
template <> struct RtTile<double> { static constexpr int TC = 16, TS = 18; };
constexpr size_t kRtTileBytes = 64 * 36 * 4;
static_assert(64 * RtTile<double>::TS * 8 == kRtTileBytes, "one tile size for both accumulation types");

template <typename T, typename XT>
__global__ __launch_bounds__(kThreads) void k_rtrn_lds(DevPlan P, const RtItem* __restrict__ items, int64_t n,
                                                       const XT* __restrict__ x, double* __restrict__ Q) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    constexpr int TC = RtTile<A>::TC, TS = RtTile<A>::TS, LPC = 64 / TC;
    constexpr int GS = TC / W;                      // groups per step
    static_assert(GS >= 1, "a step holds at least one group");
    __shared__ __align__(16) unsigned char tiles[kWaves][kRtTileBytes];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t idx = int64_t(blockIdx.x) * kWaves + wave;
    if (idx >= n) return;                           // whole wave; no workgroup barrier follows
    A* tile = reinterpret_cast<A*>(tiles[wave]);
    const RtItem wi = items[idx];
    const RChunk C = P.rchunks[wi.chunk];
    const int gend = min(wi.g0 + kRtSpan, int(C.ngroups));
    const RowPanel R = P.rp[C.panel];
    const int m = R.m;
    const bool on = lane < m;
    const A xi = on ? A(x[R.row0 + lane]) : A(0);
    const bytep base = elem_addr<T>(P.arena[E::F], C.data_off) + int64_t(on ? lane : 0) * 16;
    const int64_t gs = int64_t(m) * 16;             // bytes per group
    const int c = lane % TC, h = lane / TC;
#pragma unroll 1
    for (int g0 = wi.g0; g0 < gend; g0 += GS) {
        u32x4 raw[GS];
#pragma unroll
        for (int u = 0; u < GS; ++u)
            raw[u] = (on && g0 + u < gend) ? raw16(base + int64_t(g0 + u) * gs) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            A v[W];
            E::decode(raw[u], v);                   // rows past m and groups past gend: zeros
#pragma unroll
            for (int w = 0; w < W; ++w) tile[lane * TS + u * W + w] = v[w] * xi;
        }
        wave_sync();
        A s = A(0);
#pragma unroll
        for (int r = 0; r < 64 / LPC; ++r) s += tile[(LPC * r + h) * TS + c];
#pragma unroll
        for (int o = TC; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
        const int ncol = min(GS, gend - g0) * W;    // columns of this step
        if (h == 0 && c < ncol) {
            const int64_t gc = C.col0 + int64_t(g0) * W + c;
            const int32_t d = P.rdest[gc];
            double sv = double(s);
            if (E::kScaled) sv *= double(P.zscale[gc]);
            if (d >= 0) Q[d] = sv;
        }
        wave_sync();                                // the tile is rewritten by the next step
    }
}

// ---------------------------------------------------------------------------
// transposed: V pass, y_q += V_q s  (+ dense partials), workgroup per panel
// ---------------------------------------------------------------------------

constexpr int kTRows = 4;           // rows reduced together in the transposed V pass
constexpr int kVtLoads = 8;         // rows loaded per step of the transposed V pass

// lane = group of the item; rows in blocks of kTRows: each lane forms its
// rows' dot products with s, the wave reduces them across the groups, and
// the lanes L with L % (64 / kTRows) == 0 add the sum of row
// kTRows b + L / (64 / kTRows) to the wave's own row sums in LDS (rw, 64
// entries)
template <typename T>
__device__ inline void item_vtrn(const DevPlan& P, const ColPanel& C, const VItem& it,
                                 const double* __restrict__ s, double* rw, int lane) {
    using E = Elem<T>;
    using A = typename E::Acc;
    constexpr int W = E::W;
    const bool on = lane < it.ng;
    const int g = on ? lane : 0;
    // s values (times the column scales) in the accumulation type: a row's
    // dot product over the lane's W columns is formed as in the forward
    // passes (FP32 for the FP32-accumulated formats), the row sums across
    // lanes in FP64
    // All loads unconditional (lanes past the item's groups read group 0's
    // entries and are masked afterwards): loaded under divergent
    // conditions, each column's source and value had waited for the
    // column before, two round trips per column.
    A sv[W];
    int32_t ev[W];
    float sc[W];
    double xv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) ev[w] = P.vsrc[it.vc0 + int64_t(g) * W + w];
#pragma unroll
    for (int w = 0; w < W; ++w) sc[w] = E::kScaled ? P.vscale[it.vc0 + int64_t(g) * W + w] : 1.0f;
#pragma unroll
    for (int w = 0; w < W; ++w) xv[w] = s[(on && ev[w] >= 0) ? ev[w] : 0];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        double v = (on && ev[w] >= 0) ? xv[w] : 0.0;
        if (E::kScaled && on) v *= double(sc[w]);
        sv[w] = A(v);
    }
    const int64_t ldb = int64_t(it.ng) * 16;            // bytes per row of the item
    const bytep base = elem_addr<T>(P.arena[E::F], it.data_off) + int64_t(g) * 16;
    constexpr int NL = kVtLoads;
    static_assert(NL % kTRows == 0, "a step covers whole row blocks");
#pragma unroll 1
    for (int i0 = 0; i0 < C.m; i0 += NL) {
        // rows of the step, all loads in flight together (lanes past the
        // item's groups read group 0 and meet sv = 0; rows past m are not read)
        u32x4 raw[NL];
#pragma unroll
        for (int r = 0; r < NL; ++r)
            raw[r] = (i0 + r < C.m) ? raw16(base + int64_t(i0 + r) * ldb) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int sb = 0; sb < NL / kTRows; ++sb) {
            double d[kTRows];
#pragma unroll
            for (int r = 0; r < kTRows; ++r) {
                A v[W];
                E::decode(raw[sb * kTRows + r], v);
                A acc = A(0);
#pragma unroll
                for (int w = 0; w < W; ++w) acc += v[w] * sv[w];
                d[r] = double(acc);
            }
            const double r = xreduce<kTRows>(d, lane);
            const int row = i0 + sb * kTRows + lane / (64 / kTRows);
            if ((lane % (64 / kTRows)) == 0 && row < 64) rw[row] += r;
        }
    }
}

template <typename YT, int FAM>
__global__ __launch_bounds__(kThreads) void k_vtrn(DevPlan P, double alpha, const double* __restrict__ s,
                                                   const double* __restrict__ Q, YT* __restrict__ y) {
    __shared__ double red[kWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const ColPanel C = P.cp[blockIdx.x];
    red[wave][lane] = 0.0;          // the wave's row sums; only this wave touches its row
    for (int j = wave; j < C.nitems; j += kWaves) {
        const VItem it = P.vitems[C.item0 + j];
        HMVM_FMT_SWITCH_FAM(FAM, it.fmt, item_vtrn<T>(P, C, it, s, red[wave], lane))
    }
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
__global__ __launch_bounds__(kThreads) void k_grad_r(DevPlan P, const RtItem* __restrict__ items, int64_t n,
                                                     const XT* __restrict__ x, const XT* __restrict__ ybar,
                                                     const double* __restrict__ t, double* __restrict__ gR) {
    constexpr int W = Elem<T>::W;
    constexpr int NZ = kRtSpan * W;                      // columns of one item at most
    __shared__ double zsh[kWaves][NZ];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t idx = int64_t(blockIdx.x) * kWaves + wave;
    if (idx >= n) return;                                // whole wave: no barrier below spans waves
    const RtItem wi = items[idx];
    const RChunk C = P.rchunks[wi.chunk];
    const int gend = min(wi.g0 + kRtSpan, int(C.ngroups));
    const RowPanel R = P.rp[C.panel];
    const int m = R.m;
    // the item's z values into the wave's LDS slice: all column sources
    // first, then all values (two load round trips for the item)
    const int ncol = (gend - wi.g0) * W;
    const int64_t c0 = C.col0 + int64_t(wi.g0) * W;
    double* zs = zsh[wave];
    constexpr int PER = (NZ + 63) / 64;
    int32_t src[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int j = lane + 64 * u;
        src[u] = (j < ncol) ? P.colsrc[c0 + j] : kPadCol;
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int j = lane + 64 * u;
        const int32_t sj = src[u];
        const double z = (sj >= 0) ? double(x[sj]) : (sj == kPadCol ? 0.0 : t[~sj]);
        if (j < NZ) zs[j] = z;
    }
    wave_sync();
    if (lane >= m) return;
    const double yb = double(ybar[R.row0 + lane]);
    // a lane's W values of a group are adjacent (16-byte stores when W is
    // even: offsets are multiples of W), consecutive lanes adjacent
    for (int g = wi.g0; g < gend; ++g) {
        const double* zg = zs + (g - wi.g0) * W;
        double* d = gR + C.data_off + (int64_t(g) * m + lane) * W;
        if constexpr (W % 2 == 0) {
#pragma unroll
            for (int w = 0; w < W; w += 2) reinterpret_cast<double2*>(d)[w / 2] = double2{yb * zg[w], yb * zg[w + 1]};
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) d[w] = yb * zg[w];
        }
    }
}

struct GradPtrs {
    double* p[NFMT];
};

// dV(i, c) = x_i * s_e for the W columns of group `lane` of segment `it`:
// row by row, so that a wave writes each row of the segment (ng W values)
// as one contiguous run, each lane its W values together (16-byte stores
// when W is even; the segment's offset and row length are multiples of W)
template <int W, typename XT>
__device__ inline void grad_v_group(const DevPlan& P, const VItem& it, const XT* __restrict__ x,
                                    const double* __restrict__ s, double* __restrict__ gV, int lane) {
    const int64_t ld = int64_t(it.ng) * W;
    double sv[W];
    const int64_t vc0 = it.vc0 + int64_t(lane) * W;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int32_t e = P.vsrc[vc0 + w];
        sv[w] = (e >= 0) ? s[e] : 0.0;
    }
    double* dst = gV + it.data_off + int64_t(lane) * W;
    const int m = it.m;
    const int64_t col0 = it.x0;
    for (int i = 0; i < m; ++i) {
        const double xi = double(x[col0 + i]);
        double* d = dst + int64_t(i) * ld;
        if constexpr (W % 2 == 0) {
#pragma unroll
            for (int w = 0; w < W; w += 2) reinterpret_cast<double2*>(d)[w / 2] = double2{xi * sv[w], xi * sv[w + 1]};
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) d[w] = xi * sv[w];
        }
    }
}

// dV(i, c) = x_i * s_e, one wave per item, lane = group
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_grad_v(DevPlan P, const XT* __restrict__ x,
                                                     const double* __restrict__ s, GradPtrs gA) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t item = int64_t(blockIdx.x) * kWaves + wave;
    if (item >= P.nvitems) return;
    const VItem it = P.vitems[item];
    double* gV = nullptr;             // selected without dynamic indexing (which would go through scratch)
#pragma unroll
    for (int f = 0; f < NFMT; ++f)
        if (it.fmt == f) gV = gA.p[f];
    if (!gV || lane >= it.ng) return;
    HMVM_FMT_SWITCH(it.fmt, grad_v_group<Elem<T>::W>(P, it, x, s, gV, lane))
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
    void* arena_dev[NFMT] = {};
    // R work items (chunk, first group; at most kRtSpan groups) grouped by
    // format, for the transposed and gradient launches
    const RtItem* rtitems_dev = nullptr;
    int64_t rtitem_start[NFMT + 1] = {};
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
    UP(tstride, L.tstride);
    UP(vitems, L.vitems);
    UP(fitems, L.fitems);
    UP(vpacks, L.vpacks);
    UP(pl, L.pl);
    UP(pl_il, L.pl_il);
    UP(ql, L.ql);
    UP(ql_il, L.ql_il);
#undef UP
    D.npl = int64_t(L.pl.size());
    D.nql = int64_t(L.ql.size());
    std::vector<RtItem> items;
    for (int f = 0; f < NFMT; ++f) {
        p->rtitem_start[f] = int64_t(items.size());
        for (size_t c = 0; c < L.rchunks.size(); ++c)
            if (L.rchunks[c].fmt == f)
                for (int32_t g0 = 0; g0 < L.rchunks[c].ngroups; g0 += kRtSpan) items.push_back({int32_t(c), g0});
    }
    p->rtitem_start[NFMT] = int64_t(items.size());
    if ((rc = upload(p, items, &p->rtitems_dev)) != HMVM_OK) return rc;
    // arenas
    const void* a[NFMT] = {};
    if ((rc = upload(p, L.a64, reinterpret_cast<const double**>(&a[F64]))) != HMVM_OK) return rc;
    if ((rc = upload(p, L.a32, reinterpret_cast<const float**>(&a[F32]))) != HMVM_OK) return rc;
    if ((rc = upload(p, L.a16, reinterpret_cast<const uint16_t**>(&a[BF16]))) != HMVM_OK) return rc;
    if ((rc = upload(p, L.a8, reinterpret_cast<const uint8_t**>(&a[FP8]))) != HMVM_OK) return rc;
    for (int f = kFirstBfp; f < NFMT; ++f)
        if ((rc = upload(p, L.abfp[f - kFirstBfp], reinterpret_cast<const uint32_t**>(&a[f]))) != HMVM_OK) return rc;
    for (int f = 0; f < NFMT; ++f) { D.arena[f] = a[f]; p->arena_dev[f] = const_cast<void*>(a[f]); }
    D.nrp = int64_t(L.rp.size());
    D.ncp = int64_t(L.cp.size());
    D.nvitems = int64_t(L.vitems.size());
    D.nfitems = int64_t(L.fitems.size());
    D.nvpacks = int64_t(L.vpacks.size());
    D.nrchunks = int64_t(L.rchunks.size());
    D.ntent = L.ntent;
    D.psize = L.psize;
    D.qsize = L.qsize;
    D.nrows = L.nrows;
    D.ncols = L.ncols;
    // the host copies of the arenas are no longer needed on the device path
    L.release_arenas();
    return HMVM_OK;
}

// transposed R pass of one format
template <typename T, typename XT>
void launch_rtrn(const DevPlan& D, const RtItem* ids, int64_t n, const XT* x, double* Q, unsigned nb, hipStream_t st) {
    k_rtrn_lds<T, XT><<<nb, kThreads, 0, st>>>(D, ids, n, x, Q);
}

// calls f(std::integral_constant<int, FAM>) for the plan's format family
template <typename F>
void by_family(int fam, F&& f) {
    if (fam == kFamBfp) f(std::integral_constant<int, kFamBfp>());
    else if (fam == kFamAfl) f(std::integral_constant<int, kFamAfl>());
    else f(std::integral_constant<int, kFamHw>());
}

inline unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

inline int64_t pad32(int64_t v) { return ((v + 31) / 32) * 32; }

// workspace: partials | t (or s) + 32 spare
int64_t ws_doubles(const hmvm_plan* p, int op) {
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const HostLayout& L = p->L;
    const int64_t a = trans ? L.qsize : L.psize;
    return pad32(a) + pad32(L.ntent) + 32;
}

// t = alpha S sum(partials), one kernel (k_psum)
void launch_psum(const DevPlan& D, bool trans, double* t, double alpha, const double* part, hipStream_t st) {
    const int64_t nl = trans ? D.nql : D.npl;
    if (D.ntent == 0 || nl == 0) return;
    k_psum<<<blocks_for(nl, 256), 256, 0, st>>>(trans ? D.ql : D.pl, trans ? D.ql_il : D.pl_il, D.tmul, part, t, nl,
                                                alpha);
}

template <typename XT, typename YT>
int launch_apply(const hmvm_plan* p, double alpha, int op, const XT* x, YT* y, double* ws, hipStream_t st,
                 hipEvent_t* ev = nullptr) {
    const DevPlan& D = p->dp;
    const HostLayout& L = p->L;
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t abuf = trans ? L.qsize : L.psize;
    double* part = ws;
    double* t = ws + pad32(abuf);
    if (ev) (void)hipEventRecord(ev[0], st);
    if (!trans) {
        if (D.nvpacks > 0)
            by_family(p->L.family, [&](auto fam) {
                k_vfwd<XT, decltype(fam)::value><<<blocks_for(D.nvpacks, kWaves), kThreads, 0, st>>>(D, x, part);
            });
        if (ev) (void)hipEventRecord(ev[1], st);
        launch_psum(D, false, t, alpha, part, st);
        if (ev) (void)hipEventRecord(ev[2], st);
        if (D.nrp > 0) {
            // chunk staging needs kChunkGroups * 16 floats (FP8) or * 2 doubles (FP64): 16 KB
            by_family(p->L.family, [&](auto fam) {
                constexpr int FAM = decltype(fam)::value;
                if (p->L.max_zbytes <= kWholeZBytes)
                    k_ufwd<XT, YT, true, FAM><<<unsigned(D.nrp), kThreads, size_t(p->L.max_zbytes), st>>>(D, alpha, x,
                                                                                                          t, y);
                else
                    k_ufwd<XT, YT, false, FAM><<<unsigned(D.nrp), kThreads, size_t(kChunkGroups * 16 * 4), st>>>(
                        D, alpha, x, t, y);
            });
        }
        if (ev) (void)hipEventRecord(ev[3], st);
    } else {
        for (int f = 0; f < NFMT; ++f) {
            const int64_t n = p->rtitem_start[f + 1] - p->rtitem_start[f];
            if (n == 0) continue;
            const RtItem* ids = p->rtitems_dev + p->rtitem_start[f];
            const unsigned nb = blocks_for(n, kWaves);
            HMVM_FMT_SWITCH(f, launch_rtrn<T, XT>(D, ids, n, x, part, nb, st))
        }
        if (ev) (void)hipEventRecord(ev[1], st);
        launch_psum(D, true, t, alpha, part, st);
        if (ev) (void)hipEventRecord(ev[2], st);
        if (D.ncp > 0)
            by_family(p->L.family, [&](auto fam) {
                k_vtrn<YT, decltype(fam)::value><<<unsigned(D.ncp), kThreads, 0, st>>>(D, alpha, t, part, y);
            });
        if (ev) (void)hipEventRecord(ev[3], st);
    }
    return hipGetLastError() == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

// workspace of a call: the single right-hand-side layout, or the matrix-core
// layout for forward products with several right-hand sides
int64_t call_ws_doubles(const hmvm_plan* p, int op, int nrhs) {
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    int64_t w = ws_doubles(p, op);
    if (!trans && nrhs > 1) {
        // sizes from the host layout, so that host-only plans answer too
        DevPlan d{};
        d.psize = p->L.psize;
        d.ntent = p->L.ntent;
        d.ncols = p->L.ncols;
        w = std::max(w, mrhs_ws_doubles(d, std::min(nrhs, kMrhsMax)));
    }
    return w;
}

int apply_impl(const hmvm_plan* p, double alpha, int op, const void* x, int64_t ldx, void* y, int64_t ldy,
               int nrhs, int vdt, void* ws, int64_t ws_bytes, hipStream_t st) {
    const bool trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t nin = trans ? p->L.nrows : p->L.ncols;
    const int64_t nout = trans ? p->L.ncols : p->L.nrows;
    if (nrhs > 1 && (ldx < nin || ldy < nout)) return HMVM_ERR_LAYOUT;
    if (ws_bytes < call_ws_doubles(p, op, nrhs) * int64_t(sizeof(double))) return HMVM_ERR_INVALID_ARG;
    const size_t es = (vdt == HMVM_F64) ? 8 : 4;
    if (!trans && nrhs > 1) {
        // matrix cores, up to kMrhsMax right-hand sides per pass
        for (int r0 = 0; r0 < nrhs; r0 += kMrhsMax) {
            const int R = std::min(kMrhsMax, nrhs - r0);
            const void* xr = static_cast<const char*>(x) + size_t(r0) * size_t(ldx) * es;
            void* yr = static_cast<char*>(y) + size_t(r0) * size_t(ldy) * es;
            int rc;
            if (R == 1) {
                rc = (vdt == HMVM_F64)
                         ? launch_apply(p, alpha, op, static_cast<const double*>(xr), static_cast<double*>(yr),
                                        static_cast<double*>(ws), st)
                         : launch_apply(p, alpha, op, static_cast<const float*>(xr), static_cast<float*>(yr),
                                        static_cast<double*>(ws), st);
            } else {
                rc = launch_apply_mrhs(p->dp, p->L.only_f64() ? int(kFamF64) : p->L.family, alpha, xr, ldx, yr, ldy, R, vdt, static_cast<double*>(ws), st) == 0
                         ? HMVM_OK : HMVM_ERR_DEVICE;
            }
            if (rc != HMVM_OK) return rc;
        }
        return HMVM_OK;
    }
    // transposed (and single right-hand side): one vector at a time
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

// ---------------------------------------------------------------------------
// fixed device scratch of hmvm_apply_noalloc
// ---------------------------------------------------------------------------

// One array per GPU (each device that loads the library holds its own copy
// until the process ends).  Host code never passes `scratch` itself to a
// kernel: on the host the name stands for a host-side shadow, so its device
// address is looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_dev.
static_assert(HMVM_NOALLOC_MAX_BYTES % 256 == 0, "scratch size is a multiple of its alignment");
__device__ __align__(256) double scratch[HMVM_NOALLOC_MAX_BYTES / sizeof(double)];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};

// device address of the scratch array of device `dev`, which must be the
// current device
int scratch_address(int dev, void** out) {
    *out = nullptr;
    if (dev < 0 || dev >= kMaxScratchDevices) return HMVM_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* a = nullptr;
        if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess || !a) return HMVM_ERR_DEVICE;
        g_scratch_dev[dev] = a;
    }
    *out = g_scratch_dev[dev];
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
    case HMVM_ERR_IO: return "file missing, short, or not an H-matrix file";
    case HMVM_ERR_CORRUPT: return "H-matrix file fails its checksums";
    case HMVM_ERR_TOO_LARGE: return "workspace of the call exceeds the fixed scratch buffer (HMVM_NOALLOC_MAX_BYTES)";
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
    info->scale_bytes = L.scale_bytes;
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
    *bytes = call_ws_doubles(p, op, nrhs) * int64_t(sizeof(double));
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
    const int64_t bytes = call_ws_doubles(p, op, nrhs) * int64_t(sizeof(double));
    void* ws = nullptr;
    if (hipMallocAsync(&ws, size_t(bytes), st) != hipSuccess) return HMVM_ERR_ALLOC;
    rc = apply_impl(p, alpha, op, x, ldx, y, ldy, nrhs, vdt, ws, bytes, st);
    if (hipFreeAsync(ws, st) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    return rc;
}

int hmvm_apply_noalloc(const hmvm_plan* p, double alpha, int op, const void* x, int64_t ldx, void* y, int64_t ldy,
                       int nrhs, int vdt, void* stream) {
    int rc = check_apply_args(p, op, x, y, nrhs, vdt);
    if (rc != HMVM_OK) return rc;
    if (alpha == 0.0) return HMVM_OK;
    const int64_t bytes = call_ws_doubles(p, op, nrhs) * int64_t(sizeof(double));
    if (bytes > HMVM_NOALLOC_MAX_BYTES) return HMVM_ERR_TOO_LARGE;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    void* ws = nullptr;
    if ((rc = scratch_address(p->device, &ws)) != HMVM_OK) return rc;
    // every workspace value the kernels read is written earlier in the same
    // call, so nothing is cleared between calls
    return apply_impl(p, alpha, op, x, ldx, y, ldy, nrhs, vdt, ws, bytes, static_cast<hipStream_t>(stream));
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

int hmvm_grad(const hmvm_plan* p, const void* x, const void* ybar, void* gx, void* grad_arena[NFMT], int vdt,
              void* stream) {
    int rc = check_apply_args(p, HMVM_OP_NORMAL, x, ybar, 1, vdt);
    if (rc != HMVM_OK) return rc;
    // gx is read and written as vector elements; the gradient kernels store
    // 16 bytes (two doubles) at a time into the arenas
    if (reinterpret_cast<uintptr_t>(gx) % ((vdt == HMVM_F64) ? 8 : 4)) return HMVM_ERR_LAYOUT;
    if (grad_arena)
        for (int f = 0; f < NFMT; ++f)
            if (reinterpret_cast<uintptr_t>(grad_arena[f]) % 16) return HMVM_ERR_LAYOUT;
    DeviceGuard g(p->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const HostLayout& L = p->L;
    const DevPlan& D = p->dp;
    const int64_t np = pad32(L.psize), nq = pad32(L.qsize), ne = pad32(L.ntent) + 32;
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
        if (D.nvpacks > 0)
            by_family(p->L.family, [&](auto fam) {
                k_vfwd<XT, decltype(fam)::value><<<blocks_for(D.nvpacks, kWaves), kThreads, 0, st>>>(D, xx, Pb);
            });
        for (int f = 0; f < NFMT; ++f) {
            const int64_t n = p->rtitem_start[f + 1] - p->rtitem_start[f];
            if (n == 0) continue;
            const RtItem* ids = p->rtitems_dev + p->rtitem_start[f];
            const unsigned nb = blocks_for(n, kWaves);
            HMVM_FMT_SWITCH(f, launch_rtrn<T, XT>(D, ids, n, yb, Qb, nb, st))
        }
        launch_psum(D, false, t, 1.0, Pb, st);
        launch_psum(D, true, s, 1.0, Qb, st);
        if (gx && D.ncp > 0)
            by_family(p->L.family, [&](auto fam) {
                k_vtrn<XT, decltype(fam)::value><<<unsigned(D.ncp), kThreads, 0, st>>>(D, 1.0, s, Qb,
                                                                                       static_cast<XT*>(gx));
            });
        if (grad_arena) {
            for (int f = 0; f < NFMT; ++f) {
                if (!grad_arena[f]) continue;
                double* gA = static_cast<double*>(grad_arena[f]);
                const int64_t nr = p->rtitem_start[f + 1] - p->rtitem_start[f];
                if (nr > 0) {
                    const RtItem* ids = p->rtitems_dev + p->rtitem_start[f];
                    const unsigned nb = blocks_for(nr, kWaves);
                    HMVM_FMT_SWITCH(f, k_grad_r<T, XT><<<nb, kThreads, 0, st>>>(D, ids, nr, xx, yb, t, gA))
                }
            }
            GradPtrs gp;
            for (int f = 0; f < NFMT; ++f) gp.p[f] = static_cast<double*>(grad_arena[f]);
            if (D.nvitems > 0) k_grad_v<XT><<<blocks_for(D.nvitems, kWaves), kThreads, 0, st>>>(D, xx, s, gp);
        }
        return hipGetLastError() == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
    };
    rc = (vdt == HMVM_F64) ? run(double(0)) : run(float(0));
    if (hipFreeAsync(ws, st) != hipSuccess && rc == HMVM_OK) rc = HMVM_ERR_DEVICE;
    return rc;
}

int hmvm_plan_arena_elems(const hmvm_plan* p, int fmt, int64_t* elems) {
    if (!p || !elems || fmt < 0 || fmt >= NFMT) return HMVM_ERR_INVALID_ARG;
    *elems = p->L.arena_elems(fmt);
    return HMVM_OK;
}

int hmvm_plan_get_arena(const hmvm_plan* p, int fmt, void* dst, void* stream) {
    if (!p || !dst || fmt < 0 || fmt >= NFMT) return HMVM_ERR_INVALID_ARG;
    const int64_t bytes = p->L.bytes_by_format[fmt];
    if (bytes == 0) return HMVM_OK;
    if (p->device < 0) {
        std::memcpy(dst, p->L.arena_host(fmt), size_t(bytes));
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

int hmvm_plan_unpack_block(const hmvm_plan* p, int64_t block, const void* const host_arenas[NFMT], int as_f64,
                           double* a, double* b) {
    if (!p || !host_arenas) return HMVM_ERR_INVALID_ARG;
    return unpack_block(p->L, block, host_arenas, as_f64 != 0, a, b);
}

int hmvm_plan_arena_source(const hmvm_plan* p, int fmt, int64_t* src) {
    if (!p) return HMVM_ERR_INVALID_ARG;
    try {
        return arena_source(p->L, fmt, src);
    } catch (const std::bad_alloc&) {
        return HMVM_ERR_ALLOC;
    }
}

} // extern "C"

namespace hmvm {
const DevPlan& plan_devplan(const hmvm_plan* p) { return p->dp; }
const HostLayout& plan_layout(const hmvm_plan* p) { return p->L; }
int plan_device(const hmvm_plan* p) { return p->device; }
int64_t plan_fwd_ws_doubles(const hmvm_plan* p, int R) { return call_ws_doubles(p, HMVM_OP_NORMAL, R); }
} // namespace hmvm
