// batch.hip -- forward products of many H-matrices and right-hand sides in
// one fixed sequence of launches.
//
// Every member keeps its own plan.  A batch lists the work of all members as
// units of six kinds; one kernel per kind (and per tile count for VM and
// RM) runs every unit of that kind of every member, one unit per workgroup:
//   XT  transposed copy of X (64 indices) for members with several
//       right-hand sides
//   V1  column-panel pass of members with one right-hand side (4 packs, one
//       per wave)
//   VM  column-panel pass on the matrix cores (one segment, nrhs >= 2; one
//       kernel per count of 16-wide right-hand-side tiles)
//   PS  partial sums (a range of (entry, pair of right-hand sides) items,
//       or (entry, rhs) items for odd R)
//   R1  row-panel pass, one right-hand side (one row panel)
//   RM  row-panel pass on the matrix cores (one row panel, nrhs >= 2; one
//       kernel per tile count)
// The launches are ordered on the caller's stream, which is the only
// synchronisation: no workgroup waits for another.  The launch count is at
// most twelve whatever the number of members, and the whole sequence
// (descriptor upload plus the kernels) can be captured in a HIP graph and
// replayed.

#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "fwd.hpp"
#include "hmvm.h"
#include "internal.hpp"
#include "mrhs_dev.hpp"

using namespace hmvm;

namespace {

// matrix-core kinds come in four variants by tiles of 16 right-hand sides
// (accumulators and LDS staging sized for ceil(R / 16) tiles of the member)
enum UnitKind : int {
    K_XT = 0, K_V1, K_VM1, K_VM2, K_VM3, K_VM4, K_PS, K_R1, K_RM1, K_RM2, K_RM3, K_RM4, kKinds
};
constexpr int kConfigs = 4;                     // {member order, largest first} x {256, 1024} sum items
constexpr int kSumItems[2] = {256, 1024};       // (entry, rhs pair | rhs) items per sum unit

struct BatchMat {
    DevPlan D;
    const void* x;
    void* y;
    int64_t ldx, ldy;
    double alpha;
    double* part;
    double* t;
    double* xt;               // X^T (ncols x R) for R >= 2
    int32_t R;
    int32_t pad_;
};

struct BatchUnit {
    int32_t mat, first, count, pad_;
};

template <typename XT, int FAM>
__global__ __launch_bounds__(kThreads) void k_bv1(const BatchMat* __restrict__ mats,
                                                  const BatchUnit* __restrict__ units) {
    extern __shared__ double lds[];
    // wave-uniform wave index: the pack's table is read with scalar loads
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const BatchUnit w = units[blockIdx.x];
    const BatchMat& M = mats[w.mat];
    if (wave >= w.count) return;
    const VPack* pk = M.D.vpacks + w.first + wave;
    const XT* x = static_cast<const XT*>(M.x);
    double* xs = lds + wave * kPackXs;
    vpack_stage_x(pk, x, xs, lane);
    wave_sync();
    HMVM_FMT_SWITCH_FAM(FAM, pk->fmt, vpack_fwd<T>(M.D, pk, M.part, xs, lane))
}

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_bxt(const BatchMat* __restrict__ mats,
                                                  const BatchUnit* __restrict__ units) {
    extern __shared__ double lds[];
    const BatchUnit w = units[blockIdx.x];
    const BatchMat& M = mats[w.mat];
    xtrans_tile(static_cast<const XT*>(M.x), M.ldx, M.D.ncols, M.R, M.xt, int64_t(w.first) * 64, lds, threadIdx.x);
}

template <int NTM, int FAM>
__global__ __launch_bounds__(kThreads) void k_bvm(const BatchMat* __restrict__ mats,
                                                  const BatchUnit* __restrict__ units) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const BatchUnit w = units[blockIdx.x];
    const BatchMat& M = mats[w.mat];
    vseg_mm_dispatch<NTM, FAM, false>(M.D, M.D.fitems[w.first], M.xt, M.R, M.part, lds, tid, wave, lane);
}

// partial sums of one member: items (entry, pair of right-hand sides) when R
// is even, else (entry, right-hand side); the workspace is 256-byte aligned
static_assert(kSumItems[1] <= kBsumItems * kThreads, "a sum unit is at most kBsumItems items per thread");
__global__ __launch_bounds__(kThreads) void k_bps(const BatchMat* __restrict__ mats,
                                                  const BatchUnit* __restrict__ units) {
    const BatchUnit w = units[blockIdx.x];
    const BatchMat& M = mats[w.mat];
    const int64_t g0 = int64_t(w.first) + threadIdx.x, g1 = int64_t(w.first) + w.count;
    if (M.R % 2 == 0)
        psum_mm_items<2, kBsumItems, kBsumRound>(M.D, M.part, M.t, M.R, M.alpha, g0, kThreads, g1);
    else
        psum_mm_items<1, kBsumItems, kBsumRound>(M.D, M.part, M.t, M.R, M.alpha, g0, kThreads, g1);
}

template <typename XT, typename YT, int FAM>
__global__ __launch_bounds__(kThreads) void k_br1(const BatchMat* __restrict__ mats,
                                                  const BatchUnit* __restrict__ units) {
    extern __shared__ double lds[];
    __shared__ double red[kWaves][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const BatchUnit w = units[blockIdx.x];
    const BatchMat& M = mats[w.mat];
    const RowPanel Rw = M.D.rp[w.first];
    const double acc = rpanel_fwd_chunked<XT, FAM>(M.D, Rw, M.alpha, static_cast<const XT*>(M.x), M.t, lds, tid, wave,
                                                   lane);
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && lane < Rw.m) {
        YT* y = static_cast<YT*>(M.y);
        const double s = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
        y[Rw.row0 + lane] = YT(double(y[Rw.row0 + lane]) + s);
    }
}

template <int NTM, typename YT, int FAM>
__global__ __launch_bounds__(kThreads) void k_brm(const BatchMat* __restrict__ mats,
                                                  const BatchUnit* __restrict__ units) {
    extern __shared__ double lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const BatchUnit w = units[blockIdx.x];
    const BatchMat& M = mats[w.mat];
    const RowPanel Rw = M.D.rp[w.first];
    rpanel_mm<NTM, YT, FAM>(M.D, Rw, M.alpha, M.xt, M.t, static_cast<YT*>(M.y), M.ldy, M.R, lds, tid, wave, lane);
}

inline int64_t pad32(int64_t v) { return ((v + 31) / 32) * 32; }

} // namespace

struct hmvm_batch {
    int device = -1;
    int vdt = HMVM_F64;
    int nmat = 0;
    std::vector<const hmvm_plan*> plans;
    std::vector<int> nrhs;
    std::vector<BatchMat> proto;            // per member, pointers to vectors filled at apply time
    BatchMat* h_mats = nullptr;             // pinned host staging of the descriptors
    BatchMat* d_mats = nullptr;             // descriptors of eager launches (apply, tune)
    BatchMat* d_mats_graph = nullptr;       // the captured graph's own descriptors
    // per configuration: the units of all kinds back to back, kind k at
    // kind_ofs[c][k] .. kind_ofs[c][k + 1]
    std::vector<BatchUnit> units[kConfigs];
    int64_t kind_ofs[kConfigs][kKinds + 1] = {};
    BatchUnit* d_units[kConfigs] = {};
    size_t lds[kKinds] = {};
    int config = 1;
    int family = kFamAll;                   // the members' common format family, else all formats
    double* d_ws = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> allocs;
};

namespace {

struct Guard {
    int prev = -1;
    bool ok = true;
    explicit Guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) { ok = false; return; }
        if (prev != dev && hipSetDevice(dev) != hipSuccess) ok = false;
    }
    ~Guard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

int fill_descriptors(hmvm_batch* b, const double* alphas, const void* const* xs, const int64_t* ldx,
                     void* const* ys, const int64_t* ldy) {
    for (int i = 0; i < b->nmat; ++i) {
        if (!xs[i] || !ys[i]) return HMVM_ERR_INVALID_ARG;
        const HostLayout& L = plan_layout(b->plans[i]);
        const int R = b->nrhs[i];
        const int64_t lx = ldx ? ldx[i] : L.ncols, ly = ldy ? ldy[i] : L.nrows;
        if (R > 1 && (lx < L.ncols || ly < L.nrows)) return HMVM_ERR_LAYOUT;
        BatchMat m = b->proto[i];
        m.x = xs[i];
        m.y = ys[i];
        m.ldx = lx;
        m.ldy = ly;
        m.alpha = alphas ? alphas[i] : 1.0;
        b->h_mats[i] = m;
    }
    return HMVM_OK;
}

template <typename XT, typename YT, int FAM>
void launch_kinds(hmvm_batch* b, const BatchMat* mats, hipStream_t st) {
    const int c = b->config;
    const BatchUnit* u = b->d_units[c];
    const int64_t* o = b->kind_ofs[c];
    auto grid = [&](int k) { return unsigned(o[k + 1] - o[k]); };
    if (grid(K_XT)) k_bxt<XT><<<grid(K_XT), kThreads, b->lds[K_XT], st>>>(mats, u + o[K_XT]);
    if (grid(K_V1)) k_bv1<XT, FAM><<<grid(K_V1), kThreads, b->lds[K_V1], st>>>(mats, u + o[K_V1]);
    if (grid(K_VM1)) k_bvm<1, FAM><<<grid(K_VM1), kThreads, b->lds[K_VM1], st>>>(mats, u + o[K_VM1]);
    if (grid(K_VM2)) k_bvm<2, FAM><<<grid(K_VM2), kThreads, b->lds[K_VM2], st>>>(mats, u + o[K_VM2]);
    if (grid(K_VM3)) k_bvm<3, FAM><<<grid(K_VM3), kThreads, b->lds[K_VM3], st>>>(mats, u + o[K_VM3]);
    if (grid(K_VM4)) k_bvm<4, FAM><<<grid(K_VM4), kThreads, b->lds[K_VM4], st>>>(mats, u + o[K_VM4]);
    if (grid(K_PS)) k_bps<<<grid(K_PS), kThreads, 0, st>>>(mats, u + o[K_PS]);
    if (grid(K_R1)) k_br1<XT, YT, FAM><<<grid(K_R1), kThreads, b->lds[K_R1], st>>>(mats, u + o[K_R1]);
    if (grid(K_RM1)) k_brm<1, YT, FAM><<<grid(K_RM1), kThreads, b->lds[K_RM1], st>>>(mats, u + o[K_RM1]);
    if (grid(K_RM2)) k_brm<2, YT, FAM><<<grid(K_RM2), kThreads, b->lds[K_RM2], st>>>(mats, u + o[K_RM2]);
    if (grid(K_RM3)) k_brm<3, YT, FAM><<<grid(K_RM3), kThreads, b->lds[K_RM3], st>>>(mats, u + o[K_RM3]);
    if (grid(K_RM4)) k_brm<4, YT, FAM><<<grid(K_RM4), kThreads, b->lds[K_RM4], st>>>(mats, u + o[K_RM4]);
}

// Eager launches (mats == b->d_mats) upload the pinned descriptors first, a
// copy on the stream; a captured graph (mats == b->d_mats_graph) takes
// descriptors uploaded before its capture, so that no node of the graph
// reads the shared staging buffer, which later calls rewrite.
int launch(hmvm_batch* b, BatchMat* mats, hipStream_t st) {
    if (mats == b->d_mats &&
        hipMemcpyAsync(b->d_mats, b->h_mats, sizeof(BatchMat) * size_t(b->nmat), hipMemcpyHostToDevice, st) !=
            hipSuccess)
        return HMVM_ERR_DEVICE;
    // the matrix-core kernels come in one instantiation per format family
    auto go = [&](auto fam) {
        constexpr int FAM = decltype(fam)::value;
        if (b->vdt == HMVM_F64)
            launch_kinds<double, double, FAM>(b, mats, st);
        else
            launch_kinds<float, float, FAM>(b, mats, st);
    };
    switch (b->family) {
    case kFamHw: go(std::integral_constant<int, kFamHw>()); break;
    case kFamBfp: go(std::integral_constant<int, kFamBfp>()); break;
    case kFamF64: go(std::integral_constant<int, kFamF64>()); break;
    case kFamAfl: go(std::integral_constant<int, kFamAfl>()); break;
    default: go(std::integral_constant<int, kFamAll>()); break;
    }
    return hipGetLastError() == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

template <typename V>
int dev_alloc(hmvm_batch* b, V** out, size_t n) {
    *out = nullptr;
    if (n == 0) return HMVM_OK;
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(V)) != hipSuccess) return HMVM_ERR_ALLOC;
    b->allocs.push_back(p);
    *out = static_cast<V*>(p);
    return HMVM_OK;
}

void release(hmvm_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->h_mats) (void)hipHostFree(b->h_mats);
}

} // namespace

extern "C" {

int hmvm_batch_create(const hmvm_plan* const* plans, int nplans, const int* nrhs, int vdt, hmvm_batch** out) {
    if (!out) return HMVM_ERR_INVALID_ARG;
    *out = nullptr;
    if (!plans || nplans < 1 || (vdt != HMVM_F64 && vdt != HMVM_F32)) return HMVM_ERR_INVALID_ARG;
    const int dev = plan_device(plans[0]);
    for (int i = 0; i < nplans; ++i) {
        if (!plans[i] || plan_device(plans[i]) != dev || dev < 0) return HMVM_ERR_INVALID_ARG;
        const int R = nrhs ? nrhs[i] : 1;
        if (R < 1) return HMVM_ERR_INVALID_ARG;
        if (R > kMaxRp) return HMVM_ERR_UNSUPPORTED;
    }
    std::unique_ptr<hmvm_batch> b;
    try {
        b.reset(new hmvm_batch);
    } catch (...) {
        return HMVM_ERR_ALLOC;
    }
    b->device = dev;
    b->vdt = vdt;
    b->nmat = nplans;
    b->plans.assign(plans, plans + nplans);
    for (int i = 0; i < nplans; ++i) b->nrhs.push_back(nrhs ? nrhs[i] : 1);
    // the members' common family: FP64-only members (their own
    // instantiation when all are) fit either family; hardware formats
    // beside block floating point need the all-format kernels
    auto fam_of = [](const HostLayout& L) { return L.only_f64() ? int(kFamF64) : L.family; };
    auto join = [](int a, int c) { return a == c ? a : a == kFamF64 ? c : c == kFamF64 ? a : int(kFamAll); };
    b->family = nplans > 0 ? fam_of(plan_layout(plans[0])) : kFamHw;
    for (int i = 1; i < nplans; ++i) b->family = join(b->family, fam_of(plan_layout(plans[i])));
    Guard g(dev);
    if (!g.ok) return HMVM_ERR_DEVICE;

    // per-member descriptors and workspaces; units of every kind with their
    // algorithmic bytes (the sort key of the largest-first order)
    struct Sized { BatchUnit u; double bytes; };
    std::vector<Sized> kinds[2][kKinds];        // [sum granularity][kind]
    std::vector<int64_t> wsoff(nplans);
    int64_t wsz = 0;
    int maxRp = 16;
    for (int i = 0; i < nplans; ++i) {
        const DevPlan& D = plan_devplan(plans[i]);
        const HostLayout& L = plan_layout(plans[i]);
        const int R = b->nrhs[i];
        if (R > 1) maxRp = std::max(maxRp, (R + 15) & ~15);
        BatchMat m{};
        m.D = D;
        m.R = R;
        wsoff[i] = wsz;
        wsz += pad32(D.psize * R) + pad32(D.ntent * R) + 32 + (R > 1 ? pad32(D.ncols * R) : 0);
        for (int gsel = 0; gsel < 2; ++gsel) {
            auto& K = kinds[gsel];
            if (R == 1) {
                for (int64_t k = 0; k < D.nvpacks; k += kWaves) {
                    const int cnt = int(std::min<int64_t>(kWaves, D.nvpacks - k));
                    double by = 0;
                    for (int j = 0; j < cnt; ++j) {
                        const VPack& pk = L.vpacks[k + j];
                        for (int t = 0; t < pk.nseg; ++t) by += double(pk.ng[t]) * double(pk.m[t]) * 16.0;
                    }
                    K[K_V1].push_back({{i, int32_t(k), cnt, 0}, by});
                }
            } else {
                for (int64_t k = 0; k < D.nfitems; ++k)
                    K[K_VM1 + (R + 15) / 16 - 1].push_back(
                        {{i, int32_t(k), 1, 0}, double(L.fitems[k].ng) * 16.0 * double(L.fitems[k].m) * R});
                for (int64_t i0 = 0; i0 < D.ncols; i0 += 64)
                    K[K_XT].push_back({{i, int32_t(i0 / 64), 1, 0}, double(std::min<int64_t>(64, D.ncols - i0)) * R});
            }
            const int items = kSumItems[gsel];
            const int64_t nsum = D.ntent * (R % 2 == 0 ? R / 2 : R);
            for (int64_t k = 0; k < nsum; k += items)
                K[K_PS].push_back({{i, int32_t(k), int32_t(std::min<int64_t>(items, nsum - k)), 0},
                                   double(std::min<int64_t>(items, nsum - k))});
            for (int64_t p = 0; p < D.nrp; ++p) {
                const RowPanel& rp = L.rp[p];
                double by = 0;
                for (int64_t c = rp.chunk0; c < rp.chunk0 + rp.nchunks; ++c)
                    by += double(L.rchunks[c].ngroups) * rp.m * 16.0;
                K[R == 1 ? K_R1 : K_RM1 + (R + 15) / 16 - 1].push_back({{i, int32_t(p), 1, 0}, by * R});
            }
        }
        b->proto.push_back(m);
    }
    // configuration c: order = c / 2 (0 member order, 1 largest first),
    // sum granularity = c % 2
    for (int c = 0; c < kConfigs; ++c) {
        const int order = c / 2, gsel = c % 2;
        b->kind_ofs[c][0] = 0;
        for (int k = 0; k < kKinds; ++k) {
            std::vector<Sized> v = kinds[gsel][k];
            if (order == 1)
                std::stable_sort(v.begin(), v.end(), [](const Sized& a, const Sized& z) { return a.bytes > z.bytes; });
            for (const Sized& s : v) b->units[c].push_back(s.u);
            b->kind_ofs[c][k + 1] = int64_t(b->units[c].size());
        }
    }
    b->lds[K_XT] = mm_lds_bytes(maxRp);
    b->lds[K_V1] = size_t(kWaves) * kPackXs * sizeof(double);
    for (int k = K_VM1; k <= K_VM4; ++k) b->lds[k] = mm_lds_bytes(16 * (k - K_VM1 + 1));
    b->lds[K_R1] = size_t(kChunkGroups) * 16 * sizeof(float);
    for (int k = K_RM1; k <= K_RM4; ++k) b->lds[k] = mm_lds_bytes(16 * (k - K_RM1 + 1));

    int rc;
    if ((rc = dev_alloc(b.get(), &b->d_mats, size_t(nplans))) != HMVM_OK) { release(b.get()); return rc; }
    if ((rc = dev_alloc(b.get(), &b->d_mats_graph, size_t(nplans))) != HMVM_OK) { release(b.get()); return rc; }
    for (int c = 0; c < kConfigs; ++c)
        if ((rc = dev_alloc(b.get(), &b->d_units[c], b->units[c].size())) != HMVM_OK) { release(b.get()); return rc; }
    if ((rc = dev_alloc(b.get(), &b->d_ws, size_t(wsz))) != HMVM_OK) { release(b.get()); return rc; }
    if (hipHostMalloc(reinterpret_cast<void**>(&b->h_mats), sizeof(BatchMat) * size_t(nplans)) != hipSuccess) {
        b->h_mats = nullptr;
        release(b.get());
        return HMVM_ERR_ALLOC;
    }
    for (int c = 0; c < kConfigs; ++c)
        if (!b->units[c].empty() &&
            hipMemcpy(b->d_units[c], b->units[c].data(), sizeof(BatchUnit) * b->units[c].size(),
                      hipMemcpyHostToDevice) != hipSuccess) {
            release(b.get());
            return HMVM_ERR_DEVICE;
        }
    for (int i = 0; i < nplans; ++i) {
        const DevPlan& D = b->proto[i].D;
        const int R = b->proto[i].R;
        double* base = b->d_ws + wsoff[i];
        b->proto[i].part = base;
        b->proto[i].t = base + pad32(D.psize * R);
        b->proto[i].xt = (R > 1) ? b->proto[i].t + pad32(D.ntent * R) + 32 : nullptr;
    }
    *out = b.release();
    return HMVM_OK;
}

int hmvm_batch_destroy(hmvm_batch* b) {
    if (!b) return HMVM_OK;
    {
        Guard g(b->device);
        (void)hipDeviceSynchronize();
        release(b);
    }
    delete b;
    return HMVM_OK;
}

int hmvm_batch_apply(hmvm_batch* b, const double* alphas, const void* const* xs, const int64_t* ldx,
                     void* const* ys, const int64_t* ldy, void* stream) {
    if (!b || !xs || !ys) return HMVM_ERR_INVALID_ARG;
    Guard g(b->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the pinned descriptors may still be read by an earlier launch
    if (hipStreamSynchronize(st) != hipSuccess) return HMVM_ERR_DEVICE;
    const int rc = fill_descriptors(b, alphas, xs, ldx, ys, ldy);
    if (rc != HMVM_OK) return rc;
    return launch(b, b->d_mats, st);
}

int hmvm_batch_capture(hmvm_batch* b, const double* alphas, const void* const* xs, const int64_t* ldx,
                       void* const* ys, const int64_t* ldy, void* stream) {
    if (!b || !xs || !ys) return HMVM_ERR_INVALID_ARG;
    Guard g(b->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (hipStreamSynchronize(st) != hipSuccess) return HMVM_ERR_DEVICE;
    int rc = fill_descriptors(b, alphas, xs, ldx, ys, ldy);
    if (rc != HMVM_OK) return rc;
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    // the graph's descriptors: uploaded once, synchronously, after every
    // earlier replay (on any stream) has finished reading them
    if (hipDeviceSynchronize() != hipSuccess) return HMVM_ERR_DEVICE;
    if (hipMemcpy(b->d_mats_graph, b->h_mats, sizeof(BatchMat) * size_t(b->nmat), hipMemcpyHostToDevice) !=
        hipSuccess)
        return HMVM_ERR_DEVICE;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return HMVM_ERR_DEVICE;
    rc = launch(b, b->d_mats_graph, st);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc != HMVM_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != HMVM_OK ? rc : HMVM_ERR_DEVICE;
    }
    b->graph = graph;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
        b->exec = nullptr;
        return HMVM_ERR_DEVICE;
    }
    return HMVM_OK;
}

int hmvm_batch_replay(hmvm_batch* b, void* stream) {
    if (!b || !b->exec) return HMVM_ERR_INVALID_ARG;
    Guard g(b->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    return hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

int hmvm_batch_check(hmvm_batch* b, void* stream) {
    if (!b) return HMVM_ERR_INVALID_ARG;
    Guard g(b->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    return hipStreamSynchronize(static_cast<hipStream_t>(stream)) == hipSuccess ? HMVM_OK : HMVM_ERR_DEVICE;
}

int hmvm_batch_tune(hmvm_batch* b, const void* const* xs, const int64_t* ldx, void* const* ys, const int64_t* ldy,
                    void* stream, int* chosen) {
    if (!b || !xs || !ys) return HMVM_ERR_INVALID_ARG;
    Guard g(b->device);
    if (!g.ok) return HMVM_ERR_DEVICE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (hipStreamSynchronize(st) != hipSuccess) return HMVM_ERR_DEVICE;
    // every candidate accumulates alpha = 1 products into the caller's y,
    // which the caller resets afterwards
    int rc = fill_descriptors(b, nullptr, xs, ldx, ys, ldy);
    if (rc != HMVM_OK) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return HMVM_ERR_DEVICE;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return HMVM_ERR_DEVICE; }
    const int keep = b->config;
    float best = 1e30f;
    int best_cfg = -1;
    for (int c = 0; c < kConfigs && rc == HMVM_OK; ++c) {
        b->config = c;
        if ((rc = launch(b, b->d_mats, st)) != HMVM_OK) break;   // warm-up
        float tot = 0;
        for (int rep = 0; rep < 3 && rc == HMVM_OK; ++rep) {
            if (hipEventRecord(e0, st) != hipSuccess) { rc = HMVM_ERR_DEVICE; break; }
            rc = launch(b, b->d_mats, st);
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = HMVM_ERR_DEVICE;
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tot += ms;
        }
        if (rc == HMVM_OK && tot < best) { best = tot; best_cfg = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HMVM_OK) { b->config = keep; return rc; }
    b->config = best_cfg;
    if (chosen) *chosen = best_cfg;
    return HMVM_OK;
}

} // extern "C"
