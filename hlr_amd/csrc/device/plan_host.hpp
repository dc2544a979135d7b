// plan_host.hpp -- host-side layout builder of the device plan.
#pragma once
#include <cstdint>
#include <vector>

#include "hmvm.h"
#include "plan.hpp"

namespace hmvm {

struct HostLayout {
    int64_t nrows = 0, ncols = 0;
    int32_t precision = HMVM_PREC_FP64;
    int32_t dense_fmt = F64;
    double eps = 0;
    bool one_level = false;
    std::vector<hmvm_block> blocks;

    // panels, chunks, work items
    std::vector<RowPanel> rp;
    std::vector<RChunk> rchunks;
    std::vector<ColPanel> cp;
    std::vector<VItem> fitems;     // V segments (forward passes), by column set, unit and format
    std::vector<VItem> vitems;     // their views per column panel, panel-major (transposed, gradient)
    std::vector<VPack> vpacks;     // forward V work units

    // per R column (all row panels, chunk order)
    std::vector<int32_t> colsrc;   // >= 0: x index; < 0: ~t-entry; kPadCol: padding
    std::vector<float> zscale;     // multiplier of the gathered value (empty: all 1)
    std::vector<int32_t> rdest;    // transposed: Q index (-1: padding)
    // per V column (all column panels, format-major inside a panel)
    std::vector<int32_t> vdest;    // forward: P index (-1: padding)
    std::vector<int32_t> vsrc;     // transposed: t-entry (-1: padding)
    std::vector<float> vscale;     // multiplier of the partial sum (empty: all 1)

    // t entries (one per low-rank column l of block b)
    int64_t ntent = 0;
    std::vector<int64_t> blk_t0;   // per block: first t-entry (-1 for dense)
    std::vector<int8_t> tfmt;      // per t-entry: storage format
    std::vector<int64_t> blk_d0;   // per block: first dense column in dfmt (-1 for low-rank)
    std::vector<int8_t> dfmt;      // per dense column (block by block): storage format
    std::vector<double> tmul;      // per t-entry: S_l for U S V^T leaves, else 1
    std::vector<int32_t> pt_off, pt_cnt;   // forward partials of t-entry e
    std::vector<int32_t> qt_off, qt_cnt;   // transposed partials of t-entry e
    std::vector<int32_t> tstride;          // distance of consecutive partials of e (= rank)
    // one-kernel sums for one right-hand side: lanes of t-entry e (at most 64,
    // adjacent, inside one wave); lane i sums segments i, i + L, ... of
    // kSegLen partials and the wave adds up the entry's lanes.  Per lane: e
    // (-1: padding) and i | L << 8
    std::vector<SumLane> pl;   std::vector<int32_t> pl_il;   // forward (P)
    std::vector<SumLane> ql;   std::vector<int32_t> ql_il;   // transposed (Q)
    std::vector<int32_t> qd_off, qd_cnt;   // transposed dense partials per column
    int64_t psize = 0, qsize = 0;

    // factor arenas per format; block floating point rows as 4 words each
    std::vector<double> a64;
    std::vector<float> a32;
    std::vector<uint16_t> a16;
    std::vector<uint8_t> a8;
    std::vector<uint32_t> abfp[kNumBfp];
    int64_t nvals[NFMT] = {};      // stored values per format (arena offsets are value offsets)

    // statistics
    int64_t ndense = 0, nlowrank = 0;
    int64_t bytes_by_format[NFMT] = {};
    int64_t scale_bytes = 0;          // panel-column scales read by the kernels
    int64_t cols_by_format[NFMT] = {};
    double flops = 0, datasize = 0, bytes_rpass = 0, bytes_vpass = 0;
    int32_t max_vitem_rows = 0;
    int32_t family = kFamHw;       // format family of every stored column
    int64_t max_zbytes = 0;        // largest per-row-panel z block (bytes)

    int64_t arena_elems(int f) const { return nvals[f]; }
    // host copy of the arena of format f (null when empty or released)
    const void* arena_host(int f) const {
        switch (f) {
        case F64: return a64.empty() ? nullptr : a64.data();
        case F32: return a32.empty() ? nullptr : a32.data();
        case BF16: return a16.empty() ? nullptr : a16.data();
        case FP8: return a8.empty() ? nullptr : a8.data();
        default: return abfp[f - kFirstBfp].empty() ? nullptr : abfp[f - kFirstBfp].data();
        }
    }
    // true when every stored value is FP64
    bool only_f64() const {
        for (int f = 1; f < NFMT; ++f)
            if (nvals[f] > 0) return false;
        return true;
    }
    void release_arenas() {
        std::vector<double>().swap(a64);
        std::vector<float>().swap(a32);
        std::vector<uint16_t>().swap(a16);
        std::vector<uint8_t>().swap(a8);
        for (auto& v : abfp) std::vector<uint32_t>().swap(v);
    }
};

// validate the leaves and build the layout; returns an HMVM_* code
int build_layout(const hmvm_block* blocks, int64_t nblocks, int64_t nrows, int64_t ncols,
                 const double* arena, int64_t arena_len, const hmvm_plan_opts& opts,
                 HostLayout& L);

// dequantise one block's factors from (host copies of) the arenas
int unpack_block(const HostLayout& L, int64_t b, const void* const arenas[NFMT], bool as_f64, double* a, double* bb);

// for every stored value of format f, its offset in the plan's input arena (-1: padding)
int arena_source(const HostLayout& L, int f, int64_t* src);

// storage helpers shared with tests
uint8_t encode_e4m3(float v);
float decode_e4m3(uint8_t c);
uint16_t encode_bf16(float v);
float decode_bf16(uint16_t c);
// block floating point: step exponent of a panel column, and one value
int bfp_exp(double maxabs, int nb);
int32_t bfp_quant(double v, int k, int nb);

} // namespace hmvm
