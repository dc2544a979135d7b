// plan_build.cpp -- validation, precision choice, panel layout and packing.
//
// Host-only code.  Builds the device layout described in plan.hpp from the
// flattened leaves.  The precision rule restates the reference's VALR
// chooser: per-column tolerance u_l = delta / sigma_l from
// valr::get_tolerances (compress/valr.hh:402-512), mantissa bits
// m_l = max(1, ceil(-log2 u_l)) from afl::eps_to_rate (compress/afl.hh:81-82),
// and per-column scaling as in afl::compress_lr (compress/afl.hh:818-885);
// here m_l selects the narrowest hardware format with at least m_l mantissa
// bits (FP8-E4M3: 3, BF16: 7, FP32: 23, FP64: 52) instead of a bitstream.
//
// HMVM_PREC_BFP stores columns as block floating point, as bfp::compress_lr
// (compress/bfp.hh:765-860) does with one scale per column: here the scale
// is a power of two per panel column and the value width NB is 8, 16 or 25
// bits (16, 8 or 5 values per 16-byte row).  The width of a low-rank column
// is the narrowest whose stored U and V columns are both within the
// column's VALR tolerance u_l of the FP64 columns in the 2-norm -- the
// relative column error the VALR budget allows, which AFL's truncation to
// m_l bits also stays within -- so the VALR error budget holds.

#include "plan_host.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <omp.h>

namespace hmvm {

// ---------------------------------------------------------------------------
// number formats
// ---------------------------------------------------------------------------

uint8_t encode_e4m3(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    const uint8_t sign = uint8_t((u >> 24) & 0x80u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return 0x7f;                  // NaN
    const float a = std::fabs(v);
    if (!(a < 464.0f)) return sign | 0x7e;             // saturate at 448
    if (a < 0.015625f) {                               // below 2^-6: multiples of 2^-9
        const float q = std::nearbyint(a * 512.0f);    // exact scaling, RNE
        return sign | uint8_t(q);                      // q == 8 encodes 2^-6 (e=1, m=0)
    }
    int e = int(u >> 23) - 127;
    const uint32_t m23 = u & 0x7fffffu;
    uint32_t r = m23 >> 20;
    const uint32_t rem = m23 & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) ++r;
    if (r == 8) { r = 0; ++e; }
    uint32_t code = (uint32_t(e + 7) << 3) | r;
    if (code > 0x7eu) code = 0x7eu;
    return sign | uint8_t(code);
}

float decode_e4m3(uint8_t c) {
    const int s = c >> 7, e = (c >> 3) & 0xf, m = c & 7;
    if (e == 0xf && m == 7) return std::nanf("");
    float v = (e == 0) ? std::ldexp(float(m), -9) : std::ldexp(1.0f + float(m) / 8.0f, e - 7);
    return s ? -v : v;
}

uint16_t encode_bf16(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

float decode_bf16(uint16_t c) {
    const uint32_t u = uint32_t(c) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Step exponent k of a panel column of NB-bit values: the smallest k with
// max|v| <= (2^(NB-1) - 1) 2^k, kept in FP32's normal range [-126, 127]
// (the scale 2^k is stored as a float).  0 for an all-zero column.
int bfp_exp(double maxabs, int nb) {
    if (!(maxabs > 0) || !std::isfinite(maxabs)) return 0;
    const double qmax = std::ldexp(1.0, nb - 1) - 1.0;
    int k = int(std::ceil(std::log2(maxabs / qmax)));
    while (maxabs > qmax * std::ldexp(1.0, k)) ++k;          // guard the log2 rounding
    while (maxabs <= qmax * std::ldexp(1.0, k - 1)) --k;
    return std::min(127, std::max(-126, k));
}

// q = v / 2^k rounded to nearest (ties to even), clamped to +-(2^(NB-1) - 1)
int32_t bfp_quant(double v, int k, int nb) {
    const double qmax = std::ldexp(1.0, nb - 1) - 1.0;
    double q = std::nearbyint(std::ldexp(v, -k));
    if (!(q <= qmax)) q = (q < 0) ? -qmax : qmax;        // also catches NaN (never stored)
    if (q < -qmax) q = -qmax;
    return int32_t(q);
}

namespace {

// NB of each block floating point format and of each AFL format, narrowest first
constexpr int kNumWidths = 3;
constexpr int kBfpBits[kNumWidths] = {8, 16, 25};
constexpr int kAflBits[kNumWidths] = {12, 16, 24};
static_assert(fmt_bits(B8) == 8 && fmt_bits(B16) == 16 && fmt_bits(B25) == 25, "format table");
static_assert(fmt_bits(A12) == 12 && fmt_bits(A16) == 16 && fmt_bits(A24) == 24, "format table");
static_assert(B25 - B8 + 1 == kNumWidths && A24 - A12 + 1 == kNumWidths, "format table");

// ---- adaptive float (compress/afl.hh:220-420 with fixed code widths) ----
// A value v of a panel column with scale s is a = |v| / s + 1 >= 2; its code
// keeps the sign, the low kAflExpBits bits of a's binary exponent (stored as
// exponent - 1: a's biased FP32 exponent modulo 2^5) and the leading
// mantissa bits of a, truncated.  The all-ones code is zero.

// scale of a panel column: its smallest non-zero |v| (afl.hh:835-854),
// raised to max|v| 2^-(2^5) where the exponent field cannot span the
// column's range, rounded down to an FP32 value
float afl_scale(const double* v, int64_t m) {
    double vmax = 0, vmin = HUGE_VAL;
    for (int64_t i = 0; i < m; ++i) {
        const double a = std::fabs(v[i]);
        vmax = std::max(vmax, a);
        if (a > 0) vmin = std::min(vmin, a);
    }
    if (!(vmax > 0)) return 1.0f;
    const double s = std::max(vmin, std::ldexp(vmax, -(1 << kAflExpBits)));
    float f = float(s);
    if (double(f) > s) f = std::nextafter(f, 0.0f);
    return f > 0 ? f : FLT_MIN;
}

uint32_t afl_code(double v, float s, int nb) {
    const int m = nb - 1 - kAflExpBits;
    const uint32_t zero = (1u << nb) - 1u;
    const double a = std::fabs(v) / double(s) + 1.0;
    if (!(a >= 2.0) || !std::isfinite(a)) return zero;     // below the scale: zero
    int E;
    const double fr = std::frexp(a, &E);                   // a = fr 2^E, fr in [0.5, 1)
    if (E - 2 > (1 << kAflExpBits) - 1) return zero;       // beyond the range (not for a column's own scale)
    const uint32_t mant = uint32_t(std::floor(std::ldexp(2.0 * fr - 1.0, m)));
    return (uint32_t(v < 0) << (nb - 1)) | (uint32_t(E - 2) << m) | mant;
}

// +-(a - 1) of a code, in FP32 as the kernels decode it (scale not applied)
float afl_unscaled(uint32_t c, int nb) {
    const int m = nb - 1 - kAflExpBits;
    if (c == (1u << nb) - 1u) return 0.0f;
    const uint32_t e = (c >> m) & ((1u << kAflExpBits) - 1u), mant = c & ((1u << m) - 1u);
    const uint32_t bits = ((e | 0x80u) << 23) | (mant << (23 - m));
    float a;
    std::memcpy(&a, &bits, 4);
    const float d = a - 1.0f;
    return (c >> (nb - 1)) ? -d : d;
}

// sum of squared errors of one panel column stored with NB-bit values of
// the block floating point (afl false) or AFL (afl true) formats
double packed_err2(const double* src, int64_t m, int nb, bool afl) {
    double e2 = 0;
    if (afl) {
        const float s = afl_scale(src, m);
        for (int64_t i = 0; i < m; ++i) {
            const double d = double(afl_unscaled(afl_code(src[i], s, nb), nb)) * double(s) - src[i];
            e2 += d * d;
        }
        return e2;
    }
    double mx = 0;
    for (int64_t i = 0; i < m; ++i) mx = std::max(mx, std::fabs(src[i]));
    const int k = bfp_exp(mx, nb);
    for (int64_t i = 0; i < m; ++i) {
        const double d = std::ldexp(double(bfp_quant(src[i], k, nb)), k) - src[i];
        e2 += d * d;
    }
    return e2;
}

struct Piece {            // one panel column: m values from src
    const double* src;
    int64_t m;
};

// the narrowest packed format (block floating point, or AFL when afl) whose
// stored pieces of every column in `cols` are within tol (relative, 2-norm)
// of the column; else FP32 when rate <= 23 (its rounding error
// 2^-24 <= 2^-(rate+1) <= tol), else FP64
int packed_choose(const std::vector<std::vector<Piece>>& cols, double tol, int rate, bool afl) {
    std::vector<double> thr(cols.size());
    for (size_t c = 0; c < cols.size(); ++c) {
        double n2 = 0;
        for (const Piece& pc : cols[c])
            for (int64_t i = 0; i < pc.m; ++i) n2 += pc.src[i] * pc.src[i];
        thr[c] = (tol * tol) * n2;
    }
    for (int f = 0; f < kNumWidths; ++f) {
        const int nb = afl ? kAflBits[f] : kBfpBits[f];
        bool ok = true;
        for (size_t c = 0; c < cols.size() && ok; ++c) {
            double e2 = 0;
            for (const Piece& pc : cols[c]) e2 += packed_err2(pc.src, pc.m, nb, afl);
            ok = (e2 <= thr[c]);
        }
        if (ok) return (afl ? kFirstAfl : kFirstBfp) + f;
    }
    return rate <= 23 ? F32 : F64;
}

// write NB-bit field u (the low NB bits are kept) as field j of the 16-byte row w[0..3]
void packed_put(uint32_t* w, int j, int nb, uint32_t u) {
    u &= (1u << nb) - 1u;
    const int bit = j * nb, word = bit >> 5, off = bit & 31;
    w[word] |= u << off;
    if (off + nb > 32) w[word + 1] |= u >> (32 - off);
}

int eps_to_rate(double eps) {
    const double r = std::ceil(-std::log2(eps));
    return int(r < 1.0 ? 1.0 : r);
}

int format_for_rate(int rate) {
    if (rate <= 3) return FP8;
    if (rate <= 7) return BF16;
    if (rate <= 23) return F32;
    return F64;
}

void valr_tolerances(int64_t k, const double* sv, double eps, double* out) {
    double tol = eps * sv[0];
    double sum = 0;
    for (int64_t i = 0; i < k; ++i) sum += 1.0 / sv[i];
    const double p = 1.0 / sum, q = tol / sum;
    double t = p * p / 4 - q;
    if (t < 0) t = 0;
    const double d1 = std::fabs(p / 2 + std::sqrt(t)), d2 = std::fabs(p / 2 - std::sqrt(t));
    tol = std::min(d1, d2);
    for (int64_t l = 0; l < k; ++l) out[l] = tol / sv[l];
}

// split [a,b) into pieces of at most kPanel, as even as possible
void split_interval(int64_t a, int64_t b, std::vector<std::pair<int64_t, int32_t>>& out) {
    const int64_t len = b - a;
    if (len <= 0) return;
    const int64_t np = (len + kPanel - 1) / kPanel;
    const int64_t base = len / np, extra = len % np;
    int64_t s = a;
    for (int64_t i = 0; i < np; ++i) {
        const int64_t l = base + (i < extra ? 1 : 0);
        out.push_back({s, int32_t(l)});
        s += l;
    }
}

int64_t round_up(int64_t v, int64_t w) { return (v + w - 1) / w * w; }

// power-of-two scale so that max |v| / s <= 448
float fp8_scale(double maxabs) {
    if (!(maxabs > 0) || !std::isfinite(maxabs)) return 1.0f;
    int e = int(std::ceil(std::log2(maxabs / 448.0)));
    float s = std::ldexp(1.0f, e);
    while (maxabs / s > 448.0) s *= 2;     // guard against log2 rounding
    return s;
}

struct Seg {              // a block's contribution to one panel
    int64_t b;            // block
    int32_t pi;           // panel index inside the block's panel range
};

} // namespace

// ---------------------------------------------------------------------------
// layout
// ---------------------------------------------------------------------------

int build_layout(const hmvm_block* blocks, int64_t nb, int64_t nrows, int64_t ncols, const double* arena,
                 int64_t arena_len, const hmvm_plan_opts& opts, HostLayout& L) {
    if (nb < 0 || nrows < 0 || ncols < 0 || (nb > 0 && (!blocks || !arena))) return HMVM_ERR_INVALID_ARG;
    if (opts.precision < HMVM_PREC_FP64 || opts.precision > HMVM_PREC_AFL) return HMVM_ERR_INVALID_ARG;
    const bool packed = (opts.precision == HMVM_PREC_BFP || opts.precision == HMVM_PREC_AFL);
    const bool afl = (opts.precision == HMVM_PREC_AFL);
    const bool adaptive = (opts.precision == HMVM_PREC_VALR || packed);
    if (adaptive && !(opts.eps > 0)) return HMVM_ERR_INVALID_ARG;
    if (opts.force_format < 0 || opts.force_format > NFMT) return HMVM_ERR_INVALID_ARG;
    if (nrows > INT32_MAX || ncols > INT32_MAX) return HMVM_ERR_UNSUPPORTED;
    L = HostLayout();
    L.nrows = nrows;
    L.ncols = ncols;
    L.precision = opts.precision;
    L.eps = opts.eps;
    // dense leaves under VALR: as requested, or (dense_prec == VALR) by the
    // same rate rule as the columns, restricted to FP32 / FP64: FP32 while
    // eps needs at most its 23 mantissa bits.  Under BFP / AFL with
    // dense_prec == VALR every dense column gets the plan's packed-format
    // rule at rate(eps).
    if (adaptive && opts.dense_prec != HMVM_PREC_FP64 &&
        opts.dense_prec != HMVM_PREC_FP32 && opts.dense_prec != HMVM_PREC_VALR)
        return HMVM_ERR_INVALID_ARG;
    L.dense_fmt = (opts.precision == HMVM_PREC_FP64) ? F64
                : (opts.precision == HMVM_PREC_FP32) ? F32
                : (opts.dense_prec == HMVM_PREC_FP64) ? F64
                : (opts.dense_prec == HMVM_PREC_FP32) ? F32
                : (eps_to_rate(opts.eps) <= 23 ? F32 : F64);
    const bool dense_packed = (packed && opts.dense_prec == HMVM_PREC_VALR);
    L.blocks.assign(blocks, blocks + nb);

    // ---- validate leaves (cf. HLR_ASSERT in apply_add, check_data()) ----
    // [off, off + len) inside [0, total), compared in subtract form: offsets
    // come from files and may be anywhere in int64, so off + len could wrap
    auto inside = [](int64_t off, int64_t len, int64_t total) {
        return off >= 0 && len >= 0 && total >= 0 && len <= total && off <= total - len;
    };
    if (nrows < 0 || ncols < 0 || arena_len < 0) return HMVM_ERR_INVALID_ARG;
    for (int64_t i = 0; i < nb; ++i) {
        const hmvm_block& b = blocks[i];
        if (b.kind < HMVM_BLOCK_DENSE || b.kind > HMVM_BLOCK_LOWRANK_SV) return HMVM_ERR_BAD_BLOCK;
        if (!inside(b.row_ofs, b.nrows, nrows) || !inside(b.col_ofs, b.ncols, ncols)) return HMVM_ERR_BAD_BLOCK;
        const int64_t m = b.nrows, n = b.ncols;   // < 2^31 each: m * n and m * k cannot overflow
        if (b.kind == HMVM_BLOCK_DENSE) {
            if (m * n > 0 && !inside(b.off_a, m * n, arena_len)) return HMVM_ERR_BAD_BLOCK;
        } else {
            const int64_t k = b.rank;
            if (k < 0 || k > std::min(m, n)) return HMVM_ERR_BAD_BLOCK;
            if (k > 0) {
                if (!inside(b.off_a, m * k, arena_len)) return HMVM_ERR_BAD_BLOCK;
                if (!inside(b.off_b, n * k, arena_len)) return HMVM_ERR_BAD_BLOCK;
            }
            if (b.kind == HMVM_BLOCK_LOWRANK_SV && k > 0 && !inside(b.off_s, k, arena_len)) return HMVM_ERR_BAD_BLOCK;
            if (b.off_s >= 0 && k > 0 && !inside(b.off_s, k, arena_len)) return HMVM_ERR_BAD_BLOCK;
        }
    }
    if (opts.check_finite) {
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
        for (int64_t i = 0; i < nb; ++i) {
            const hmvm_block& b = blocks[i];
            auto scan = [&](int64_t off, int64_t len) {
                for (int64_t t = 0; t < len; ++t)
                    if (!std::isfinite(arena[off + t])) return 1;
                return 0;
            };
            if (b.kind == HMVM_BLOCK_DENSE) bad |= scan(b.off_a, int64_t(b.nrows) * b.ncols);
            else if (b.rank > 0) {
                bad |= scan(b.off_a, int64_t(b.nrows) * b.rank);
                bad |= scan(b.off_b, int64_t(b.ncols) * b.rank);
                if (b.kind == HMVM_BLOCK_LOWRANK_SV) bad |= scan(b.off_s, b.rank);
            }
        }
        if (bad) return HMVM_ERR_NONFINITE;
    }

    // ---- panels ----
    std::vector<int64_t> rb = {0, nrows}, cb = {0, ncols};
    for (int64_t i = 0; i < nb; ++i) {
        rb.push_back(blocks[i].row_ofs); rb.push_back(blocks[i].row_ofs + blocks[i].nrows);
        cb.push_back(blocks[i].col_ofs); cb.push_back(blocks[i].col_ofs + blocks[i].ncols);
    }
    std::sort(rb.begin(), rb.end()); rb.erase(std::unique(rb.begin(), rb.end()), rb.end());
    std::sort(cb.begin(), cb.end()); cb.erase(std::unique(cb.begin(), cb.end()), cb.end());
    std::vector<std::pair<int64_t, int32_t>> rpan, cpan;
    for (size_t i = 0; i + 1 < rb.size(); ++i) split_interval(rb[i], rb[i + 1], rpan);
    for (size_t i = 0; i + 1 < cb.size(); ++i) split_interval(cb[i], cb[i + 1], cpan);
    std::vector<int64_t> rstart(rpan.size() + 1), cstart(cpan.size() + 1);
    for (size_t i = 0; i < rpan.size(); ++i) rstart[i] = rpan[i].first;
    rstart[rpan.size()] = nrows;
    for (size_t i = 0; i < cpan.size(); ++i) cstart[i] = cpan[i].first;
    cstart[cpan.size()] = ncols;
    auto find_panel = [](const std::vector<int64_t>& st, int64_t v) {
        return int64_t(std::lower_bound(st.begin(), st.end(), v) - st.begin());
    };
    const int64_t NR = int64_t(rpan.size()), NC = int64_t(cpan.size());
    std::vector<int64_t> bp0(nb), bp1(nb), bq0(nb), bq1(nb);
    for (int64_t i = 0; i < nb; ++i) {
        const hmvm_block& b = blocks[i];
        bp0[i] = find_panel(rstart, b.row_ofs);
        bp1[i] = find_panel(rstart, b.row_ofs + b.nrows);
        bq0[i] = find_panel(cstart, b.col_ofs);
        bq1[i] = find_panel(cstart, b.col_ofs + b.ncols);
        if (b.nrows == 0 || b.ncols == 0) { bp1[i] = bp0[i]; bq1[i] = bq0[i]; }
    }
    // units of a low-rank block: its column panels in runs of at most
    // kUnitPanels from its first panel (at most kUnitRows rows); blocks with
    // the same column index set share their units, and their V segments of
    // the wide formats (FP64, FP32: W <= 4).  Columns of the narrow formats
    // (16-byte rows of 5 to 16 values) stay in one segment per column panel
    // and format across all blocks of the panel, so that a segment still
    // fills a wave's lanes.
    auto wide = [](int f) { return fmt_width(f) <= 4; };
    std::vector<int64_t> nunit(nb, 0);
    for (int64_t i = 0; i < nb; ++i)
        if (blocks[i].kind != HMVM_BLOCK_DENSE && blocks[i].rank > 0)
            nunit[i] = (bq1[i] - bq0[i] + kUnitPanels - 1) / kUnitPanels;
    auto unit_panels = [&](int64_t i, int64_t u) {
        const int64_t a = bq0[i] + u * kUnitPanels;
        return std::make_pair(a, std::min<int64_t>(a + kUnitPanels, bq1[i]));
    };

    // ---- t entries and their formats ----
    L.blk_t0.assign(nb, -1);
    int64_t ne = 0;
    for (int64_t i = 0; i < nb; ++i) {
        if (blocks[i].kind == HMVM_BLOCK_DENSE) { ++L.ndense; continue; }
        ++L.nlowrank;
        L.blk_t0[i] = ne;
        ne += blocks[i].rank;
    }
    if (ne > INT32_MAX / 2) return HMVM_ERR_UNSUPPORTED;
    L.ntent = ne;
    L.tfmt.assign(ne, int8_t(F64));
    L.tmul.assign(ne, 1.0);
    {
        std::vector<double> tol;
        for (int64_t i = 0; i < nb; ++i) {
            const hmvm_block& b = blocks[i];
            if (b.kind == HMVM_BLOCK_DENSE || b.rank == 0) continue;
            const int64_t e0 = L.blk_t0[i];
            if (b.kind == HMVM_BLOCK_LOWRANK_SV)
                for (int64_t l = 0; l < b.rank; ++l) L.tmul[e0 + l] = arena[b.off_s + l];
            int f = F64;
            if (opts.precision == HMVM_PREC_FP32) f = F32;
            for (int64_t l = 0; l < b.rank; ++l) L.tfmt[e0 + l] = int8_t(f);
            if (opts.precision == HMVM_PREC_VALR && b.off_s >= 0) {
                tol.resize(b.rank);
                valr_tolerances(b.rank, arena + b.off_s, opts.eps, tol.data());
                for (int64_t l = 0; l < b.rank; ++l) L.tfmt[e0 + l] = int8_t(format_for_rate(eps_to_rate(tol[l])));
            }
            if (opts.force_format > 0)
                for (int64_t l = 0; l < b.rank; ++l) L.tfmt[e0 + l] = int8_t(opts.force_format - 1);
        }
    }
    // block floating point / AFL: U and V column l of a block share one
    // width, the narrowest whose panel columns meet the 2-norm bound for
    // the column's VALR tolerance
    auto upieces = [&](int64_t i, int64_t l) {
        const hmvm_block& b = blocks[i];
        std::vector<Piece> v;
        for (int64_t p = bp0[i]; p < bp1[i]; ++p)
            v.push_back({arena + b.off_a + l * int64_t(b.nrows) + (rstart[p] - b.row_ofs), rpan[p].second});
        return v;
    };
    if (packed && opts.force_format == 0) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < nb; ++i) {
            const hmvm_block& b = blocks[i];
            if (b.kind == HMVM_BLOCK_DENSE || b.rank == 0 || b.off_s < 0) continue;
            std::vector<double> tol(b.rank);
            valr_tolerances(b.rank, arena + b.off_s, opts.eps, tol.data());
            for (int64_t l = 0; l < b.rank; ++l) {
                std::vector<std::vector<Piece>> cols(2);
                cols[0] = upieces(i, l);
                for (int64_t q = bq0[i]; q < bq1[i]; ++q)
                    cols[1].push_back({arena + b.off_b + l * int64_t(b.ncols) + (cstart[q] - b.col_ofs),
                                       cpan[q].second});
                L.tfmt[L.blk_t0[i] + l] = int8_t(packed_choose(cols, tol[l], eps_to_rate(tol[l]), afl));
            }
        }
    }
    for (int64_t e = 0; e < ne; ++e) L.cols_by_format[L.tfmt[e]]++;
    // dense columns: one format per column of each dense leaf
    L.blk_d0.assign(nb, -1);
    {
        int64_t nd = 0;
        for (int64_t i = 0; i < nb; ++i)
            if (blocks[i].kind == HMVM_BLOCK_DENSE) { L.blk_d0[i] = nd; nd += blocks[i].ncols; }
        L.dfmt.assign(nd, int8_t(L.dense_fmt));
        if (dense_packed) {
            // a dense column has no VALR tolerance of its own: eps's rate plus
            // one bit, the bound of a rounding format with rate(eps) bits
            const int rate = eps_to_rate(opts.eps);
            const double dtol = std::ldexp(1.0, -(rate + 1));
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t i = 0; i < nb; ++i) {
                if (blocks[i].kind != HMVM_BLOCK_DENSE) continue;
                for (int64_t j = 0; j < blocks[i].ncols; ++j) {
                    std::vector<std::vector<Piece>> cols(1, upieces(i, j));
                    L.dfmt[L.blk_d0[i] + j] = int8_t(packed_choose(cols, dtol, rate, afl));
                }
            }
        }
    }

    // ---- overlap check: per row panel, column intervals must be disjoint ----
    std::vector<std::vector<Seg>> rseg(NR), cseg(NC);
    for (int64_t i = 0; i < nb; ++i) {
        for (int64_t p = bp0[i]; p < bp1[i]; ++p) rseg[p].push_back({i, int32_t(p - bp0[i])});
        if (nunit[i] > 0)
            for (int64_t q = bq0[i]; q < bq1[i]; ++q) cseg[q].push_back({i, int32_t(q - bq0[i])});
    }
    {
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
        for (int64_t p = 0; p < NR; ++p) {
            std::vector<std::pair<int64_t, int64_t>> iv;
            for (const Seg& s : rseg[p]) {
                const hmvm_block& b = blocks[s.b];
                if (b.ncols > 0) iv.push_back({b.col_ofs, b.col_ofs + b.ncols});
            }
            std::sort(iv.begin(), iv.end());
            for (size_t t = 1; t < iv.size(); ++t)
                if (iv[t].first < iv[t - 1].second) bad = 1;
        }
        if (bad) return HMVM_ERR_OVERLAP;
    }

    // column index sets of the low-rank blocks; a set's columns of a wide
    // format go to tall segments (one per unit) when they fill at least half
    // a wave (kItemGroups / 2 groups), else to the per-panel segments
    // column index sets of the low-rank blocks, in order of first column
    struct ColSet {
        int64_t b0;                    // a block of the set (its panels and units)
        std::vector<int64_t> blocks;   // blocks of the set, in block order
        int64_t unit0 = 0;             // first unit of the set in the unit table
    };
    std::vector<ColSet> csets;
    {
        std::map<std::pair<int64_t, int64_t>, int64_t> id;
        for (int64_t i = 0; i < nb; ++i) {
            if (nunit[i] == 0) continue;
            auto key = std::make_pair(int64_t(blocks[i].col_ofs), int64_t(blocks[i].ncols));
            auto it = id.find(key);
            if (it == id.end()) it = id.emplace(key, int64_t(-1)).first;
            (void)it;
        }
        int64_t k = 0;
        for (auto& kv : id) kv.second = k++;
        csets.resize(size_t(k));
        for (int64_t i = 0; i < nb; ++i) {
            if (nunit[i] == 0) continue;
            ColSet& cs = csets[size_t(id[std::make_pair(int64_t(blocks[i].col_ofs), int64_t(blocks[i].ncols))])];
            if (cs.blocks.empty()) cs.b0 = i;
            cs.blocks.push_back(i);
        }
    }
    std::vector<int64_t> blk_set(nb, -1);
    std::vector<std::array<char, NFMT>> settall(csets.size());
    for (size_t c = 0; c < csets.size(); ++c) {
        int64_t ncol[NFMT] = {};
        for (int64_t i : csets[c].blocks) {
            blk_set[i] = int64_t(c);
            for (int64_t l = 0; l < blocks[i].rank; ++l) ncol[L.tfmt[L.blk_t0[i] + l]]++;
        }
        for (int f = 0; f < NFMT; ++f)
            settall[c][f] = wide(f) && round_up(ncol[f], fmt_width(f)) / fmt_width(f) >= kItemGroups / 2;
    }
    auto tall = [&](int64_t i, int f) { return blk_set[i] >= 0 && settall[size_t(blk_set[i])][f] != 0; };

    // ---- forward partials P and transposed partials Q, block-major:
    //      entry (panel index r, column l) of block b at base_b + r * k_b + l,
    //      so the sums over r read consecutive l with unit stride ----
    L.pt_off.assign(ne, 0); L.pt_cnt.assign(ne, 0);
    L.qt_off.assign(ne, 0); L.qt_cnt.assign(ne, 0);
    L.tstride.assign(ne, 0);
    {
        int64_t pos = 0, qpos = 0;
        for (int64_t i = 0; i < nb; ++i) {
            if (L.blk_t0[i] < 0) continue;
            const int64_t k = blocks[i].rank;
            for (int64_t l = 0; l < k; ++l) {
                const int64_t e = L.blk_t0[i] + l;
                L.pt_off[e] = int32_t(pos + l);
                L.pt_cnt[e] = int32_t(tall(i, L.tfmt[e]) ? nunit[i] : bq1[i] - bq0[i]);
                L.qt_off[e] = int32_t(qpos + l); L.qt_cnt[e] = int32_t(bp1[i] - bp0[i]);
                L.tstride[e] = int32_t(k);
            }
            pos += k * (bq1[i] - bq0[i]);
            qpos += k * (bp1[i] - bp0[i]);
        }
        L.psize = pos;
        // lanes of the one-kernel sums: an entry's lanes never straddle a wave
        auto lanes = [&](const std::vector<int32_t>& off, const std::vector<int32_t>& cnt, std::vector<SumLane>& le,
                         std::vector<int32_t>& lil) {
            le.clear(); lil.clear();
            const SumLane pad{0, 0, 0, -1};
            for (int64_t e = 0; e < ne; ++e) {
                const int64_t nseg = (int64_t(cnt[e]) + kSegLen - 1) / kSegLen;
                const int32_t Le = int32_t(std::min<int64_t>(std::max<int64_t>(nseg, 1), 64));
                if (int64_t(le.size() % 64) + Le > 64)
                    while (le.size() % 64) { le.push_back(pad); lil.push_back(0); }
                for (int32_t i = 0; i < Le; ++i) {
                    le.push_back({off[e], cnt[e], L.tstride[e], int32_t(e)});
                    lil.push_back(i | (Le << 8));
                }
            }
            while (le.size() % 64) { le.push_back(pad); lil.push_back(0); }
        };
        lanes(L.pt_off, L.pt_cnt, L.pl, L.pl_il);
        lanes(L.qt_off, L.qt_cnt, L.ql, L.ql_il);
        // dense transposed partials: per global column, one per (dense block, row panel)
        L.qd_cnt.assign(ncols, 0);
        L.qd_off.assign(ncols, 0);
        for (int64_t i = 0; i < nb; ++i) {
            if (blocks[i].kind != HMVM_BLOCK_DENSE) continue;
            for (int64_t j = 0; j < blocks[i].ncols; ++j) L.qd_cnt[blocks[i].col_ofs + j] += int32_t(bp1[i] - bp0[i]);
        }
        for (int64_t j = 0; j < ncols; ++j) { L.qd_off[j] = int32_t(qpos); qpos += L.qd_cnt[j]; }
        L.qsize = qpos;
        if (L.qsize > INT32_MAX / 2 || L.psize > INT32_MAX / 2) return HMVM_ERR_UNSUPPORTED;
    }
    // per (dense block, row panel, column) destination of its transposed
    // partial, assigned in block order then panel order
    std::vector<int64_t> dstart(nb, -1);    // index into ddest of the block's first entry
    std::vector<int32_t> ddest;
    {
        std::vector<int32_t> cur(ncols, 0);
        int64_t cnt = 0;
        for (int64_t i = 0; i < nb; ++i)
            if (blocks[i].kind == HMVM_BLOCK_DENSE) cnt += (bp1[i] - bp0[i]) * int64_t(blocks[i].ncols);
        ddest.resize(cnt);
        int64_t pos = 0;
        for (int64_t i = 0; i < nb; ++i) {
            if (blocks[i].kind != HMVM_BLOCK_DENSE) continue;
            dstart[i] = pos;
            for (int64_t p = bp0[i]; p < bp1[i]; ++p)
                for (int64_t j = 0; j < blocks[i].ncols; ++j) {
                    const int64_t g = blocks[i].col_ofs + j;
                    ddest[pos++] = L.qd_off[g] + cur[g]++;
                }
        }
    }

    // ---- row panels: column counts per format, chunks ----
    L.rp.resize(NR);
    std::vector<std::array<int64_t, NFMT>> rcols(NR);
    for (int64_t p = 0; p < NR; ++p) {
        std::array<int64_t, NFMT> c{};
        for (const Seg& s : rseg[p]) {
            const hmvm_block& b = blocks[s.b];
            if (b.kind == HMVM_BLOCK_DENSE)
                for (int64_t j = 0; j < b.ncols; ++j) c[L.dfmt[L.blk_d0[s.b] + j]]++;
            else for (int64_t l = 0; l < b.rank; ++l) c[L.tfmt[L.blk_t0[s.b] + l]]++;
        }
        rcols[p] = c;
    }
    // chunk table and offsets
    int64_t arena_pos[NFMT] = {};
    int64_t colpos = 0;
    std::vector<std::array<int64_t, NFMT>> rcol0(NR);        // first per-column index of format f in panel p
    for (int64_t p = 0; p < NR; ++p) {
        RowPanel& P = L.rp[p];
        P.row0 = rpan[p].first;
        P.m = rpan[p].second;
        P.chunk0 = int64_t(L.rchunks.size());
        for (int f = 0; f < NFMT; ++f) {
            const int W = fmt_width(f);
            const int64_t kp = round_up(rcols[p][f], W);
            rcol0[p][f] = colpos;
            const int64_t ng = kp / W;
            for (int64_t g0 = 0; g0 < ng; g0 += kChunkGroups) {
                RChunk c;
                c.fmt = f;
                c.ngroups = int32_t(std::min<int64_t>(kChunkGroups, ng - g0));
                c.data_off = arena_pos[f] + g0 * P.m * W;
                c.col0 = colpos + g0 * W;
                c.panel = int32_t(p);
                c.zoff = 0;
                L.rchunks.push_back(c);
            }
            arena_pos[f] += kp * P.m;
            colpos += kp;
        }
        P.nchunks = int32_t(int64_t(L.rchunks.size()) - P.chunk0);
        int64_t zb = 0;
        for (int64_t c = P.chunk0; c < P.chunk0 + P.nchunks; ++c) {
            RChunk& ch = L.rchunks[c];
            const int ab = (ch.fmt == F64) ? 8 : 4;        // accumulation type size
            zb = round_up(zb, 16);
            ch.zoff = int32_t(zb);
            zb += int64_t(ch.ngroups) * fmt_width(ch.fmt) * ab;
        }
        P.zbytes = int32_t(round_up(zb, 16));
        L.max_zbytes = std::max<int64_t>(L.max_zbytes, P.zbytes);
    }
    const int64_t nrcols = colpos;

    // ---- V segments: per column index set (its blocks), unit and format ----
    int64_t nunits = 0;
    for (ColSet& cs : csets) { cs.unit0 = nunits; nunits += nunit[cs.b0]; }
    // first segment of each (unit, format); -1 when the unit has none
    std::vector<std::array<int64_t, NFMT>> useg0(static_cast<size_t>(nunits));
    int64_t vpos = 0;
    int64_t varena_pos[NFMT];
    for (int f = 0; f < NFMT; ++f) varena_pos[f] = arena_pos[f];
    L.fitems.clear();
    for (const ColSet& cs : csets) {
        for (int64_t u = 0; u < nunit[cs.b0]; ++u) {
            const auto pq = unit_panels(cs.b0, u);
            const int64_t ux0 = cstart[pq.first], um = cstart[pq.second] - cstart[pq.first];
            int64_t ncol[NFMT] = {};
            for (int64_t i : cs.blocks)
                for (int64_t l = 0; l < blocks[i].rank; ++l) ncol[L.tfmt[L.blk_t0[i] + l]]++;
            for (int f = 0; f < NFMT; ++f) {
                useg0[size_t(cs.unit0 + u)][f] = -1;
                if (ncol[f] == 0 || !tall(cs.b0, f)) continue;
                const int W = fmt_width(f);
                const int64_t ng = round_up(ncol[f], W) / W;
                useg0[size_t(cs.unit0 + u)][f] = int64_t(L.fitems.size());
                // segments of kItemGroups groups, each a contiguous um x (ng W)
                // row-major block starting on a 128-byte boundary
                for (int64_t g = 0; g < ng; g += kItemGroups) {
                    VItem v;
                    v.q = int32_t(pq.first);
                    v.fmt = f;
                    v.ng = int32_t(std::min<int64_t>(kItemGroups, ng - g));
                    v.m = int32_t(um);
                    varena_pos[f] = round_up(varena_pos[f], 8 * int64_t(W));
                    v.data_off = varena_pos[f];
                    v.vc0 = vpos;
                    v.x0 = ux0;
                    varena_pos[f] += um * int64_t(v.ng) * W;
                    vpos += int64_t(v.ng) * W;
                    L.fitems.push_back(v);
                }
            }
        }
    }
    // narrow formats: one run of segments per column panel and format, over
    // all blocks of the panel
    std::vector<std::array<int64_t, NFMT>> qseg0(static_cast<size_t>(NC));
    for (int64_t q = 0; q < NC; ++q) {
        int64_t ncol[NFMT] = {};
        for (const Seg& sg : cseg[q])
            for (int64_t l = 0; l < blocks[sg.b].rank; ++l) {
                const int f = L.tfmt[L.blk_t0[sg.b] + l];
                if (!tall(sg.b, f)) ncol[f]++;
            }
        for (int f = 0; f < NFMT; ++f) {
            qseg0[size_t(q)][f] = -1;
            if (ncol[f] == 0) continue;
            const int W = fmt_width(f);
            const int64_t ng = round_up(ncol[f], W) / W;
            qseg0[size_t(q)][f] = int64_t(L.fitems.size());
            for (int64_t g = 0; g < ng; g += kItemGroups) {
                VItem v;
                v.q = int32_t(q);
                v.fmt = f;
                v.ng = int32_t(std::min<int64_t>(kItemGroups, ng - g));
                v.m = cpan[q].second;
                varena_pos[f] = round_up(varena_pos[f], 8 * int64_t(W));
                v.data_off = varena_pos[f];
                v.vc0 = vpos;
                v.x0 = cstart[q];
                varena_pos[f] += int64_t(v.m) * v.ng * W;
                vpos += int64_t(v.ng) * W;
                L.fitems.push_back(v);
            }
        }
    }
    // per column panel: views of the segments covering it
    L.cp.resize(NC);
    {
        std::vector<std::vector<VItem>> views(NC);
        for (const VItem& sg : L.fitems)
            for (int64_t q = sg.q; q < NC && cstart[q] < sg.x0 + sg.m; ++q) {
                VItem v = sg;
                v.q = int32_t(q);
                v.m = cpan[q].second;
                v.x0 = cstart[q];
                v.data_off = sg.data_off + (cstart[q] - sg.x0) * int64_t(sg.ng) * fmt_width(sg.fmt);
                views[q].push_back(v);
            }
        L.vitems.clear();
        for (int64_t q = 0; q < NC; ++q) {
            ColPanel& C = L.cp[q];
            C.col0 = cpan[q].first;
            C.m = cpan[q].second;
            C.item0 = int64_t(L.vitems.size());
            C.nitems = int32_t(views[q].size());
            L.vitems.insert(L.vitems.end(), views[q].begin(), views[q].end());
            L.max_vitem_rows = std::max<int32_t>(L.max_vitem_rows, C.m);
        }
    }
    // forward packs: segments of one format (by decreasing rows, so that the
    // lanes of a pack stream similar row counts), <= 64 groups and <= kPackXs
    // x values per pack
    for (int f = 0; f < NFMT; ++f) {
        VPack cur{};
        int64_t xused = 0;
        cur.fmt = f;
        auto flush = [&]() {
            if (cur.nseg > 0) L.vpacks.push_back(cur);
            cur = VPack{};
            cur.fmt = f;
            xused = 0;
        };
        std::vector<int64_t> order;
        for (int64_t j = 0; j < int64_t(L.fitems.size()); ++j)
            if (L.fitems[j].fmt == f) order.push_back(j);
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t a, int64_t b) { return L.fitems[a].m > L.fitems[b].m; });
        for (int64_t oj : order) {
            const VItem& v = L.fitems[oj];
            if (cur.nseg == kPackSegs || cur.gofs[cur.nseg] + v.ng > kItemGroups || xused + v.m + 1 > kPackXs) flush();
            const int j = cur.nseg;
            cur.gofs[j + 1] = cur.gofs[j] + v.ng;
            cur.m[j] = v.m;
            cur.ng[j] = v.ng;
            cur.xoff[j] = int32_t(xused);
            cur.x0[j] = int32_t(v.x0);
            cur.vc0[j] = v.vc0;
            cur.data_off[j] = v.data_off;
            xused += v.m + 1;          // odd distance between slices (LDS banks)
            ++cur.nseg;
        }
        flush();
    }
    const int64_t nvcols = vpos;
    if (nrcols > INT32_MAX || nvcols > INT32_MAX) return HMVM_ERR_UNSUPPORTED;

    // ---- allocate arenas and tables ----
    for (int f = 0; f < NFMT; ++f) L.nvals[f] = varena_pos[f];
    L.a64.assign(varena_pos[F64], 0.0);
    L.a32.assign(varena_pos[F32], 0.0f);
    L.a16.assign(varena_pos[BF16], 0);
    L.a8.assign(varena_pos[FP8], 0);
    for (int f = kFirstBfp; f < NFMT; ++f) L.abfp[f - kFirstBfp].assign(varena_pos[f] / fmt_width(f) * 4, 0u);
    L.colsrc.assign(nrcols, kPadCol);
    L.rdest.assign(nrcols, -1);
    L.vdest.assign(nvcols, -1);
    L.vsrc.assign(nvcols, -1);
    const bool need_scale = (opts.precision != HMVM_PREC_FP64) || (opts.force_format > 0);
    if (need_scale) {
        L.zscale.assign(nrcols, 1.0f);
        L.vscale.assign(nvcols, 1.0f);
    }

    // store m values src[0..m) of one panel column in format f at value
    // offsets o, o + ld, ..., o + (m-1) ld; returns the column's scale
    auto store_vals = [&](int f, int64_t o, int64_t ld, int64_t m, const double* src) -> float {
        switch (f) {
        case F64: for (int64_t i = 0; i < m; ++i) L.a64[o + i * ld] = src[i]; return 1.0f;
        case F32: for (int64_t i = 0; i < m; ++i) L.a32[o + i * ld] = float(src[i]); return 1.0f;
        case BF16: for (int64_t i = 0; i < m; ++i) L.a16[o + i * ld] = encode_bf16(float(src[i])); return 1.0f;
        case FP8: {
            double mx = 0;
            for (int64_t i = 0; i < m; ++i) mx = std::max(mx, std::fabs(src[i]));
            const float s = fp8_scale(mx);
            for (int64_t i = 0; i < m; ++i) L.a8[o + i * ld] = encode_e4m3(float(src[i] / double(s)));
            return s;
        }
        default: {
            const int nb = fmt_bits(f), W = fmt_width(f);
            uint32_t* words = L.abfp[f - kFirstBfp].data();
            if (fmt_is_afl(f)) {
                const float s = afl_scale(src, m);
                for (int64_t i = 0; i < m; ++i) {
                    const int64_t v = o + i * ld;
                    packed_put(words + (v / W) * 4, int(v % W), nb, afl_code(src[i], s, nb));
                }
                return s;
            }
            double mx = 0;
            for (int64_t i = 0; i < m; ++i) mx = std::max(mx, std::fabs(src[i]));
            const int k = bfp_exp(mx, nb);
            for (int64_t i = 0; i < m; ++i) {
                const int64_t v = o + i * ld;
                packed_put(words + (v / W) * 4, int(v % W), nb, uint32_t(bfp_quant(src[i], k, nb)));
            }
            return std::ldexp(1.0f, k);
        }
        }
    };
    // column c of a format group block starting at value offset `base` with
    // groups of W columns, each m x W row-major
    auto store_col = [&](int f, int64_t base, int64_t m, int64_t W, int64_t c, const double* src) {
        const int64_t g = c / W, w = c % W;
        return store_vals(f, base + g * m * W + w, W, m, src);
    };

    // ---- fill row panels (parallel over panels) ----
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = 0; p < NR; ++p) {
        const RowPanel& P = L.rp[p];
        const int64_t m = P.m;
        int64_t cnt[NFMT] = {};
        // value base of format f block of this panel = data_off of its first chunk
        int64_t fbase[NFMT];
        for (int f = 0; f < NFMT; ++f) fbase[f] = -1;
        for (int64_t c = P.chunk0; c < P.chunk0 + P.nchunks; ++c)
            if (fbase[L.rchunks[c].fmt] < 0) fbase[L.rchunks[c].fmt] = L.rchunks[c].data_off;
        std::vector<double> tmp(m);
        for (const Seg& s : rseg[p]) {
            const hmvm_block& b = blocks[s.b];
            const int64_t r0 = P.row0 - b.row_ofs;     // first row of the panel inside the block
            if (b.kind == HMVM_BLOCK_DENSE) {
                for (int64_t j = 0; j < b.ncols; ++j) {
                    const int f = L.dfmt[L.blk_d0[s.b] + j], W = fmt_width(f);
                    const int64_t c = cnt[f]++;
                    const double* src = arena + b.off_a + j * int64_t(b.nrows) + r0;
                    const float sc = store_col(f, fbase[f], m, W, c, src);
                    const int64_t gc = rcol0[p][f] + c;
                    L.colsrc[gc] = int32_t(b.col_ofs + j);
                    if (need_scale) L.zscale[gc] = sc;
                    L.rdest[gc] = ddest[dstart[s.b] + int64_t(s.pi) * b.ncols + j];
                }
            } else {
                const int64_t e0 = L.blk_t0[s.b];
                for (int64_t l = 0; l < b.rank; ++l) {
                    const int f = L.tfmt[e0 + l], W = fmt_width(f);
                    const int64_t c = cnt[f]++;
                    const double* src = arena + b.off_a + l * int64_t(b.nrows) + r0;
                    const float sc = store_col(f, fbase[f], m, W, c, src);
                    const int64_t gc = rcol0[p][f] + c;
                    L.colsrc[gc] = ~int32_t(e0 + l);
                    if (need_scale) L.zscale[gc] = sc;
                    L.rdest[gc] = L.qt_off[e0 + l] + s.pi * int32_t(b.rank);
                }
            }
        }
    }

    // ---- fill V segments (row-major, lane = group), parallel over column sets ----
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t ci = 0; ci < int64_t(csets.size()); ++ci) {
        const ColSet& cs = csets[size_t(ci)];
        for (int64_t u = 0; u < nunit[cs.b0]; ++u) {
            int64_t cnt[NFMT] = {};
            for (int64_t i : cs.blocks) {
                const hmvm_block& b = blocks[i];
                const int64_t e0 = L.blk_t0[i];
                for (int64_t l = 0; l < b.rank; ++l) {
                    const int f = L.tfmt[e0 + l];
                    if (!tall(i, f)) continue;
                    const int64_t per = int64_t(kItemGroups) * fmt_width(f);     // columns per segment
                    const int64_t c = cnt[f]++;
                    const VItem& sg = L.fitems[size_t(useg0[size_t(cs.unit0 + u)][f] + c / per)];
                    const int64_t cc = c % per;
                    const double* src = arena + b.off_b + l * int64_t(b.ncols) + (sg.x0 - b.col_ofs);
                    const float sc = store_vals(f, sg.data_off + cc, int64_t(sg.ng) * fmt_width(f), sg.m, src);
                    const int64_t vc = sg.vc0 + cc;
                    L.vdest[vc] = L.pt_off[e0 + l] + int32_t(u) * int32_t(b.rank);
                    L.vsrc[vc] = int32_t(e0 + l);
                    if (need_scale) L.vscale[vc] = sc;
                }
            }
        }
    }

#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t q = 0; q < NC; ++q) {
        int64_t cnt[NFMT] = {};
        const int64_t x0 = cstart[q], m = cpan[q].second;
        for (const Seg& sgm : cseg[q]) {
            const hmvm_block& b = blocks[sgm.b];
            const int64_t e0 = L.blk_t0[sgm.b];
            for (int64_t l = 0; l < b.rank; ++l) {
                const int f = L.tfmt[e0 + l];
                if (tall(sgm.b, f)) continue;
                const int64_t per = int64_t(kItemGroups) * fmt_width(f);
                const int64_t c = cnt[f]++;
                const VItem& sg = L.fitems[size_t(qseg0[size_t(q)][f] + c / per)];
                const int64_t cc = c % per;
                const double* src = arena + b.off_b + l * int64_t(b.ncols) + (x0 - b.col_ofs);
                const float sc = store_vals(f, sg.data_off + cc, int64_t(sg.ng) * fmt_width(f), m, src);
                const int64_t vc = sg.vc0 + cc;
                L.vdest[vc] = L.pt_off[e0 + l] + sgm.pi * int32_t(b.rank);
                L.vsrc[vc] = int32_t(e0 + l);
                if (need_scale) L.vscale[vc] = sc;
            }
        }
    }

    // ---- one-level (TLR) detection: row and column intervals do not nest ----
    {
        auto disjoint = [&](bool rows) {
            std::vector<std::pair<int64_t, int64_t>> iv;
            for (int64_t i = 0; i < nb; ++i) {
                const hmvm_block& b = blocks[i];
                if (rows) iv.push_back({b.row_ofs, b.row_ofs + b.nrows});
                else iv.push_back({b.col_ofs, b.col_ofs + b.ncols});
            }
            std::sort(iv.begin(), iv.end());
            iv.erase(std::unique(iv.begin(), iv.end()), iv.end());
            for (size_t t = 1; t < iv.size(); ++t)
                if (iv[t].first < iv[t - 1].second) return false;
            return true;
        };
        L.one_level = disjoint(true) && disjoint(false);
    }

    // ---- statistics ----
    for (int f = 0; f < NFMT; ++f) L.bytes_by_format[f] = fmt_store_bytes(f, L.nvals[f]);
    {
        bool hw = false, bfp = false, aflf = false;
        for (int f = 0; f < NFMT; ++f) {
            if (L.nvals[f] == 0) continue;
            hw |= (f == BF16 || f == FP8);
            bfp |= fmt_is_bfp(f);
            aflf |= fmt_is_afl(f);
        }
        // e.g. BF16 forced beside block floating point dense leaves
        if (int(hw) + int(bfp) + int(aflf) > 1) return HMVM_ERR_INVALID_ARG;
        L.family = bfp ? kFamBfp : aflf ? kFamAfl : kFamHw;
    }
    // bytes per stored value (16 bytes per row of W values)
    auto vbytes = [](int f) { return 16.0 / double(fmt_width(f)); };
    double fl = 0, ds = 0, br = 0, bv = 0;
    for (int64_t i = 0; i < nb; ++i) {
        const hmvm_block& b = blocks[i];
        const double m = b.nrows, n = b.ncols;
        if (b.kind == HMVM_BLOCK_DENSE) {
            fl += 2 * m * n + 2 * m;
            for (int64_t j = 0; j < b.ncols; ++j) {
                const double by = vbytes(L.dfmt[L.blk_d0[i] + j]) * m;
                ds += by;
                br += by;
            }
        } else {
            const double k = b.rank;
            fl += 2 * k * (m + n) + 2 * (m + n);
            if (b.kind == HMVM_BLOCK_LOWRANK_SV) ds += 8.0 * k;   // S (matrix/lrsvmatrix.hh:364-370)
            for (int64_t l = 0; l < b.rank; ++l) {
                const int f = L.tfmt[L.blk_t0[i] + l];
                ds += vbytes(f) * (m + n);
                br += vbytes(f) * m;
                bv += vbytes(f) * n;
            }
        }
    }
    // the scale of an FP8 or block floating point panel column counts as
    // data (4 bytes per column of such a chunk or segment, which the kernels
    // read; other columns' table entries are never read)
    int64_t s8r = 0, s8v = 0;
    for (const RChunk& c : L.rchunks)
        if (fmt_scaled(c.fmt)) s8r += int64_t(c.ngroups) * fmt_width(c.fmt);
    for (const VItem& v : L.fitems)
        if (fmt_scaled(v.fmt)) s8v += int64_t(v.ng) * fmt_width(v.fmt);
    L.scale_bytes = 4 * (s8r + s8v);
    br += 4.0 * double(s8r);
    bv += 4.0 * double(s8v);
    ds += double(L.scale_bytes);
    L.bytes_rpass = br;
    L.bytes_vpass = bv;
    L.flops = fl;
    L.datasize = ds;
    return HMVM_OK;
}

// ---------------------------------------------------------------------------
// unpack (host copies of the arenas)
// ---------------------------------------------------------------------------

template <int NB>
static int32_t bfp_get(const uint32_t* words, int64_t v) {
    constexpr int W = 128 / NB;
    const uint32_t* r = words + (v / W) * 4;
    const uint32_t w[4] = {r[0], r[1], r[2], r[3]};
    return bfp_field<NB>(w, int(v % W));
}

template <int NB>
static double afl_get(const uint32_t* words, int64_t v) {
    constexpr int W = 128 / NB;
    const uint32_t* r = words + (v / W) * 4;
    const uint32_t w[4] = {r[0], r[1], r[2], r[3]};
    return double(afl_unscaled(afl_field<NB>(w, int(v % W)), NB));
}

// stored value idx of format f (without its panel-column scale)
static double load_elem(int f, const void* const arenas[NFMT], int64_t idx, bool as_f64) {
    if (as_f64) return static_cast<const double*>(arenas[f])[idx];
    const uint32_t* w = static_cast<const uint32_t*>(arenas[f]);
    switch (f) {
    case F64: return static_cast<const double*>(arenas[F64])[idx];
    case F32: return static_cast<const float*>(arenas[F32])[idx];
    case BF16: return decode_bf16(static_cast<const uint16_t*>(arenas[BF16])[idx]);
    case FP8: return decode_e4m3(static_cast<const uint8_t*>(arenas[FP8])[idx]);
    case B8: return bfp_get<8>(w, idx);
    case B16: return bfp_get<16>(w, idx);
    case B25: return bfp_get<25>(w, idx);
    case A12: return afl_get<12>(w, idx);
    case A16: return afl_get<16>(w, idx);
    default: return afl_get<24>(w, idx);
    }
}

int unpack_block(const HostLayout& L, int64_t bi, const void* const arenas[NFMT], bool as_f64, double* a,
                 double* bb) {
    if (bi < 0 || bi >= int64_t(L.blocks.size()) || !a) return HMVM_ERR_INVALID_ARG;
    const hmvm_block& b = L.blocks[bi];
    const bool dense = (b.kind == HMVM_BLOCK_DENSE);
    if (!dense && !bb) return HMVM_ERR_INVALID_ARG;
    for (int f = 0; f < NFMT; ++f)
        if (L.bytes_by_format[f] > 0 && !arenas[f]) return HMVM_ERR_INVALID_ARG;
    const int64_t e0 = L.blk_t0[bi];
    // R side: walk the row panels the block covers
    for (size_t p = 0; p < L.rp.size(); ++p) {
        const RowPanel& P = L.rp[p];
        if (P.row0 < b.row_ofs || P.row0 >= b.row_ofs + b.nrows) continue;
        for (int64_t ci = P.chunk0; ci < P.chunk0 + P.nchunks; ++ci) {
            const RChunk& C = L.rchunks[ci];
            const int W = fmt_width(C.fmt);
            for (int64_t c = 0; c < int64_t(C.ngroups) * W; ++c) {
                const int32_t src = L.colsrc[C.col0 + c];
                int64_t col = -1;
                if (dense && src >= 0 && src >= b.col_ofs && src < b.col_ofs + b.ncols) col = src - b.col_ofs;
                if (!dense && src < 0 && src != kPadCol) {
                    const int64_t e = ~int64_t(src);
                    if (e >= e0 && e < e0 + b.rank) col = e - e0;
                }
                if (col < 0) continue;
                const double sc = (as_f64 || L.zscale.empty()) ? 1.0 : L.zscale[C.col0 + c];
                const int64_t g = c / W, w = c % W;
                for (int64_t i = 0; i < P.m; ++i) {
                    const double v = load_elem(C.fmt, arenas, C.data_off + g * P.m * W + i * W + w, as_f64) * sc;
                    a[col * b.nrows + (P.row0 - b.row_ofs) + i] = v;
                }
            }
        }
    }
    if (dense) return HMVM_OK;
    for (size_t q = 0; q < L.cp.size(); ++q) {
        const ColPanel& C = L.cp[q];
        if (C.col0 < b.col_ofs || C.col0 >= b.col_ofs + b.ncols) continue;
        for (int64_t j = C.item0; j < C.item0 + C.nitems; ++j) {
            const VItem& v = L.vitems[j];
            const int f = v.fmt, W = fmt_width(f);
            const int64_t ld = int64_t(v.ng) * W;
            for (int64_t cc = 0; cc < ld; ++cc) {
                const int64_t vc = v.vc0 + cc;
                const int32_t e = L.vsrc[vc];
                if (e < e0 || e >= e0 + b.rank) continue;
                const int64_t l = e - e0;
                const double sc = (as_f64 || L.vscale.empty()) ? 1.0 : L.vscale[vc];
                for (int64_t i = 0; i < C.m; ++i)
                    bb[l * b.ncols + (C.col0 - b.col_ofs) + i] = load_elem(f, arenas, v.data_off + i * ld + cc, as_f64) * sc;
            }
        }
    }
    return HMVM_OK;
}

} // namespace hmvm

namespace hmvm {

// For every stored value of format f: its offset in the FP64 arena that
// hmvm_plan_create was given (the leaves' own column-major layout), or -1
// for padding.  One walk over the R chunks and V segments.
int arena_source(const HostLayout& L, int f, int64_t* src) {
    if (f < 0 || f >= NFMT || (L.nvals[f] > 0 && !src)) return HMVM_ERR_INVALID_ARG;
    const int64_t nv = L.nvals[f];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nv; ++k) src[k] = -1;
    if (nv == 0) return HMVM_OK;
    const int64_t nb = int64_t(L.blocks.size());
    // t-entry -> block
    std::vector<int64_t> ent_block(static_cast<size_t>(L.ntent), -1);
    for (int64_t b = 0; b < nb; ++b)
        if (L.blk_t0[b] >= 0)
            for (int64_t l = 0; l < L.blocks[b].rank; ++l) ent_block[size_t(L.blk_t0[b] + l)] = b;
    // dense blocks of each row panel, by first column
    const int64_t NR = int64_t(L.rp.size());
    std::vector<int64_t> prow0(static_cast<size_t>(NR));
    for (int64_t p = 0; p < NR; ++p) prow0[size_t(p)] = L.rp[p].row0;
    std::vector<std::vector<std::pair<int64_t, int64_t>>> pdense(static_cast<size_t>(NR));
    for (int64_t b = 0; b < nb; ++b) {
        const hmvm_block& B = L.blocks[b];
        if (B.kind != HMVM_BLOCK_DENSE) continue;
        auto it = std::lower_bound(prow0.begin(), prow0.end(), B.row_ofs);
        for (; it != prow0.end() && *it < B.row_ofs + B.nrows; ++it) pdense[size_t(it - prow0.begin())].push_back({B.col_ofs, b});
    }
    for (auto& v : pdense) std::sort(v.begin(), v.end());
    const int W = fmt_width(f);
    // R side (row panels write disjoint runs)
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : bad)
    for (int64_t p = 0; p < NR; ++p) {
        const RowPanel& P = L.rp[p];
        const auto& dl = pdense[size_t(p)];
        for (int64_t ci = P.chunk0; ci < P.chunk0 + P.nchunks; ++ci) {
            const RChunk& C = L.rchunks[ci];
            if (C.fmt != f) continue;
            for (int64_t c = 0; c < int64_t(C.ngroups) * W; ++c) {
                const int32_t s = L.colsrc[C.col0 + c];
                if (s == kPadCol) continue;
                int64_t b = -1, col = -1;
                if (s >= 0) {
                    auto it = std::upper_bound(dl.begin(), dl.end(), std::make_pair(int64_t(s), INT64_MAX));
                    if (it == dl.begin()) { bad = 1; continue; }
                    --it;
                    b = it->second;
                    col = s - L.blocks[b].col_ofs;
                    if (col >= L.blocks[b].ncols) { bad = 1; continue; }
                } else {
                    const int64_t e = ~int64_t(s);
                    b = ent_block[size_t(e)];
                    col = e - L.blk_t0[b];
                }
                const hmvm_block& B = L.blocks[b];
                const int64_t g = c / W, w = c % W;
                for (int64_t i = 0; i < P.m; ++i)
                    src[C.data_off + g * P.m * W + i * W + w] = B.off_a + col * B.nrows + (P.row0 - B.row_ofs) + i;
            }
        }
    }
    if (bad) return HMVM_ERR_KERNEL;
    // V side (column panels write disjoint segments)
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < int64_t(L.cp.size()); ++q) {
        const ColPanel& C = L.cp[q];
        for (int64_t j = C.item0; j < C.item0 + C.nitems; ++j) {
            const VItem& v = L.vitems[j];
            if (v.fmt != f) continue;
            const int64_t ld = int64_t(v.ng) * W;
            for (int64_t cc = 0; cc < ld; ++cc) {
                const int32_t e = L.vsrc[v.vc0 + cc];
                if (e < 0) continue;
                const int64_t b = ent_block[size_t(e)];
                const hmvm_block& B = L.blocks[b];
                const int64_t l = e - L.blk_t0[b];
                for (int64_t i = 0; i < C.m; ++i)
                    src[v.data_off + i * ld + cc] = B.off_b + l * B.ncols + (C.col0 - B.col_ofs) + i;
            }
        }
    }
    return HMVM_OK;
}

} // namespace hmvm
