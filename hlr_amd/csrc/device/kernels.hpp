// kernels.hpp -- CDNA4 (gfx950) kernels of the H-/TLR-matrix-vector product.
//
// One wavefront = 64 lanes.  Factor data is read with one 16-byte load per
// lane (W = 16 / sizeof(element) values), decoded to the accumulation type,
// and never re-read: the forward product touches every stored byte once.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "plan.hpp"

namespace hmvm {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// device view of a plan (passed by value)
struct DevPlan {
    const RowPanel* rp;
    const RChunk* rchunks;
    const ColPanel* cp;
    const VItem* vitems;      // panel views of the V segments
    const VItem* fitems;      // V segments
    const int32_t* colsrc;
    const float* zscale;      // null: all 1
    const int32_t* rdest;
    const int32_t* vdest;
    const int32_t* vsrc;
    const float* vscale;      // null: all 1
    const double* tmul;
    const int32_t* pt_off;
    const int32_t* pt_cnt;
    const int32_t* qt_off;
    const int32_t* qt_cnt;
    const int32_t* tstride;
    const VPack* vpacks;
    const SumLane* pl;        // one-kernel sums: lanes per t-entry (see the plan builder)
    const int32_t* pl_il;
    const SumLane* ql;
    const int32_t* ql_il;
    int64_t npl, nql;
    int64_t nvpacks;
    const int32_t* qd_off;
    const int32_t* qd_cnt;
    const void* arena[NFMT];
    int64_t nrp, ncp, nvitems, nfitems, nrchunks, ntent, psize, qsize, nrows, ncols;
};

// ---------------------------------------------------------------------------
// 16-byte loads of W elements, decoded
// ---------------------------------------------------------------------------

template <typename T> struct Elem;

// raw(): the 16-byte non-temporal load alone; decode(): its W values in the
// accumulation type; load() = decode(raw()).  Splitting the two lets a loop
// keep several raw loads in flight while it decodes earlier ones.
__device__ inline u32x4 raw16(const void* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
}

// Factor rows are addressed by value offsets (multiples of W): the row of
// offset `off` starts at byte (off / W) * 16 of its format's arena.
using bytep = const unsigned char*;

template <> struct Elem<double> {
    static constexpr int F = F64;
    static constexpr int W = 2;
    static constexpr bool kScaled = false;
    using Acc = double;
    __device__ static inline void decode(u32x4 v, double* out) {
        out[0] = __hiloint2double(int(v.y), int(v.x));
        out[1] = __hiloint2double(int(v.w), int(v.z));
    }
    __device__ static inline void load(bytep p, double* out) {
        const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p));
        out[0] = v.x; out[1] = v.y;
    }
};

template <> struct Elem<float> {
    static constexpr int F = F32;
    static constexpr int W = 4;
    static constexpr bool kScaled = false;
    using Acc = float;
    __device__ static inline void decode(u32x4 v, float* out) {
        out[0] = __uint_as_float(v.x); out[1] = __uint_as_float(v.y);
        out[2] = __uint_as_float(v.z); out[3] = __uint_as_float(v.w);
    }
    __device__ static inline void load(bytep p, float* out) {
        const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
        out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
    }
};

template <> struct Elem<uint16_t> {   // BF16
    static constexpr int F = BF16;
    static constexpr int W = 8;
    static constexpr bool kScaled = false;
    using Acc = float;
    __device__ static inline void decode(u32x4 v, float* out) {
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            out[2 * i] = __uint_as_float(w[i] << 16);
            out[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
        }
    }
    __device__ static inline void load(bytep p, float* out) { decode(raw16(p), out); }
};

template <> struct Elem<uint8_t> {    // FP8 E4M3 (OCP)
    static constexpr int F = FP8;
    static constexpr int W = 16;
    static constexpr bool kScaled = true;
    using Acc = float;
    __device__ static inline void decode(u32x4 v, float* out) {
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(int(w[i]), false);
            const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(int(w[i]), true);
            out[4 * i] = lo.x; out[4 * i + 1] = lo.y; out[4 * i + 2] = hi.x; out[4 * i + 3] = hi.y;
        }
    }
    __device__ static inline void load(bytep p, float* out) { decode(raw16(p), out); }
};

// block floating point with NB-bit values (tag type; the arena holds packed rows)
template <int NB> struct Bfp {};
constexpr int bfp_format(int nb) {
    return nb == 8 ? B8 : nb == 16 ? B16 : nb == 25 ? B25 : -1;
}

template <int NB> struct Elem<Bfp<NB>> {
    static constexpr int F = bfp_format(NB);
    static_assert(F >= kFirstBfp, "no storage format with this width");
    static constexpr int W = 128 / NB;
    static constexpr bool kScaled = true;
    using Acc = float;
    // |q| < 2^24, so the conversion is exact
    __device__ static inline void decode(u32x4 v, float* out) {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < W; ++j) out[j] = float(bfp_field<NB>(w, j));
    }
    __device__ static inline void load(bytep p, float* out) { decode(raw16(p), out); }
};

// adaptive float with NB-bit codes (tag type; the arena holds packed rows)
template <int NB> struct Afl {};
constexpr int afl_format(int nb) {
    return nb == 12 ? A12 : nb == 16 ? A16 : nb == 24 ? A24 : -1;
}

template <int NB> struct Elem<Afl<NB>> {
    static constexpr int F = afl_format(NB);
    static_assert(F >= kFirstAfl, "no storage format with this width");
    static constexpr int W = 128 / NB;
    static constexpr bool kScaled = true;
    using Acc = float;
    static constexpr int M = NB - 1 - kAflExpBits;
    // +-(a - 1) with a = 2^(e + 1) (1 + mant 2^-M) built from its FP32 bits
    // (biased exponent 128 + e), the subtraction in FP32; the all-ones code
    // is zero.  The column scale rides on z (or on the partial sums).
    __device__ static inline void decode(u32x4 v, float* out) {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        constexpr uint32_t kZero = (1u << NB) - 1u;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const uint32_t c = afl_field<NB>(w, j);
            const uint32_t mant = c & ((1u << M) - 1u);
            const uint32_t e = (c >> M) & ((1u << kAflExpBits) - 1u);
            const float a = __uint_as_float(((e | 0x80u) << 23) | (mant << (23 - M)));
            const float d = a - 1.0f;
            out[j] = (c == kZero) ? 0.0f : ((c >> (NB - 1)) ? -d : d);
        }
    }
    __device__ static inline void load(bytep p, float* out) { decode(raw16(p), out); }
};

// start of the row holding value offset `off` (a multiple of W) in an arena
template <typename T>
__device__ inline bytep elem_addr(const void* arena, int64_t off) {
    return static_cast<bytep>(arena) + (off / Elem<T>::W) * 16;
}

// Like HMVM_FMT_SWITCH, restricted to family FAM (a compile-time constant);
// formats outside it do nothing.
#define HMVM_FMT_SWITCH_FAM(FAM, f, ...)                                                  \
    switch (f) {                                                                          \
    case F64: { using T = double; __VA_ARGS__; } break;                                   \
    case F32: if constexpr (FAM != kFamF64) { using T = float; __VA_ARGS__; } break;      \
    case BF16: if constexpr (FAM == kFamAll || FAM == kFamHw) { using T = uint16_t; __VA_ARGS__; } break; \
    case FP8: if constexpr (FAM == kFamAll || FAM == kFamHw) { using T = uint8_t; __VA_ARGS__; } break;   \
    case B8: if constexpr (FAM == kFamAll || FAM == kFamBfp) { using T = Bfp<8>; __VA_ARGS__; } break;    \
    case B16: if constexpr (FAM == kFamAll || FAM == kFamBfp) { using T = Bfp<16>; __VA_ARGS__; } break;  \
    case B25: if constexpr (FAM == kFamAll || FAM == kFamBfp) { using T = Bfp<25>; __VA_ARGS__; } break;  \
    case A12: if constexpr (FAM == kFamAll || FAM == kFamAfl) { using T = Afl<12>; __VA_ARGS__; } break;  \
    case A16: if constexpr (FAM == kFamAll || FAM == kFamAfl) { using T = Afl<16>; __VA_ARGS__; } break;  \
    default: if constexpr (FAM == kFamAll || FAM == kFamAfl) { using T = Afl<24>; __VA_ARGS__; } break;   \
    }

// Runs the statements with T bound to the element type of storage format f.
#define HMVM_FMT_SWITCH(f, ...)                                      \
    switch (f) {                                                     \
    case F64: { using T = double; __VA_ARGS__; } break;              \
    case F32: { using T = float; __VA_ARGS__; } break;               \
    case BF16: { using T = uint16_t; __VA_ARGS__; } break;           \
    case FP8: { using T = uint8_t; __VA_ARGS__; } break;             \
    case B8: { using T = Bfp<8>; __VA_ARGS__; } break;               \
    case B16: { using T = Bfp<16>; __VA_ARGS__; } break;             \
    case B25: { using T = Bfp<25>; __VA_ARGS__; } break;             \
    case A12: { using T = Afl<12>; __VA_ARGS__; } break;             \
    case A16: { using T = Afl<16>; __VA_ARGS__; } break;             \
    default: { using T = Afl<24>; __VA_ARGS__; } break;              \
    }

// Workgroup i of a launch is dispatched to XCD i % 8.  xcd_item(i, n, B)
// reorders items so that each XCD runs consecutive items in runs of B (run
// r goes to XCD r % 8): neighbouring items then share that XCD's L2, while
// every XCD still takes runs from the whole range.  A bijection on [0, n):
// the last n mod 8B items keep their places.
__device__ inline int64_t xcd_item(int64_t i, int64_t n, int B) {
    const int64_t full = n / (8 * B) * (8 * B);
    if (i >= full) return i;
    const int64_t x = i % 8, j = i / 8;
    return ((j / B) * 8 + x) * B + (j % B);
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum v[k] over the 64 lanes for k = 0..K-1 (K a power of two <= 64):
// log2(K) halving steps (each lane keeps one half and adds its partner's),
// then log2(64 / K) plain steps.  Afterwards lane L holds the sum for
// k = L / (64 / K); the 64 / K lanes of a group hold the same value.
// About one shuffle per summed index for any K, with K values per lane
// live instead of 64.  The order is fixed, so results are reproducible.
template <int K, typename A>
__device__ inline A xreduce(A (&v)[K], int lane) {
    static_assert(K >= 1 && K <= 64 && (K & (K - 1)) == 0, "K must be a power of two <= 64");
#pragma unroll
    for (int s = 32, h = K / 2; h >= 1; s >>= 1, h >>= 1) {
        const bool up = (lane & s) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            const A send = up ? v[j] : v[j + h];
            const A keep = up ? v[j + h] : v[j];
            v[j] = keep + __shfl_xor(send, s, 64);
        }
    }
    A r = v[0];
#pragma unroll
    for (int s = 32 / K; s >= 1; s >>= 1) r += __shfl_xor(r, s, 64);
    return r;
}

} // namespace hmvm
