/*
 * hmvm_oracle.c -- CPU FP64 restatement of HLR's mul_vec path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the CPU baseline -- never as the product path.
 *
 * Functions and the reference code each one restates:
 *   oracle_mul_vec        hlr::mul_vec / seq::mul_vec, leaf order of the
 *                         row-major recursion (arith/mulvec.hh:27-64,72-81;
 *                         seq/arith.hh:40-72), leaves via apply_add
 *                         (matrix/lrmatrix.hh:625-679,
 *                         matrix/dense_matrix.hh:459-481) and mulvec_lr
 *                         (arith/blas.hh:880-949, 1030-1103)
 *   oracle_tlr_mul_vec    hlr::tlr::mul_vec (arith/mulvec.hh:882-919) and the
 *                         TBB variant with a private t_i per block row
 *                         (tbb/arith.hh:998-1042)
 *   oracle_mul_vec_chunk  hlr::tbb::mul_vec_chunk: parallel leaves, private
 *                         yt per leaf, 64-entry locked chunk updates
 *                         (tbb/arith.hh:46-61; tbb/detail/arith.hh:21-23,
 *                         95-175)
 *   oracle_mul_vec_flops  mul_vec_flops (arith/mulvec.hh:795-862) with
 *                         FLOPS_DGEMV (utils/flops.hh:41-42,212)
 *   oracle_mul_vec_datasize  mul_vec_datasize (arith/mulvec.hh:867-874) with
 *                         data_byte_size (matrix/lrmatrix.hh:552-558,
 *                         matrix/dense_matrix.hh:389-395)
 *   oracle_valr_tolerances  valr::get_tolerances (compress/valr.hh:402-512)
 *   oracle_eps_to_rate    afl::eps_to_rate (compress/afl.hh:81-82)
 *   oracle_round_*        round-to-nearest-even into BF16 / FP8-E4M3 (OCP)
 *   oracle_bfp_*          block floating point columns: bfp::compress_lr
 *                         (compress/bfp.hh:765-860) with a power-of-two step
 *                         per panel column and value widths of 8, 16 or 25
 *                         bits; the width of a column is the
 *                         narrowest whose stored column is within a
 *                         given relative 2-norm tolerance (the VALR
 *                         tolerance u_l of a low-rank column)
 *   oracle_dense_expand   sum of the leaves placed at their offsets
 *   oracle_uniform_mul_vec  hlr::uniform::mul_vec (arith/uniform.hh:114-145;
 *                         arith/detail/uniform.hh:53-193) and hlr::h2::mul_vec
 *                         (arith/h2.hh:18-44; arith/detail/h2.hh:29-322):
 *                         forward transform, couplings and dense leaves in
 *                         block order, backward transform
 *   oracle_uniform_dense_expand  dense leaves plus U_r S V_c^H per coupling,
 *                         nested bases expanded through their transfers
 *
 * Pins (tests/): dense expansion against direct kernel evaluation and
 * numpy products; linearity, accumulate, alpha = 0, <Ax,w> = <x,A^T w>,
 * TLR == H on one-level trees, brute force on hand-built trees; rounding
 * against torch's float8_e4m3fn / bfloat16 conversions and the OCP table;
 * block floating point against numpy's rint, brute-force minimality of the
 * step and of the chosen width, and exactly representable columns;
 * uniform-H / H^2 products against numpy products of the dense expansion
 * and of a hand-assembled expansion (bases, transfers and couplings
 * multiplied out in numpy), and H^2 against the equal uniform-H matrix
 * whose bases are the expanded nested ones.
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hmvm.h"

#define ORACLE_CHUNK 64

/* y(m) += alpha * A(m x n, col-major, lda) * x(n) -- netlib dgemv 'N' order */
static void gemv_n(int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                   const double* x, double* y) {
    for (int64_t j = 0; j < n; ++j) {
        const double t = alpha * x[j];
        if (t == 0.0) continue;
        const double* a = A + j * lda;
        for (int64_t i = 0; i < m; ++i) y[i] += t * a[i];
    }
}

/* y(n) += alpha * A(m x n)^T * x(m) -- netlib dgemv 'T' order */
static void gemv_t(int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                   const double* x, double* y) {
    for (int64_t j = 0; j < n; ++j) {
        const double* a = A + j * lda;
        double t = 0.0;
        for (int64_t i = 0; i < m; ++i) t += a[i] * x[i];
        y[j] += alpha * t;
    }
}

/* leaf apply_add(alpha, x, y, op) for real data (transposed == adjoint) */
static void leaf_apply(const hmvm_block* b, const double* arena, double alpha, int op,
                       const double* x, double* y, double* t) {
    const int64_t m = b->nrows, n = b->ncols, k = b->rank;
    const int trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    if (b->kind == HMVM_BLOCK_DENSE) {
        const double* D = arena + b->off_a;
        if (!trans) gemv_n(m, n, alpha, D, m, x, y);
        else        gemv_t(m, n, alpha, D, m, x, y);
        return;
    }
    const double* U = arena + b->off_a;
    const double* V = arena + b->off_b;
    const double* S = (b->kind == HMVM_BLOCK_LOWRANK_SV) ? arena + b->off_s : NULL;
    if (k == 0) return;
    memset(t, 0, sizeof(double) * (size_t)k);
    if (!trans) {
        /* t := V^H x ; t := alpha (S) t ; y := y + U t */
        gemv_t(n, k, 1.0, V, n, x, t);
        for (int64_t l = 0; l < k; ++l) t[l] *= S ? alpha * S[l] : alpha;
        gemv_n(m, k, 1.0, U, m, t, y);
    } else {
        /* t := U^T x ; t := alpha (S) t ; y := y + V t */
        gemv_t(m, k, 1.0, U, m, x, t);
        for (int64_t l = 0; l < k; ++l) t[l] *= S ? alpha * S[l] : alpha;
        gemv_n(n, k, 1.0, V, n, t, y);
    }
}

static int64_t max_rank(const hmvm_block* blocks, int64_t nb) {
    int64_t k = 1;
    for (int64_t i = 0; i < nb; ++i)
        if (blocks[i].kind != HMVM_BLOCK_DENSE && blocks[i].rank > k) k = blocks[i].rank;
    return k;
}

static int64_t max_dim(const hmvm_block* blocks, int64_t nb) {
    int64_t d = 1;
    for (int64_t i = 0; i < nb; ++i) {
        if (blocks[i].nrows > d) d = blocks[i].nrows;
        if (blocks[i].ncols > d) d = blocks[i].ncols;
    }
    return d;
}

/*
 * y += alpha op(A) x, sequential, leaves in the given (recursion) order.
 * x and y hold nrhs columns with leading dimensions ldx / ldy.
 */
int oracle_mul_vec(const hmvm_block* blocks, int64_t nb, const double* arena, double alpha,
                   int op, const double* x, int64_t ldx, double* y, int64_t ldy, int nrhs) {
    if (alpha == 0.0) return 0;
    const int trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    double* t = (double*)malloc(sizeof(double) * (size_t)max_rank(blocks, nb));
    if (!t) return HMVM_ERR_ALLOC;
    for (int r = 0; r < nrhs; ++r)
        for (int64_t i = 0; i < nb; ++i) {
            const hmvm_block* b = &blocks[i];
            const int64_t xo = trans ? b->row_ofs : b->col_ofs;
            const int64_t yo = trans ? b->col_ofs : b->row_ofs;
            leaf_apply(b, arena, alpha, op, x + r * ldx + xo, y + r * ldy + yo, t);
        }
    free(t);
    return 0;
}

/*
 * TLR product: per block row i a private t_i, accumulated over the block
 * columns and then added to y_i (tbb/arith.hh:998-1042).  Block rows are
 * identified by row_ofs (op normal) or col_ofs (transposed); leaves must be
 * grouped by block row, which the one-level tilings are (row-major order).
 * nthreads > 1 runs block rows in parallel like the TBB version.
 */
int oracle_tlr_mul_vec(const hmvm_block* blocks, int64_t nb, const double* arena, double alpha,
                       int op, const double* x, double* y, int nthreads) {
    if (alpha == 0.0) return 0;
    const int trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    /* group leaves by output block: for op normal the list is already grouped,
       for transposed build column groups */
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb ? nb : 1));
    int64_t* gstart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    if (!order || !gstart) { free(order); free(gstart); return HMVM_ERR_ALLOC; }
    int64_t ng = 0, pos = 0;
    char* done = (char*)calloc((size_t)(nb ? nb : 1), 1);
    if (!done) { free(order); free(gstart); return HMVM_ERR_ALLOC; }
    for (int64_t i = 0; i < nb; ++i) {
        if (done[i]) continue;
        const int64_t key = trans ? blocks[i].col_ofs : blocks[i].row_ofs;
        gstart[ng++] = pos;
        for (int64_t j = i; j < nb; ++j) {
            const int64_t kj = trans ? blocks[j].col_ofs : blocks[j].row_ofs;
            if (!done[j] && kj == key) { order[pos++] = j; done[j] = 1; }
        }
    }
    gstart[ng] = pos;
    free(done);
    const int64_t kmax = max_rank(blocks, nb), dmax = max_dim(blocks, nb);
    int err = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        double* t = (double*)malloc(sizeof(double) * (size_t)kmax);
        double* ti = (double*)malloc(sizeof(double) * (size_t)dmax);
        if (!t || !ti) {
#pragma omp atomic write
            err = HMVM_ERR_ALLOC;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int64_t g = 0; g < ng; ++g) {
                const hmvm_block* b0 = &blocks[order[gstart[g]]];
                const int64_t len = trans ? b0->ncols : b0->nrows;
                const int64_t yo = trans ? b0->col_ofs : b0->row_ofs;
                memset(ti, 0, sizeof(double) * (size_t)len);
                for (int64_t p = gstart[g]; p < gstart[g + 1]; ++p) {
                    const hmvm_block* b = &blocks[order[p]];
                    const int64_t xo = trans ? b->row_ofs : b->col_ofs;
                    leaf_apply(b, arena, alpha, op, x + xo, ti, t);
                }
                for (int64_t i = 0; i < len; ++i) y[yo + i] += ti[i];
            }
        }
        free(t);
        free(ti);
    }
    free(order);
    free(gstart);
    return err;
}

/*
 * Restatement of tbb::mul_vec_chunk: every leaf computes yt = alpha op(B) x
 * into a private temporary and adds it to y in 64-entry chunks, each chunk
 * under its own lock.  The lock vector is allocated per call, as in the
 * reference.  Summation order across leaves is not deterministic.
 */
int oracle_mul_vec_chunk(const hmvm_block* blocks, int64_t nb, const double* arena, double alpha,
                         int op, const double* x, double* y, int64_t ny, int nthreads) {
    if (alpha == 0.0) return 0;
    const int trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const int64_t nchunks = ny / ORACLE_CHUNK + 1;
    omp_lock_t* locks = (omp_lock_t*)malloc(sizeof(omp_lock_t) * (size_t)nchunks);
    if (!locks) return HMVM_ERR_ALLOC;
    for (int64_t c = 0; c < nchunks; ++c) omp_init_lock(&locks[c]);
    const int64_t kmax = max_rank(blocks, nb), dmax = max_dim(blocks, nb);
    int err = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        double* t = (double*)malloc(sizeof(double) * (size_t)kmax);
        double* yt = (double*)malloc(sizeof(double) * (size_t)dmax);
        if (!t || !yt) {
#pragma omp atomic write
            err = HMVM_ERR_ALLOC;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int64_t i = 0; i < nb; ++i) {
                const hmvm_block* b = &blocks[i];
                const int64_t xo = trans ? b->row_ofs : b->col_ofs;
                const int64_t yo = trans ? b->col_ofs : b->row_ofs;
                const int64_t len = trans ? b->ncols : b->nrows;
                memset(yt, 0, sizeof(double) * (size_t)len);
                leaf_apply(b, arena, alpha, op, x + xo, yt, t);
                /* update(): chunks of the global index space */
                int64_t start = yo;
                const int64_t last = yo + len - 1;
                int64_t chunk = start / ORACLE_CHUNK;
                int64_t end = (chunk + 1) * ORACLE_CHUNK - 1;
                if (end > last) end = last;
                while (start <= end) {
                    omp_set_lock(&locks[chunk]);
                    for (int64_t g = start; g <= end; ++g) y[g] += yt[g - yo];
                    omp_unset_lock(&locks[chunk]);
                    ++chunk;
                    start = end + 1;
                    end = end + ORACLE_CHUNK;
                    if (end > last) end = last;
                }
            }
        }
        free(t);
        free(yt);
    }
    for (int64_t c = 0; c < nchunks; ++c) omp_destroy_lock(&locks[c]);
    free(locks);
    return err;
}

/* seconds for nmvm calls of oracle_mul_vec_chunk (alpha = 2, x given) */
double oracle_time_chunk(const hmvm_block* blocks, int64_t nb, const double* arena,
                         const double* x, double* y, int64_t ny, int nthreads, int nmvm) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < nmvm; ++i)
        oracle_mul_vec_chunk(blocks, nb, arena, 2.0, HMVM_OP_NORMAL, x, y, ny, nthreads);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* FLOPS_DGEMV(m, n) = m n + 2 m (mults) + m n (adds) */
static double flops_dgemv(double m, double n) { return 2.0 * m * n + 2.0 * m; }

double oracle_mul_vec_flops(const hmvm_block* blocks, int64_t nb, int op) {
    const int trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    double f = 0.0;
    for (int64_t i = 0; i < nb; ++i) {
        const double m = trans ? blocks[i].ncols : blocks[i].nrows;
        const double n = trans ? blocks[i].nrows : blocks[i].ncols;
        if (blocks[i].kind == HMVM_BLOCK_DENSE) f += flops_dgemv(m, n);
        else f += flops_dgemv(n, blocks[i].rank) + flops_dgemv(m, blocks[i].rank);
    }
    return f;
}

/* data bytes of the leaves + 2 sizeof(double) (nrows + ncols) */
double oracle_mul_vec_datasize(const hmvm_block* blocks, int64_t nb, int64_t nrows, int64_t ncols) {
    double s = 0.0;
    for (int64_t i = 0; i < nb; ++i) {
        const double m = blocks[i].nrows, n = blocks[i].ncols;
        if (blocks[i].kind == HMVM_BLOCK_DENSE) s += 8.0 * m * n;
        else if (blocks[i].kind == HMVM_BLOCK_LOWRANK_SV) s += 8.0 * blocks[i].rank * (1.0 + m + n);
        else s += 8.0 * blocks[i].rank * (m + n);
    }
    return s + 2.0 * 8.0 * (double)(nrows + ncols);
}

/* A (nrows x ncols col-major) = sum of leaves at their offsets */
int oracle_dense_expand(const hmvm_block* blocks, int64_t nb, const double* arena,
                        int64_t nrows, int64_t ncols, double* A) {
    memset(A, 0, sizeof(double) * (size_t)(nrows * ncols));
    for (int64_t i = 0; i < nb; ++i) {
        const hmvm_block* b = &blocks[i];
        const int64_t m = b->nrows, n = b->ncols, k = b->rank;
        if (b->row_ofs < 0 || b->col_ofs < 0 || b->row_ofs + m > nrows || b->col_ofs + n > ncols)
            return HMVM_ERR_BAD_BLOCK;
        for (int64_t j = 0; j < n; ++j) {
            double* a = A + (b->col_ofs + j) * nrows + b->row_ofs;
            if (b->kind == HMVM_BLOCK_DENSE) {
                const double* d = arena + b->off_a + j * m;
                for (int64_t r = 0; r < m; ++r) a[r] += d[r];
            } else {
                for (int64_t l = 0; l < k; ++l) {
                    double v = arena[b->off_b + l * n + j];
                    if (b->kind == HMVM_BLOCK_LOWRANK_SV) v *= arena[b->off_s + l];
                    const double* u = arena + b->off_a + l * m;
                    for (int64_t r = 0; r < m; ++r) a[r] += u[r] * v;
                }
            }
        }
    }
    return 0;
}

/*
 * valr::get_tolerances for relative accuracy eps in the spectral norm
 * (norm = sv(0)), including the correction that solves for eps' with
 * sum(1/sigma) eps'^2 + eps' = delta via p/q (compress/valr.hh:456-499);
 * returns tol / sigma_l per column.
 */
void oracle_valr_tolerances(int64_t k, const double* sv, double eps, double* out) {
    if (k <= 0) return;
    double tol = eps * sv[0];
    double sum = 0.0;
    for (int64_t i = 0; i < k; ++i) sum += 1.0 / sv[i];
    const double p = 1.0 / sum, q = tol / sum;
    double t = p * p / 4.0 - q;
    if (t < 0) t = 0;
    const double d1 = fabs(p / 2.0 + sqrt(t)), d2 = fabs(p / 2.0 - sqrt(t));
    tol = (d1 < d2) ? d1 : d2;
    for (int64_t l = 0; l < k; ++l) out[l] = tol / sv[l];
}

/* mantissa bits needed for relative accuracy eps: max(1, ceil(-log2 eps)) */
int oracle_eps_to_rate(double eps) {
    const double r = ceil(-log2(eps));
    return (int)(r < 1.0 ? 1.0 : r);
}

/* storage format for a column with mantissa-bit requirement `rate`:
   0 = FP64 (52), 1 = FP32 (23), 2 = BF16 (7), 3 = FP8-E4M3 (3) */
int oracle_format_for_rate(int rate) {
    if (rate <= 3) return 3;
    if (rate <= 7) return 2;
    if (rate <= 23) return 1;
    return 0;
}

/* round-to-nearest-even of a double to BF16, returned as double */
double oracle_round_bf16(double v) {
    const float f = (float)v;              /* the device converts from FP32 too */
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;   /* inf / nan unchanged */
    const uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

/*
 * OCP FP8 E4M3: 1 sign, 4 exponent (bias 7), 3 mantissa bits, no infinity,
 * max finite 448, subnormals down to 2^-9.  Round to nearest even,
 * saturating at +-448.  Returns the 8-bit code.
 */
uint8_t oracle_encode_e4m3(double v) {
    const float f = (float)v;
    uint8_t sign = (f < 0 || (f == 0 && signbit(f))) ? 0x80 : 0x00;
    double a = fabs((double)f);
    if (isnan(f)) return 0x7f;
    if (a >= 464.0) return sign | 0x7e;                 /* saturate (448 = 0x7e) */
    /* grid spacing for |a|: exponent e with 2^e <= a < 2^(e+1), e >= -6 */
    int e = (a > 0) ? (int)floor(log2(a)) : -6;
    if (e < -6) e = -6;
    const double q = ldexp(1.0, e - 3);                  /* spacing */
    double r = a / q;
    double fl = floor(r);
    double diff = r - fl;
    double mant;
    if (diff > 0.5) mant = fl + 1;
    else if (diff < 0.5) mant = fl;
    else mant = (fmod(fl, 2.0) == 0.0) ? fl : fl + 1;
    double val = mant * q;
    if (val > 448.0) val = 448.0;
    /* encode val */
    if (val == 0.0) return sign;
    int ee = (int)floor(log2(val));
    if (ee < -6) {                                       /* subnormal: m * 2^-9 */
        const int m = (int)(val / ldexp(1.0, -9) + 0.5);
        return sign | (uint8_t)m;
    }
    const int m = (int)(val / ldexp(1.0, ee - 3) + 0.5) - 8;
    return sign | (uint8_t)(((ee + 7) << 3) | m);
}

double oracle_decode_e4m3(uint8_t c) {
    const int s = c >> 7, e = (c >> 3) & 0xf, m = c & 7;
    double v;
    if (e == 0xf && m == 7) return NAN;
    if (e == 0) v = ldexp((double)m, -9);
    else v = ldexp(1.0 + m / 8.0, e - 7);
    return s ? -v : v;
}

double oracle_round_e4m3(double v) { return oracle_decode_e4m3(oracle_encode_e4m3(v)); }

/*
 * Block floating point.  A panel column of NB-bit values stores q_i with
 * v_i ~ q_i 2^k, |q_i| <= 2^(NB-1) - 1, one step 2^k per panel column.
 */

/* smallest k with maxabs <= (2^(nb-1) - 1) 2^k, clamped to [-126, 127]
   (the step is kept as a normal FP32 number); 0 for maxabs == 0 */
int oracle_bfp_exp(double maxabs, int nb) {
    if (!(maxabs > 0) || isinf(maxabs)) return 0;
    const double qmax = ldexp(1.0, nb - 1) - 1.0;
    int e;
    (void)frexp(maxabs, &e);                     /* 2^(e-1) <= maxabs < 2^e */
    int k = e - nb - 1;                          /* qmax 2^k < 2^(e-2): too small */
    while (maxabs > ldexp(qmax, k)) ++k;
    if (k < -126) k = -126;
    if (k > 127) k = 127;
    return k;
}

/* nearest multiple q 2^k of v (ties to even q), |q| <= 2^(nb-1) - 1 */
double oracle_bfp_round(double v, int k, int nb) {
    const double qmax = ldexp(1.0, nb - 1) - 1.0;
    const double r = ldexp(v, -k);
    const double fl = floor(r), diff = r - fl;
    double q;
    if (diff > 0.5) q = fl + 1;
    else if (diff < 0.5) q = fl;
    else q = (fmod(fl, 2.0) == 0.0) ? fl : fl + 1;
    if (q > qmax) q = qmax;
    if (q < -qmax) q = -qmax;
    return ldexp(q, k);
}

/* value widths of the formats HMVM_FMT_B8, HMVM_FMT_B16, HMVM_FMT_B25 */
static const int oracle_bfp_bits[3] = {8, 16, 25};

/*
 * Storage format of one column group: ncols columns, values concatenated in
 * `vals`; column c consists of npieces[c] panel columns whose lengths follow
 * one another in `plen`.  Returns the first of HMVM_FMT_B8, _B16, _B25
 * for which every column's stored pieces are within tol (relative) of the
 * column in the 2-norm, else HMVM_FMT_F32 for rate <= 23, else HMVM_FMT_F64.
 */
int oracle_bfp_format(int ncols, const double* vals, const int64_t* npieces, const int64_t* plen, double tol,
                      int rate) {
    for (int f = 0; f < 3; ++f) {
        const int nb = oracle_bfp_bits[f];
        int ok = 1;
        const double* v = vals;
        const int64_t* pl = plen;
        for (int c = 0; c < ncols; ++c) {
            const double* v0 = v;
            const int64_t* pl0 = pl;
            double n2 = 0.0, e2 = 0.0;
            for (int64_t p = 0; p < npieces[c]; ++p, ++pl)
                for (int64_t i = 0; i < *pl; ++i, ++v) n2 += v[0] * v[0];
            v = v0;
            pl = pl0;
            for (int64_t p = 0; p < npieces[c]; ++p, ++pl) {
                double mx = 0.0;
                for (int64_t i = 0; i < *pl; ++i) mx = fmax(mx, fabs(v[i]));
                const int k = oracle_bfp_exp(mx, nb);
                for (int64_t i = 0; i < *pl; ++i, ++v) {
                    const double d = oracle_bfp_round(v[0], k, nb) - v[0];
                    e2 += d * d;
                }
            }
            if (!(e2 <= (tol * tol) * n2)) ok = 0;
        }
        if (ok) return HMVM_FMT_B8 + f;
    }
    return rate <= 23 ? HMVM_FMT_F32 : HMVM_FMT_F64;
}

/* ------------------------------------------------------------------------ */
/* uniform-H (shared bases) and H^2 (nested bases)                          */
/* ------------------------------------------------------------------------ */

/* depth of every cluster in its basis tree (-1 on a cycle or a bad parent) */
static int cluster_depths(const hmvm_cluster* cb, int64_t n, int* depth) {
    for (int64_t i = 0; i < n; ++i) {
        int d = 0;
        int64_t p = cb[i].parent;
        while (p >= 0) {
            if (p >= n || ++d > n) return -1;
            p = cb[p].parent;
        }
        depth[i] = d;
    }
    return 0;
}

/*
 * oracle_uniform_mul_vec: y += alpha op(M) x for a uniform-H or H^2 matrix.
 *   forward transform (scalar_to_uniform, arith/detail/uniform.hh:121-146
 *     for shared bases, arith/detail/h2.hh:238-286 for nested ones): the
 *     coefficients of every cluster of the input tree, xh_c = B_c^H x|c for
 *     an explicit basis, xh_c = sum over sons s of E_s^H xh_s otherwise
 *     (sons before parents);
 *   leaves in block order (arith/detail/uniform.hh:53-118,
 *     arith/detail/h2.hh:29-89): a coupling adds alpha op(S) xh to the
 *     output coefficients (uniform_lrmatrix::uni_apply_add), a dense leaf
 *     adds alpha op(D) x|in to y|out directly;
 *   backward transform (add_uniform_to_scalar, arith/detail/uniform.hh:
 *     171-193, arith/detail/h2.hh:289-322): parents before sons, the
 *     accumulated coefficient s_c = yh_c + E_c s_parent, and y|c += B_c s_c
 *     for an explicit basis.
 * The input tree is the column tree for op normal, the row tree for
 * transposed / adjoint (real values: conjugate == normal).
 */
int oracle_uniform_mul_vec(const hmvm_cluster* rowcb, int64_t nr, const hmvm_cluster* colcb, int64_t nc,
                           const hmvm_ublock* blocks, int64_t nb, const double* arena, double alpha, int op,
                           const double* x, double* y) {
    if (alpha == 0.0) return 0;
    const int trans = (op == HMVM_OP_TRANSPOSED || op == HMVM_OP_ADJOINT);
    const hmvm_cluster* in = trans ? rowcb : colcb;
    const hmvm_cluster* out = trans ? colcb : rowcb;
    const int64_t nin = trans ? nr : nc, nout = trans ? nc : nr;
    int* din = (int*)calloc((size_t)(nin + 1), sizeof(int));
    int* dout = (int*)calloc((size_t)(nout + 1), sizeof(int));
    int64_t* ein = (int64_t*)calloc((size_t)(nin + 1), sizeof(int64_t));
    int64_t* eout = (int64_t*)calloc((size_t)(nout + 1), sizeof(int64_t));
    int rc = 0;
    double *xh = NULL, *yh = NULL;
    if (!din || !dout || !ein || !eout || cluster_depths(in, nin, din) || cluster_depths(out, nout, dout)) {
        rc = 1;
        goto done;
    }
    for (int64_t i = 0; i < nin; ++i) ein[i + 1] = ein[i] + in[i].rank;
    for (int64_t i = 0; i < nout; ++i) eout[i + 1] = eout[i] + out[i].rank;
    xh = (double*)calloc((size_t)(ein[nin] + 1), sizeof(double));
    yh = (double*)calloc((size_t)(eout[nout] + 1), sizeof(double));
    if (!xh || !yh) { rc = 1; goto done; }
    int maxd = 0;
    for (int64_t i = 0; i < nin; ++i) maxd = din[i] > maxd ? din[i] : maxd;
    for (int64_t i = 0; i < nout; ++i) maxd = dout[i] > maxd ? dout[i] : maxd;
    /* forward transform, deepest clusters first */
    for (int d = maxd; d >= 0; --d)
        for (int64_t i = 0; i < nin; ++i) {
            const hmvm_cluster* c = &in[i];
            if (din[i] != d || c->rank == 0) continue;
            if (c->off_basis >= 0)
                gemv_t(c->size, c->rank, 1.0, arena + c->off_basis, c->size, x + c->ofs, xh + ein[i]);
            if (c->parent >= 0 && c->off_transfer >= 0 && in[c->parent].rank > 0)
                /* xh_parent += E^H xh_c, E rank x parent rank */
                gemv_t(c->rank, in[c->parent].rank, 1.0, arena + c->off_transfer, c->rank, xh + ein[i],
                       xh + ein[c->parent]);
        }
    /* leaves in block order */
    for (int64_t b = 0; b < nb; ++b) {
        const hmvm_ublock* B = &blocks[b];
        if (B->kind == HMVM_BLOCK_DENSE) {
            if (!trans) gemv_n(B->nrows, B->ncols, alpha, arena + B->off, B->nrows, x + B->col_ofs, y + B->row_ofs);
            else gemv_t(B->nrows, B->ncols, alpha, arena + B->off, B->nrows, x + B->row_ofs, y + B->col_ofs);
        } else {
            const int64_t kr = rowcb[B->rowcb].rank, kc = colcb[B->colcb].rank;
            if (kr == 0 || kc == 0) continue;
            if (!trans)
                gemv_n(kr, kc, alpha, arena + B->off, kr, xh + ein[B->colcb], yh + eout[B->rowcb]);
            else
                gemv_t(kr, kc, alpha, arena + B->off, kr, xh + ein[B->rowcb], yh + eout[B->colcb]);
        }
    }
    /* backward transform, parents first: s_c = yh_c + E_c s_parent */
    for (int d = 0; d <= maxd; ++d)
        for (int64_t i = 0; i < nout; ++i) {
            const hmvm_cluster* c = &out[i];
            if (dout[i] != d || c->rank == 0) continue;
            if (c->parent >= 0 && c->off_transfer >= 0 && out[c->parent].rank > 0)
                gemv_n(c->rank, out[c->parent].rank, 1.0, arena + c->off_transfer, c->rank, yh + eout[c->parent],
                       yh + eout[i]);
            if (c->off_basis >= 0)
                gemv_n(c->size, c->rank, 1.0, arena + c->off_basis, c->size, yh + eout[i], y + c->ofs);
        }
done:
    free(din); free(dout); free(ein); free(eout); free(xh); free(yh);
    return rc;
}

/* dense expansion of a uniform-H / H^2 matrix into A (nrows x ncols,
   column-major): dense leaves, and B_r S B_c^H per coupling with the
   explicit bases of nested clusters expanded through the transfers */
static void expand_basis(const hmvm_cluster* cb, int64_t ncb, int64_t i, const double* arena, double* out,
                         int64_t ld) {
    /* out (cluster size x rank, leading dimension ld) = explicit basis, or
       sum over sons s: rows of s get (basis of s) E_s */
    const hmvm_cluster* c = &cb[i];
    if (c->off_basis >= 0) {
        for (int64_t l = 0; l < c->rank; ++l)
            for (int64_t r = 0; r < c->size; ++r) out[l * ld + r] = arena[c->off_basis + l * c->size + r];
        return;
    }
    for (int64_t l = 0; l < c->rank; ++l)
        for (int64_t r = 0; r < c->size; ++r) out[l * ld + r] = 0.0;
    for (int64_t s = 0; s < ncb; ++s) {
        if (cb[s].parent != i || cb[s].rank == 0 || cb[s].off_transfer < 0) continue;
        const hmvm_cluster* so = &cb[s];
        double* bs = (double*)calloc((size_t)so->size * (size_t)so->rank + 1, sizeof(double));
        expand_basis(cb, ncb, s, arena, bs, so->size);
        /* out[rows of s, :] += bs (size_s x rank_s) * E_s (rank_s x rank_c) */
        for (int64_t l = 0; l < c->rank; ++l)
            for (int64_t j = 0; j < so->rank; ++j) {
                const double e = arena[so->off_transfer + l * so->rank + j];
                for (int64_t r = 0; r < so->size; ++r) out[l * ld + (so->ofs - c->ofs) + r] += bs[j * so->size + r] * e;
            }
        free(bs);
    }
}

int oracle_uniform_dense_expand(const hmvm_cluster* rowcb, int64_t nr, const hmvm_cluster* colcb, int64_t nc,
                                const hmvm_ublock* blocks, int64_t nb, const double* arena, int64_t nrows,
                                int64_t ncols, double* A) {
    memset(A, 0, sizeof(double) * (size_t)(nrows * ncols));
    for (int64_t b = 0; b < nb; ++b) {
        const hmvm_ublock* B = &blocks[b];
        if (B->kind == HMVM_BLOCK_DENSE) {
            for (int64_t j = 0; j < B->ncols; ++j)
                for (int64_t i = 0; i < B->nrows; ++i)
                    A[(B->col_ofs + j) * nrows + B->row_ofs + i] += arena[B->off + j * B->nrows + i];
            continue;
        }
        const hmvm_cluster* r = &rowcb[B->rowcb];
        const hmvm_cluster* c = &colcb[B->colcb];
        if (r->rank == 0 || c->rank == 0) continue;
        double* U = (double*)calloc((size_t)r->size * (size_t)r->rank, sizeof(double));
        double* V = (double*)calloc((size_t)c->size * (size_t)c->rank, sizeof(double));
        if (!U || !V) { free(U); free(V); return 1; }
        expand_basis(rowcb, nr, B->rowcb, arena, U, r->size);
        expand_basis(colcb, nc, B->colcb, arena, V, c->size);
        const double* S = arena + B->off;
        for (int64_t j = 0; j < c->size; ++j)
            for (int64_t i = 0; i < r->size; ++i) {
                double v = 0.0;
                for (int64_t p = 0; p < r->rank; ++p)
                    for (int64_t q = 0; q < c->rank; ++q) v += U[p * r->size + i] * S[q * r->rank + p] * V[q * c->size + j];
                A[(c->ofs + j) * nrows + r->ofs + i] += v;
            }
        free(U);
        free(V);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* adaptive float (AFL) columns                                             */
/* ------------------------------------------------------------------------ */

/*
 * The AFL codec of compress/afl.hh:220-332 (compress) and :333-420
 * (decompress), restated for fixed code widths: a value v of a column
 * with scale s is stored as a = |v| / s + 1 >= 2, keeping the sign, the low
 * ORACLE_AFL_EBITS bits of a's binary exponent (a's exponent minus one, as
 * the reference keeps the low bits of the biased exponent above 2^7) and
 * the leading m bits of a's mantissa, truncated; code = (sign, exponent,
 * mantissa) from the high bits down in an NB = 1 + ebits + m bit field; the
 * all-ones code is zero.  Decoding rebuilds a in FP32 and returns
 * sign * fl32(a - 1) * s, the reference's ( fival.f - 1 ) * scale.
 * The scale is the column's smallest non-zero |v| (afl.hh:835-854), raised
 * to max|v| / 2^(2^ebits) when the exponent bits cannot span the column's
 * range; values below the scale are stored as zero.
 */
#define ORACLE_AFL_EBITS 5

float oracle_afl_scale(const double* v, int64_t n) {
    double vmax = 0.0, vmin = INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const double a = fabs(v[i]);
        if (a > vmax) vmax = a;
        if (a > 0.0 && a < vmin) vmin = a;
    }
    if (!(vmax > 0.0)) return 1.0f;
    double s = vmin;
    const double lo = ldexp(vmax, -(1 << ORACLE_AFL_EBITS));
    if (s < lo) s = lo;
    float f = (float)s;
    if ((double)f > s) f = nextafterf(f, 0.0f);       /* not above the exact scale */
    if (!(f > 0.0f)) f = FLT_MIN;
    return f;
}

uint32_t oracle_afl_encode(double v, float scale, int nb) {
    const int m = nb - 1 - ORACLE_AFL_EBITS;
    const uint32_t zero = (nb >= 32) ? 0xffffffffu : ((1u << nb) - 1u);
    const double a = fabs(v) / (double)scale + 1.0;
    if (!(a >= 2.0) || !isfinite(a)) return zero;
    int E;
    const double fr = frexp(a, &E);                   /* a = fr 2^E, fr in [0.5, 1) */
    const int ex = E - 1;                             /* a in [2^ex, 2^(ex+1)), ex >= 1 */
    if (ex - 1 > (1 << ORACLE_AFL_EBITS) - 1) return zero;
    const double frac = fr * 2.0 - 1.0;               /* mantissa in [0, 1) */
    const uint32_t zm = (uint32_t)floor(ldexp(frac, m));
    const uint32_t sign = (v < 0.0) ? 1u : 0u;
    return (sign << (nb - 1)) | ((uint32_t)(ex - 1) << m) | zm;
}

double oracle_afl_decode(uint32_t code, float scale, int nb) {
    const int m = nb - 1 - ORACLE_AFL_EBITS;
    const uint32_t zero = (nb >= 32) ? 0xffffffffu : ((1u << nb) - 1u);
    if (code == zero) return 0.0;
    const uint32_t mant = code & ((1u << m) - 1u);
    const uint32_t ex = (code >> m) & ((1u << ORACLE_AFL_EBITS) - 1u);
    const uint32_t sign = code >> (nb - 1);
    const uint32_t bits = ((ex | 0x80u) << 23) | (mant << (23 - m));
    float a;
    memcpy(&a, &bits, sizeof(a));
    const float d = a - 1.0f;                         /* FP32, as the reference */
    return (sign ? -1.0 : 1.0) * (double)d * (double)scale;
}

/* stored value of v: decode(encode(v)) */
double oracle_afl_round(double v, float scale, int nb) {
    return oracle_afl_decode(oracle_afl_encode(v, scale, nb), scale, nb);
}

/*
 * Format of the columns `vals` (ncols columns of npieces[c] pieces of
 * plen[..] values each, concatenated) sharing one AFL width: the narrowest
 * of 12, 16, 24 bits whose stored pieces (scale per piece) keep every column
 * within tol of itself in the 2-norm, relative; else FP32 while rate <= 23,
 * else FP64.  The widths follow the per-column mantissa choice of
 * compress/afl.hh:818-880 (compress_lr) with VALR tolerances, rounded up
 * to three code widths.
 */
static const int oracle_afl_bits[3] = {12, 16, 24};

int oracle_afl_format(int ncols, const double* vals, const int64_t* npieces, const int64_t* plen, double tol,
                      int rate) {
    for (int f = 0; f < 3; ++f) {
        const int nb = oracle_afl_bits[f];
        int ok = 1;
        const double* v = vals;
        const int64_t* pl = plen;
        for (int c = 0; c < ncols; ++c) {
            double n2 = 0.0, e2 = 0.0;
            for (int64_t p = 0; p < npieces[c]; ++p, ++pl) {
                const float s = oracle_afl_scale(v, *pl);
                for (int64_t i = 0; i < *pl; ++i, ++v) {
                    const double d = oracle_afl_round(v[0], s, nb) - v[0];
                    n2 += v[0] * v[0];
                    e2 += d * d;
                }
            }
            if (!(e2 <= (tol * tol) * n2)) ok = 0;
        }
        if (ok) return HMVM_FMT_A12 + f;
    }
    return rate <= 23 ? HMVM_FMT_F32 : HMVM_FMT_F64;
}
