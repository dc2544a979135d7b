/*
 * hmvm.h -- C-ABI of the MI355X H-/TLR-matrix-vector product.
 *
 * This is the drop-in boundary for HLR's mul_vec path.  Every entry point is
 * extern "C", takes plain pointers and sizes, and reports failure through an
 * int return code (never an exception across the ABI).
 *
 * Semantics shared by all entry points (mirroring the reference):
 *   - accumulate: y += alpha * op(A) * x, y is never zeroed
 *       (arith/mulvec.hh:27-64, tbb/arith.hh:312-340)
 *   - alpha == 0 returns immediately without touching y
 *       (tbb/arith.hh:54-55, tbb/arith.hh:334-335)
 *   - op in {normal, transposed, adjoint, conjugate}; for real data
 *     transposed == adjoint and conjugate == normal
 *       (arith/blas.hh:880-949, matrix/lrmatrix.hh:681-746)
 *   - vectors are in the internal (cluster) ordering and are addressed by
 *     global index minus the matrix offset (tbb/detail/arith.hh:168-169)
 *   - calls are stream ordered and asynchronous; a plan is immutable after
 *     creation and may be used from several streams at once.
 */
#ifndef HMVM_H
#define HMVM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* error codes                                                              */
/* ------------------------------------------------------------------------ */
enum {
    HMVM_OK                 = 0,
    HMVM_ERR_INVALID_ARG    = 1,   /* null pointer, bad op, nrhs < 1, ...       */
    HMVM_ERR_BAD_BLOCK      = 2,   /* offset out of range, rank > min(m,n), ... */
    HMVM_ERR_OVERLAP        = 3,   /* leaf blocks overlap or leave a hole       */
    HMVM_ERR_NONFINITE      = 4,   /* NaN/Inf in a factor (cf. check_data())    */
    HMVM_ERR_ALLOC          = 5,   /* device or host allocation failed          */
    HMVM_ERR_DEVICE         = 6,   /* HIP runtime error                         */
    HMVM_ERR_UNSUPPORTED    = 7,   /* combination not implemented               */
    HMVM_ERR_LAYOUT         = 8,   /* leading dimension too small / misaligned  */
    HMVM_ERR_KERNEL         = 9,   /* a kernel reported a failed internal check */
    HMVM_ERR_IO             = 10,  /* file missing, short, or not an H-matrix file */
    HMVM_ERR_CORRUPT        = 11,  /* H-matrix file fails its checksums         */
    HMVM_ERR_TOO_LARGE      = 12   /* workspace beyond HMVM_NOALLOC_MAX_BYTES   */
};

/* matrix operation, values as Hpro::matop_t */
enum {
    HMVM_OP_NORMAL     = 0,
    HMVM_OP_TRANSPOSED = 1,
    HMVM_OP_ADJOINT    = 2,
    HMVM_OP_CONJUGATE  = 3
};

/* leaf kinds */
enum {
    HMVM_BLOCK_DENSE      = 0,   /* D, m x n column-major           (matrix/dense_matrix.hh:36-70) */
    HMVM_BLOCK_LOWRANK    = 1,   /* U V^H, U m x k, V n x k col-maj (matrix/lrmatrix.hh:39-66)     */
    HMVM_BLOCK_LOWRANK_SV = 2    /* U diag(S) V^H, S k reals        (matrix/lrsvmatrix.hh:41-61)   */
};

/* factor storage policy of a plan */
enum {
    HMVM_PREC_FP64 = 0,   /* plain path: every factor and accumulation in FP64          */
    HMVM_PREC_FP32 = 1,   /* every factor FP32, FP32 accumulation inside a panel        */
    HMVM_PREC_VALR = 2,   /* per-column FP64/FP32/BF16/FP8-E4M3 chosen from singular
                             values against eps (compress/valr.hh:402-512,
                             compress/afl.hh:81-82,854-856); dense leaves FP32          */
    HMVM_PREC_BFP  = 3,   /* per-column block floating point (compress/bfp.hh:765-860
                             with VALR tolerances): the narrowest of the B8, B16,
                             B25 formats whose stored column, U and V alike, is within
                             the column's VALR tolerance of the FP64 column in the
                             2-norm; FP32 / FP64 when none is.  Dense leaves per
                             column within 2^-(rate(eps)+1)                          */
    HMVM_PREC_AFL  = 4    /* per-column adaptive float (compress/afl.hh:220-420,
                             compress_lr :818-880 with VALR tolerances): the narrowest
                             of the A12, A16, A24 formats whose stored column, U
                             and V alike, is within the column's VALR tolerance in
                             the 2-norm; FP32 / FP64 when none is.  Dense leaves as
                             under HMVM_PREC_BFP, with the AFL formats              */
};

/* vector element type at the boundary */
enum {
    HMVM_F64 = 0,
    HMVM_F32 = 1
};

/*
 * One leaf of the flattened block tree.
 * Replaces the leaf objects of Hpro::TBlockMatrix (arith/mulvec.hh:42-58):
 * row_ofs/col_ofs are the first indices of row_is()/col_is() in the internal
 * ordering, off_* are element offsets into the host factor arena.
 */
typedef struct hmvm_block {
    int64_t row_ofs;
    int64_t col_ofs;
    int32_t nrows;    /* m                                         */
    int32_t ncols;    /* n                                         */
    int32_t rank;     /* k (ignored for dense leaves)              */
    int32_t kind;     /* HMVM_BLOCK_*                              */
    int64_t off_a;    /* dense: D; low-rank: U                     */
    int64_t off_b;    /* low-rank: V                               */
    int64_t off_s;    /* singular values (k), or -1 when unknown   */
} hmvm_block;

/* storage formats of factor columns (index of the plan's arenas) */
enum {
    HMVM_FMT_F64  = 0,
    HMVM_FMT_F32  = 1,
    HMVM_FMT_BF16 = 2,
    HMVM_FMT_FP8  = 3,    /* OCP FP8 E4M3 with a power-of-two scale per panel column */
    /* block floating point: NB-bit two's-complement integers q packed
       floor(128 / NB) to a 16-byte row (value j in bits [j NB, j NB + NB) of
       the row as a little-endian 128-bit integer), value q * 2^k with one
       power-of-two step 2^k per panel column                                 */
    HMVM_FMT_B8   = 4,    /* NB =  8, 16 values per row */
    HMVM_FMT_B16  = 5,    /* NB = 16,  8 values per row */
    HMVM_FMT_B25  = 6,    /* NB = 25,  5 values per row */
    /* adaptive float (AFL): NB-bit codes packed as the block floating point
       values; code = (sign, 5 exponent bits e, m mantissa bits) from the
       high bit down, standing for +-(a - 1) s with a = 2^(e+1) (1 + mant 2^-m)
       evaluated in FP32, s an FP32 scale per panel column; all-ones = 0    */
    HMVM_FMT_A12  = 7,    /* NB = 12, m =  6, 10 values per row */
    HMVM_FMT_A16  = 8,    /* NB = 16, m = 10,  8 values per row */
    HMVM_FMT_A24  = 9     /* NB = 24, m = 18,  5 values per row */
};
#define HMVM_NFMT 10

typedef struct hmvm_plan_opts {
    int32_t precision;    /* HMVM_PREC_*                                          */
    int32_t dense_prec;   /* dense leaves under HMVM_PREC_VALR / _BFP / _AFL:
                             HMVM_PREC_FP64, HMVM_PREC_FP32, or HMVM_PREC_VALR =
                             by eps (VALR: FP32 while ceil(-log2 eps) <= 23, else
                             FP64; BFP / AFL: the plan's column rule at rate(eps)) */
    double  eps;          /* compression tolerance for HMVM_PREC_VALR / _BFP / _AFL */
    int32_t check_finite; /* reject NaN/Inf factors at creation (default 1)       */
    int32_t force_format; /* 0: formats from `precision`; 1 + HMVM_FMT_*: store
                             every low-rank column in that format (experiments
                             and format tests).  A plan may not hold columns of
                             two of the families BF16 / FP8, block floating
                             point and AFL (HMVM_ERR_INVALID_ARG)                 */
} hmvm_plan_opts;

typedef struct hmvm_plan_info {
    int64_t nrows, ncols;
    int64_t nblocks, ndense, nlowrank;
    int64_t nrow_panels, ncol_panels;
    int64_t factor_bytes;        /* device bytes of all stored factors           */
    int64_t bytes_by_format[HMVM_NFMT]; /* factor bytes per HMVM_FMT_*       */
    int64_t cols_by_format[HMVM_NFMT];  /* low-rank columns per HMVM_FMT_*   */
    int64_t scale_bytes;         /* panel-column scales of the FP8 and block
                                    floating point formats (4 B each)            */
    double  flops;               /* mul_vec_flops for nrhs = 1                   */
    double  datasize;            /* mul_vec_datasize restated for this storage   */
    int64_t index_bytes;         /* per-column source/destination index tables   */
    int32_t one_level;           /* 1 when the leaves form a one-level tiling    */
    int32_t reserved;
    double  bytes_rpass;         /* algorithmic factor bytes of the row-panel
                                    pass: dense leaves + U factors (+ scales)    */
    double  bytes_vpass;         /* algorithmic factor bytes of the column-panel
                                    pass: V factors (+ scales)                   */
} hmvm_plan_info;

typedef struct hmvm_plan  hmvm_plan;
typedef struct hmvm_batch hmvm_batch;

/*
 * Build a device plan from a flattened block tree.
 * Replaces construction-time preparation that the reference does implicitly
 * through the TMatrix hierarchy consumed by hlr::tbb::mul_vec
 * (tbb/arith.hh:312-340); the factor arena is host memory (FP64) that is
 * converted and uploaded once.  `device` selects the HIP device.
 */
int hmvm_plan_create(const hmvm_block* blocks, int64_t nblocks,
                     int64_t nrows, int64_t ncols,
                     const double* arena, int64_t arena_len,
                     const hmvm_plan_opts* opts, int device,
                     hmvm_plan** out);

int hmvm_plan_destroy(hmvm_plan* plan);

int hmvm_plan_get_info(const hmvm_plan* plan, hmvm_plan_info* info);

/*
 * y += alpha * op(A) * x  for nrhs right-hand sides.
 * Replaces hlr::tbb::mul_vec (tbb/arith.hh:312-340) for nrhs == 1 and the
 * multi-RHS leaf apply_add (matrix/lrmatrix.hh:681-746,
 * matrix/dense_matrix.hh:483-496) lifted to the whole H-matrix for nrhs > 1.
 * x and y are device pointers, column-major with leading dimensions ldx/ldy
 * (in elements) of type vec_dtype.  Temporaries are stream-ordered
 * allocations on `stream` (a hipStream_t, NULL = default stream).
 */
int hmvm_apply(const hmvm_plan* plan, double alpha, int op,
               const void* x, int64_t ldx, void* y, int64_t ldy,
               int nrhs, int vec_dtype, void* stream);

/*
 * hmvm_apply for one right-hand side that also records HIP events on
 * `stream` around its three kernel phases and, after synchronising the
 * stream, returns their durations in milliseconds:
 *   ms[0]  reduction pass   (forward: V^T x per column panel;
 *                            transposed: R^T x per row panel)
 *   ms[1]  partial sums     (t = alpha S sum P, or s = alpha S sum Q)
 *   ms[2]  output pass      (forward: y += R z; transposed: y += V s + dense)
 * For measurement; no counterpart in the reference.
 */
int hmvm_apply_timed(const hmvm_plan* plan, double alpha, int op,
                     const void* x, void* y, int vec_dtype, void* stream,
                     float ms[3]);

/* bytes of workspace hmvm_apply_ws needs for (op, nrhs) */
int hmvm_workspace_size(const hmvm_plan* plan, int op, int nrhs, int64_t* bytes);

/* hmvm_apply with a caller-provided device workspace (no allocation) */
int hmvm_apply_ws(const hmvm_plan* plan, double alpha, int op,
                  const void* x, int64_t ldx, void* y, int64_t ldy,
                  int nrhs, int vec_dtype, void* workspace, int64_t ws_bytes,
                  void* stream);

/* size of the fixed device scratch buffer of hmvm_apply_noalloc (per GPU) */
#define HMVM_NOALLOC_MAX_BYTES (64LL * 1024 * 1024)

/*
 * hmvm_apply that allocates nothing: no hipMalloc, hipFree or allocating
 * library call.  Its workspace is a fixed device array of
 * HMVM_NOALLOC_MAX_BYTES that the library holds on every GPU it runs on
 * until the process ends.  Arguments and results are those of hmvm_apply,
 * with the same kernels, so the results are bitwise those of hmvm_apply.
 * A call whose workspace (hmvm_workspace_size) exceeds
 * HMVM_NOALLOC_MAX_BYTES fails with HMVM_ERR_TOO_LARGE before any device
 * work.  All calls on one GPU share that GPU's array: calls that overlap in
 * time on one GPU (different streams or host threads) race and give wrong
 * results; serialise them.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).
 */
int hmvm_apply_noalloc(const hmvm_plan* plan, double alpha, int op,
                       const void* x, int64_t ldx, void* y, int64_t ldy,
                       int nrhs, int vec_dtype, void* stream);

/*
 * Tile-low-rank product.  Replaces hlr::tlr::mul_vec (arith/mulvec.hh:882-919,
 * tbb/arith.hh:998-1042).  A TLR matrix is an H-matrix with a one-level tree,
 * so this is hmvm_apply on a plan whose blocks form a one-level tiling; it
 * fails with HMVM_ERR_INVALID_ARG when the plan is not one-level.
 */
int hmvm_tlr_apply(const hmvm_plan* plan, double alpha, int op,
                   const void* x, void* y, int vec_dtype, void* stream);

/*
 * Adjoint of y = A x with respect to x and to every factor.
 * Given x and the output cotangent ybar (device, FP64 or FP32 per vec_dtype):
 *   gx     += A^T ybar                          (transposed H-MVM)
 *   dU_b    = ybar_b (V_b^T x_b)^T              (V^T x recomputed, not stored)
 *   dV_b    = x_b (U_b^T ybar_b)^T
 *   dD_b    = ybar_b x_b^T
 * Gradients are written into device buffers of FP64 values in the plan's own
 * arena layout (one per storage format, hmvm_plan_arena_elems values each);
 * pass NULL to skip.  Each buffer must be 16-byte aligned and gx aligned to
 * its element size (HMVM_ERR_LAYOUT otherwise).  For FP64, FP32 and BF16
 * storage an entry is the gradient with respect to the stored value; for
 * the scaled formats (FP8, block floating point, AFL) it is the gradient with
 * respect to the dequantised value (stored value times its column scale).
 * No counterpart in the reference (it has no differentiation); the transposed
 * product follows arith/blas.hh:910-948.
 */
int hmvm_grad(const hmvm_plan* plan, const void* x, const void* ybar,
              void* gx, void* grad_arena[HMVM_NFMT], int vec_dtype, void* stream);

/* number of stored values of the arena of storage format f (0..HMVM_NFMT-1);
   its bytes are bytes_by_format[f] (16 bytes per floor(128 / bits) values) */
int hmvm_plan_arena_elems(const hmvm_plan* plan, int fmt, int64_t* elems);

/* copy a device arena out of / into the plan (device to device) */
int hmvm_plan_get_arena(const hmvm_plan* plan, int fmt, void* dst, void* stream);
int hmvm_plan_set_arena(hmvm_plan* plan, int fmt, const void* src, void* stream);

/*
 * Expand one block's factors from host copies of the plan arenas back into
 * FP64 column-major form (U m x k into `a` and V n x k into `b`, or D m x n
 * into `a`), undoing the panel layout.  as_f64 == 0: the arenas hold stored
 * factors (decoded and multiplied by their per-column scales); as_f64 != 0:
 * every arena holds FP64 values in the same layout, e.g. gradients from
 * hmvm_grad, taken as they are.
 */
int hmvm_plan_unpack_block(const hmvm_plan* plan, int64_t block,
                           const void* const host_arenas[HMVM_NFMT], int as_f64,
                           double* a, double* b);

/*
 * The plan's layout as a map: for each of the hmvm_plan_arena_elems(fmt)
 * stored values of format fmt, src[k] is the offset, in the FP64 arena that
 * hmvm_plan_create was given, of the value it stores (the leaves' own
 * column-major factors at off_a / off_b), or -1 for padding.  Host-side;
 * works on host-only plans.  With it a caller moves whole arenas between
 * the leaves' layout and the plan's (gradients from hmvm_grad, factor
 * directions for set_arena) without per-block calls.  No counterpart in the
 * reference (its leaves keep their factors; matrix/lrmatrix.hh:175-204).
 */
int hmvm_plan_arena_source(const hmvm_plan* plan, int fmt, int64_t* src);

/* ------------------------------------------------------------------------ */
/* batching: many plans, many right-hand sides, one launch sequence         */
/* ------------------------------------------------------------------------ */

int hmvm_batch_create(const hmvm_plan* const* plans, int nplans,
                      const int* nrhs, int vec_dtype, hmvm_batch** out);
int hmvm_batch_destroy(hmvm_batch* batch);

/*
 * y_i += alpha_i A_i x_i for every plan i (forward products; nrhs[i] in
 * 1..64 right-hand sides, column-major with ldx[i] / ldy[i]).  The work of
 * all members runs as a fixed sequence of at most twelve stream-ordered
 * kernels (transposed copy of X, column-panel pass for one and for several
 * right-hand sides, two partial-sum stages, row-panel pass for one and for
 * several right-hand sides; the matrix-core passes once per count of
 * 16-wide right-hand-side tiles), one unit of work per workgroup, whatever
 * the number of members.  No counterpart in the
 * reference (it applies one matrix at a time, programs/mvm/mvm.hh:151-204).
 */
int hmvm_batch_apply(hmvm_batch* batch, const double* alphas,
                     const void* const* xs, const int64_t* ldx,
                     void* const* ys, const int64_t* ldy, void* stream);

/*
 * Capture hmvm_batch_apply into a HIP graph, replay it.  The graph keeps the
 * alphas, vector pointers and leading dimensions given at capture (in a
 * descriptor buffer of its own), whatever later apply, tune or capture calls
 * pass; it runs the schedule chosen at capture.  Applies and replays of one
 * batch share its workspace: calls on different streams must not overlap.
 */
int hmvm_batch_capture(hmvm_batch* batch, const double* alphas,
                       const void* const* xs, const int64_t* ldx,
                       void* const* ys, const int64_t* ldy, void* stream);
int hmvm_batch_replay(hmvm_batch* batch, void* stream);

/*
 * In-project autotuner: times the candidate schedules (unit order {member
 * order, largest unit first} x partial-sum items per workgroup {256, 1024})
 * on `stream` with the given vectors (y accumulates alpha = 1 products, the
 * caller resets it) and keeps the fastest; *chosen receives
 * 2 * order + index of the item count.
 */
int hmvm_batch_tune(hmvm_batch* batch, const void* const* xs,
                    const int64_t* ldx, void* const* ys,
                    const int64_t* ldy, void* stream, int* chosen);

/* synchronise `stream` and report a device error of the batch's launches */
int hmvm_batch_check(hmvm_batch* batch, void* stream);

/* ------------------------------------------------------------------------ */
/* consumers of the product                                                 */
/* ------------------------------------------------------------------------ */

/*
 * Spectral norm of the plan's matrix by power iteration, replacing
 * hlr::norm::spectral(arithmetic, A, atol, amax_it, squared)
 * (arith/norm.hh:89-101, arith/detail/norm.hh:448-527): same start vector
 * (std::default_random_engine, uniform in [-1, 1], normalised), same
 * iteration (squared: y = A^H A x, lambda = sqrt(|<x, y>|); otherwise
 * y = A x, lambda = |y|, square matrices only), same stopping rules and
 * defaults (atol == 0: sqrt(eps); amax_it == 0: max(5, min(nrows, ncols) /
 * 10)).  Every product is an hmvm_apply on `stream`; synchronises `stream`.
 * *iterations (may be NULL) receives the number of steps taken.
 */
int hmvm_spectral_norm(const hmvm_plan* plan, double atol, int64_t amax_it,
                       int squared, void* stream, double* norm,
                       int64_t* iterations);

/*
 * 2-norm of I - A W by the squared power iteration of hmvm_spectral_norm
 * (same start vector, stopping rules and defaults), replacing
 * hlr::norm::inv_error_2(arithmetic, A, A_inv) (arith/norm.hh:115-137),
 * which runs norm::spectral on the sum operator I - A A_inv.  A and W are
 * n x n plans on one device (W an approximate inverse of A, in any storage
 * precision); per step t = x - A (W x), y = t - W^H (A^H t).  Synchronises
 * `stream`.
 */
int hmvm_inv_error_2(const hmvm_plan* A, const hmvm_plan* W, double atol,
                     int64_t amax_it, void* stream, double* error,
                     int64_t* iterations);

/*
 * Linear iteration x <- x - W (A x - b), replacing linear_iteration(A, W)
 * (programs/iterrefine.hh:20-80), where W is the inverse of an H-LU
 * factorisation in FP64 or FP32 (:124-187); here W is any n x n plan on
 * A's device, e.g. an approximate inverse stored in mixed precision, while
 * every residual A x - b is computed with A's plan.  b, x and sol are device
 * vectors of n doubles; x holds the start value and receives the result.
 * Per step: c = W r, x -= c, |e| = |sol - x|, r = A x - b, with r = A x - b
 * for the start value (the reference program starts from r = b at x = 0).
 * Stops when |e| <= tol with sol given, as the reference (tol == 0: 1e-8),
 * or when |r| <= tol |b| with sol == NULL, or after max_steps steps
 * (0: 100).  history (may be NULL; 3 doubles per step, 300 when
 * max_steps == 0) receives per step
 * |r| before the step, |c| and |e| after it (NaN without sol), the three
 * norms the reference prints.  *steps (may be NULL) receives the number of
 * steps taken.  Synchronises `stream` once per step.
 */
int hmvm_linear_iteration(const hmvm_plan* A, const hmvm_plan* W,
                          const double* b, double* x, const double* sol,
                          double tol, int64_t max_steps, void* stream,
                          int64_t* steps, double* history);

/* ------------------------------------------------------------------------ */
/* on-disk H-matrix                                                         */
/* ------------------------------------------------------------------------ */

/*
 * Store a flattened H- or TLR-matrix (the leaf records and the FP64 arena
 * that hmvm_plan_create takes) in a file, replacing hlr::io::hpro::write /
 * read (utils/io.hh:59-80), which go through HLIBpro's own file formats.
 * The file is a 72-byte header (magic "HMVMHMAT", version, nrows, ncols,
 * nblocks, arena_len and checksums of the header, the records and the
 * arena), nblocks records of 9 int64 (row_ofs, col_ofs, nrows, ncols, rank,
 * kind, off_a, off_b, off_s), then arena_len doubles, little-endian.
 * Writing does not validate the leaves (plan creation does); a failed write
 * removes the file.
 */
int hmvm_hmatrix_write(const char* path, const hmvm_block* blocks,
                       int64_t nblocks, int64_t nrows, int64_t ncols,
                       const double* arena, int64_t arena_len);

/* sizes stored in an H-matrix file (any pointer may be NULL) */
int hmvm_hmatrix_read_info(const char* path, int64_t* nrows, int64_t* ncols,
                           int64_t* nblocks, int64_t* arena_len);

/*
 * Read an H-matrix file into caller buffers of exactly the stored sizes
 * (HMVM_ERR_INVALID_ARG otherwise).  HMVM_ERR_IO: missing or short file,
 * or not this format; HMVM_ERR_CORRUPT: a checksum fails, a record field
 * is out of range, or data follows the arena.
 */
int hmvm_hmatrix_read(const char* path, hmvm_block* blocks, int64_t nblocks,
                      double* arena, int64_t arena_len);

/* hmvm_hmatrix_read + hmvm_plan_create, with the host copies freed after */
int hmvm_plan_create_from_file(const char* path, const hmvm_plan_opts* opts,
                               int device, hmvm_plan** out);

/* ------------------------------------------------------------------------ */
/* uniform-H and H^2 matrices: shared or nested cluster bases               */
/* ------------------------------------------------------------------------ */

/*
 * One cluster of a row or column cluster-basis tree
 * (matrix/shared_cluster_basis.hh, matrix/nested_cluster_basis.hh).
 * Shared bases (uniform-H): every cluster with rank > 0 has an explicit
 * basis and parent == -1.  Nested bases (H^2): leaves have an explicit
 * basis; a cluster with a parent has the transfer matrix E (rank x parent
 * rank, column-major) with which its parent's basis restricted to its
 * indices is (own basis) * E (cb.transfer_mat(i) of the parent, son i).
 */
typedef struct hmvm_cluster {
    int64_t ofs;           /* first index of the cluster                        */
    int32_t size;          /* number of indices                                 */
    int32_t rank;          /* basis columns k (0: no basis)                     */
    int32_t parent;        /* nested bases: index of the parent, else -1        */
    int32_t pad_;
    int64_t off_basis;     /* explicit basis, size x rank column-major, or -1   */
    int64_t off_transfer;  /* nested bases: E, rank x parent rank, or -1        */
    int64_t off_s;         /* singular values of the basis (rank), or -1        */
} hmvm_cluster;

/* leaf of a uniform-H / H^2 matrix: a dense block, or a coupling matrix S
   (row basis rank x column basis rank, column-major) between row cluster
   `rowcb` and column cluster `colcb`, whose index sets are the block's
   (matrix/uniform_lrmatrix.hh, matrix/h2_lrmatrix.hh: M = U S V^H) */
typedef struct hmvm_ublock {
    int64_t row_ofs;
    int64_t col_ofs;
    int32_t nrows;
    int32_t ncols;
    int32_t kind;          /* HMVM_BLOCK_DENSE or HMVM_BLOCK_COUPLING           */
    int32_t rowcb;         /* coupling: index into the row clusters             */
    int32_t colcb;         /* coupling: index into the column clusters          */
    int32_t pad_;
    int64_t off;           /* dense: D (m x n); coupling: S; column-major       */
} hmvm_ublock;

#define HMVM_BLOCK_COUPLING 3

typedef struct hmvm_uplan hmvm_uplan;

/*
 * Device plan of a uniform-H (shared bases) or H^2 (nested bases) matrix.
 * The product runs as stream-ordered device passes: forward transform of
 * x (explicit bases, then transfers level by level up the column tree),
 * coupling products, backward transform (transfers down the row tree, then
 * explicit bases) and the dense near field.  opts as in hmvm_plan_create;
 * under HMVM_PREC_VALR the bases' singular values (off_s) choose per-column
 * formats.  device -1 builds and checks a host-only plan (no apply).
 */
int hmvm_uplan_create(const hmvm_cluster* rowcb, int64_t nrowcb,
                      const hmvm_cluster* colcb, int64_t ncolcb,
                      const hmvm_ublock* blocks, int64_t nblocks,
                      int64_t nrows, int64_t ncols,
                      const double* arena, int64_t arena_len,
                      const hmvm_plan_opts* opts, int device, hmvm_uplan** out);
int hmvm_uplan_destroy(hmvm_uplan* plan);

/* statistics: factor bytes of bases, transfers, couplings and dense
   leaves, mul_vec_flops and mul_vec_datasize (arith/uniform.hh,
   arith/h2.hh:140-178) */
typedef struct hmvm_uplan_info {
    int64_t nrows, ncols;
    int64_t nrow_clusters, ncol_clusters, ncoupling, ndense;
    int32_t nested;              /* 1: H^2 (transfer matrices), 0: uniform-H      */
    int32_t levels;              /* transfer levels (row tree + column tree)      */
    int64_t basis_bytes;         /* explicit bases, as stored                     */
    int64_t transfer_bytes;
    int64_t coupling_bytes;
    int64_t dense_bytes;
    double  flops;
    double  datasize;
} hmvm_uplan_info;
int hmvm_uplan_get_info(const hmvm_uplan* plan, hmvm_uplan_info* info);

/*
 * y += alpha * op(M) * x for FP64 device vectors (one right-hand side).
 * Replaces hlr::uniform::mul_vec (arith/uniform.hh:114-145) and
 * hlr::h2::mul_vec (arith/h2.hh:18-44): alpha == 0 returns at once,
 * op as for hmvm_apply.  Workspace is a stream-ordered allocation.
 */
int hmvm_uplan_apply(const hmvm_uplan* plan, double alpha, int op,
                     const double* x, double* y, void* stream);

/* ------------------------------------------------------------------------ */
/* Low-rank leaves built on the device                                       */
/* ------------------------------------------------------------------------ */

/* kernel functions a(i, j) of the generator (include/hlrgen.h kernels) */
enum {
    HMVM_KFN_LOG          = 0,   /* log|x_i - x_j|                                  */
    HMVM_KFN_LOG_GALERKIN = 1,   /* Galerkin log entries on a uniform 1D grid       */
    HMVM_KFN_MATERN       = 2,   /* sigma^2 (1 + s) exp(-s), s = sqrt(3) r / ell    */
    HMVM_KFN_INVR         = 3    /* 1 / |x_i - x_j|                                 */
};

typedef struct hmvm_kernel_fn {
    int32_t        kind;         /* HMVM_KFN_*                                     */
    int32_t        pad_;
    int64_t        n;            /* points                                         */
    double         h;            /* 1D cell width (log kernels)                    */
    double         ell, sigma;   /* Matern length scale, standard deviation        */
    const double*  coords;       /* host, n x 3 in internal order                  */
    const int64_t* perm;         /* host, internal -> external (Galerkin log)      */
} hmvm_kernel_fn;

typedef struct hmvm_lr_leaf {
    int64_t row0, col0;          /* first internal row / column                    */
    int32_t nrows, ncols;        /* disjoint row and column ranges                 */
} hmvm_lr_leaf;

typedef struct hmvm_lowrank hmvm_lowrank;

/*
 * Low-rank factors of the kernel blocks `leaves` on device `device`: ACA
 * with maximal-pivot search to relative accuracy eps, then (truncate != 0)
 * SVD truncation to eps: U (m x k) and V (n x k) orthonormal, S descending.
 * Without truncation U and V are the ACA factors and S is all ones.
 * Replaces the leaf-wise build of seq/matrix.hh:51-176 with
 * bem/aca.hh:59-78 (approx/aca.hh:117-434, approx/svd.hh:90-209); the host
 * generator's hlrgen_build runs the same steps on the CPU.  ACA ranks above
 * 128 give HMVM_ERR_UNSUPPORTED; a leaf holding a diagonal entry,
 * HMVM_ERR_BAD_BLOCK.  The factors are returned in host memory.
 */
int hmvm_lowrank_build(const hmvm_kernel_fn* fn, const hmvm_lr_leaf* leaves, int64_t nleaves, double eps,
                       int truncate, int device, hmvm_lowrank** out);
int hmvm_lowrank_ranks(const hmvm_lowrank* lr, int32_t* ranks);
/* copy leaf i's U to arena[off_a[i]..], V to arena[off_b[i]..] and (off_s
   not NULL, off_s[i] >= 0) S to arena[off_s[i]..], column-major; fold != 0
   multiplies S into the factor with fewer rows (V on ties), as
   HLRGEN_LR_UV stores it */
int hmvm_lowrank_copy(const hmvm_lowrank* lr, const int64_t* off_a, const int64_t* off_b, const int64_t* off_s,
                      int fold, double* arena, int64_t arena_len);
/* device time of the build (seconds), from HIP events around it */
double hmvm_lowrank_seconds(const hmvm_lowrank* lr);
/* of which the kernels, and the copy of the factors to the host */
int hmvm_lowrank_timing(const hmvm_lowrank* lr, double* kernel_seconds, double* copy_seconds);
void hmvm_lowrank_free(hmvm_lowrank* lr);

const char* hmvm_error_string(int code);

#ifdef __cplusplus
}
#endif

#endif /* HMVM_H */
