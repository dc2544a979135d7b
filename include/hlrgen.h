/*
 * hlrgen.h -- C-ABI of the H-/TLR-matrix generator (host C++, OpenMP).
 *
 * Builds the H-matrices the product runs on: seeded point sets, kernel
 * functions, cardinality-balanced binary space partitioning, standard
 * geometric admissibility, ACA with maximal-pivot search and SVD truncation.
 * The output is the flattened leaf list consumed by hmvm_plan_create, in the
 * order in which hlr::mul_vec visits the leaves (row-major recursion,
 * arith/mulvec.hh:40-58), plus an FP64 factor arena.
 *
 * Reference counterparts: programs/common/gen_problem.hh:30-121 (problems),
 * src/cluster/h.cc:22-42 and src/cluster/tlr.cc:22-29 (clustering),
 * programs/common/common.hh:279-290 (admissibility),
 * seq/matrix.hh:51-176 (block-wise build), bem/aca.hh:59-78 (ACA + truncate),
 * approx/aca.hh:117-434, approx/svd.hh:90-209.
 */
#ifndef HLRGEN_H
#define HLRGEN_H

#include <stdint.h>
#include "hmvm.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HLRGEN_KERNEL_LOG        = 0,  /* log|x-y|, diagonal log(h) - 3/2 (cell average)     */
    HLRGEN_KERNEL_LOG_GALERKIN = 1,/* reference logkernel: Galerkin cell integrals        */
    HLRGEN_KERNEL_MATERN     = 2,  /* Matern nu = 1.5 closed form                          */
    HLRGEN_KERNEL_INVR       = 3   /* 1/|x-y|, diagonal = mean of nn_diag nearest entries  */
};

enum {
    HLRGEN_GEOM_LINE   = 0,   /* n points uniform on [0,1] (seeded)               */
    HLRGEN_GEOM_GRID1D = 1,   /* cell centres h(i+1/2), h = 1/n                    */
    HLRGEN_GEOM_CUBE   = 2,   /* uniform in [0,1]^3                                */
    HLRGEN_GEOM_SPHERE = 3    /* uniform on the unit sphere (normalised Gaussians) */
};

enum {
    HLRGEN_FORMAT_H   = 0,    /* BSP cluster tree, standard admissibility          */
    HLRGEN_FORMAT_TLR = 1     /* one-level tiling, off-diagonal tiles low-rank     */
};

enum {
    HLRGEN_LR_UV  = 0,        /* S folded into the smaller factor (approx/svd.hh:206-207) */
    HLRGEN_LR_USV = 1         /* orthonormal U, V and separate S (lrsvmatrix)              */
};

typedef struct hlrgen_config {
    int32_t  kernel;
    int32_t  geometry;
    int64_t  n;
    uint64_t seed;
    int32_t  ntile;         /* leaf size (H) or tile size (TLR)                     */
    int32_t  format;        /* HLRGEN_FORMAT_*                                      */
    double   eta;           /* admissibility parameter                              */
    double   eps;           /* relative ACA / truncation accuracy                   */
    double   matern_ell;    /* Matern length scale                                  */
    double   matern_sigma;  /* Matern standard deviation                            */
    int32_t  nn_diag;       /* neighbours averaged on the 1/r diagonal              */
    int32_t  lrformat;      /* HLRGEN_LR_*                                          */
    int32_t  nthreads;      /* 0 = OpenMP default                                   */
    int32_t  tlr_std_adm;   /* TLR: 1 = standard admissibility, 0 = off-diagonal    */
    int32_t  verbose;
    int32_t  skip_lowrank;  /* 1: leave the low-rank leaves at rank 0 (their
                               factors are built elsewhere, e.g. on the device
                               by hmvm_lowrank_build); dense leaves as always */
} hlrgen_config;

typedef struct hlrgen_matrix hlrgen_matrix;

/* fill *cfg with the defaults (leaf 64, eta 2, eps 1e-6, ...) */
void hlrgen_default_config(hlrgen_config* cfg);

int  hlrgen_build(const hlrgen_config* cfg, hlrgen_matrix** out);
void hlrgen_free(hlrgen_matrix* m);

int64_t            hlrgen_n(const hlrgen_matrix* m);
int64_t            hlrgen_nblocks(const hlrgen_matrix* m);
const hmvm_block*  hlrgen_blocks(const hlrgen_matrix* m);
int64_t            hlrgen_arena_len(const hlrgen_matrix* m);
const double*      hlrgen_arena(const hlrgen_matrix* m);
/* internal -> external index map and coordinates (internal order, dim 3) */
const int64_t*     hlrgen_perm(const hlrgen_matrix* m);
const double*      hlrgen_coords(const hlrgen_matrix* m);
/* seconds spent in clustering / block build */
double             hlrgen_build_seconds(const hlrgen_matrix* m);

/* evaluate kernel entries a(i,j) for internal indices (for checks) */
int  hlrgen_eval(const hlrgen_matrix* m, int64_t nr, const int64_t* rows,
                 int64_t nc, const int64_t* cols, double* out /* nr x nc col-major */);

/* seeded vectors: uniform in [lo,hi] or N(0,1) (gaussian != 0), std::mt19937_64 */
void hlrgen_random_vector(uint64_t seed, int64_t n, double lo, double hi,
                          int gaussian, double* out);

/* ------------------------------------------------------------------------ */
/* numerical building blocks, exported for tests                             */
/* ------------------------------------------------------------------------ */

/* Householder QR of A (m x k, col-major, m >= k): Q (m x k), R (k x k) */
int  hlrgen_qr(int64_t m, int64_t k, const double* A, double* Q, double* R);
/* one-sided Jacobi SVD of A (k x k): A = U diag(S) V^T, S descending */
int  hlrgen_svd(int64_t k, const double* A, double* U, double* S, double* V);
/* truncate U V^T (m x r, n x r) to relative accuracy eps; returns new rank
 * into *k_out and writes orthonormal Uo (m x k), So (k), orthonormal Vo (n x k) */
int  hlrgen_truncate(int64_t m, int64_t n, int64_t r, const double* U, const double* V,
                     double eps, double* Uo, double* So, double* Vo, int64_t* k_out);
/* ACA of the kernel block rows x cols to eps; returns rank in *k_out;
 * U (nr x kmax), V (nc x kmax) must hold kmax columns */
int  hlrgen_aca(const hlrgen_matrix* m, int64_t nr, const int64_t* rows,
                int64_t nc, const int64_t* cols, double eps, int64_t kmax,
                double* U, double* V, int64_t* k_out);

#ifdef __cplusplus
}
#endif

#endif /* HLRGEN_H */
