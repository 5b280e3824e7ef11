/*
 * oracle.h -- CPU restatement of the reference SPLATT algorithms on the
 * sparse-CPD path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else. The product (libsplatt) never links or calls it.
 *
 * Every function names the reference file:line it restates. Pinning (which
 * fixture, closed form or brute force checks each function) is listed in the
 * tests that load it (tests/test_oracle_*.py).
 */
#ifndef SPLATT_ORACLE_H
#define SPLATT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#define OR_MAX_NMODES 8

typedef struct
{
  uint64_t nmodes;
  uint64_t nnz;
  uint64_t dims[OR_MAX_NMODES];
  uint64_t * ind[OR_MAX_NMODES];
  double * vals;
} or_coo;

typedef struct
{
  uint64_t nfibs[OR_MAX_NMODES];
  uint64_t * fptr[OR_MAX_NMODES];
  uint64_t * fids[OR_MAX_NMODES];  /* fids[0] NULL when untiled */
  double * vals;
} or_tile;

typedef struct
{
  uint64_t nnz;
  uint64_t nmodes;
  uint64_t dims[OR_MAX_NMODES];
  uint64_t dim_perm[OR_MAX_NMODES];
  uint64_t dim_iperm[OR_MAX_NMODES];
  int which_tile;            /* 0 = NOTILE, 1 = DENSETILE */
  uint64_t ntiles;
  uint64_t ntiled_modes;
  uint64_t tile_dims[OR_MAX_NMODES];
  or_tile * pt;
} or_csf;

/* io: src/io.c:62-108, :273-348 */
or_coo * or_tns_read(char const * path);
or_coo * or_coo_alloc(uint64_t nmodes, uint64_t nnz);
void or_coo_free(or_coo * tt);
/* src/sptensor.c:190-252; returns number of removed slices */
uint64_t or_remove_empty(or_coo * tt);

/* src/csf.c:93-218, :608-651; type: 0 small-first, 1 big-first,
 * 2 in-order-minus-one, 3 sorted-minus-one */
void or_mode_order(uint64_t const * dims, uint64_t nmodes, int type,
    uint64_t mode, uint64_t * perm);

/* One CSF, src/csf.c:382-416 (untiled) and :427-501 + src/tile.c:260-342
 * (dense tiling). If reorder_tt != 0 the COO is left in CSF order, as the
 * reference's tt_sort/tt_densetile leave it. */
or_csf * or_csf_build(or_coo * tt, uint64_t const * perm, int which_tile,
    uint64_t tilelevel, uint64_t nthreads, int reorder_tt);
void or_csf_free(or_csf * ct);
/* src/csf.c:753 */
double or_csf_frobsq(or_csf const * ct);

/* MTTKRP by brute force over the COO in long double (the reference's own
 * test oracle: tests/mttkrp_test.c:69 with src/mttkrp.c:1699). */
void or_mttkrp_coo(or_coo const * tt, double ** mats, uint64_t rank,
    uint64_t mode, double * out);
void or_mttkrp_coo_f(or_coo const * tt, float ** mats, uint64_t rank,
    uint64_t mode, float * out);

/* CSF MTTKRP, src/mttkrp.c:324-1330 (root / internal / leaf DFS), run on
 * `nthreads` OpenMP threads with the reference's lock pool
 * (src/mutex_pool.h:95-141) guarding shared output rows. */
void or_mttkrp_csf(or_csf const * ct, double ** mats, uint64_t rank,
    uint64_t mode, double * out, int nthreads);
void or_mttkrp_csf_f(or_csf const * ct, float ** mats, uint64_t rank,
    uint64_t mode, float * out, int nthreads);

/* glibc rand() based initialisation, src/util.c:15-40 (caller seeds). */
void or_fill_rand(double * vals, uint64_t n);

/* CPD-ALS, src/cpd/cpd.c:161-332 with src/cpd/admm.c:217-345 (closed form,
 * unconstrained) and src/matrix.c. factors[m] (dims[m] x rank, row-major)
 * hold the initial values and receive the result; lambda (rank) receives
 * the weights; fits (max_its) receives 1 - rel-errsq per iteration.
 * csf[m_csf] is used for mode m via mode_csf[m]. Returns iterations run. */
int or_cpd_als(or_csf * const * csfs, uint64_t const * mode_csf,
    uint64_t nmodes, uint64_t rank, double ** factors, double * lambda,
    uint64_t max_its, double tol, double * fits, int nthreads);

/* A constraint as the CPD restatement runs it: kind 0 = unconstrained
 * closed form, 1 = closed form with diag_add on the Gram diagonal
 * (src/cpd/constraints/frob.c:25-36), 2 = ADMM (src/cpd/admm.c:346-406,
 * one chunk) with prox 1 = max-col-norm (constraints/maxcolnorm.c:32-61),
 * 2 = non-negative max-col-norm (maxcolnorm-nonneg.c:32-66),
 * 3 = non-negative (nonneg.c), 4 = L1 soft threshold at lam / rho
 * (lasso.c:30-52), 5 = rows onto the probability simplex (simplex.cc:32-104),
 * 6 = column smoothness with multiplier lam (smooth.c:30-184), 7 = columns
 * onto the probability simplex (simplex.cc:120-158), 8 = non-negative with
 * a Frobenius penalty lam (nonneg_frob.c:32-55), 9 = non-negative with an
 * L1 penalty lam (nonneg_lasso.c:25-45), 0 = none. chunk > 0 runs ADMM in
 * independent chunks of
 * that many rows (admm.c:355-400; the caller passes 0 unless the prox is
 * row-separable). */
typedef struct { int kind; int prox; double diag_add; uint64_t chunk; double lam; } or_con;

/* or_cpd_als with per-mode constraints; `unconstrained` is the reference's
 * ws->unconstrained (every constraint UNCONSTRAINED: normalise columns into
 * lambda). max_inner/inner_tol drive the ADMM modes. */
int or_cpd_als_con(or_csf * const * csfs, uint64_t const * mode_csf,
    uint64_t nmodes, uint64_t rank, double ** factors, double * lambda,
    uint64_t max_its, double tol, uint64_t max_inner, double inner_tol,
    or_con const * cons, int unconstrained, double * fits, int nthreads);

/* ADMM for one factor (src/cpd/admm.c:240-308, :346-406 with one chunk):
 * G (R x R, full, including any regularisation) is overwritten by its
 * Cholesky factor after rho = trace(G)/R is added to the diagonal. P (the
 * factor, I x R) and D (its dual) are updated in place. Returns the number
 * of inner iterations; *rho_out receives rho when not NULL. */
uint64_t or_admm(int prox, double const * M, double * P, double * D,
    double * G, uint64_t I, uint64_t R, uint64_t max_inner, double inner_tol,
    double * rho_out, double lam);

/* The same with the rows split into chunks of `chunk` rows (the last one
 * ragged), each iterated to its own convergence (src/cpd/admm.c:355-400).
 * Returns the sum over chunks of iterations x rows. */
uint64_t or_admm_chunked(int prox, double const * M, double * P, double * D,
    double * G, uint64_t I, uint64_t R, uint64_t chunk, uint64_t max_inner,
    double inner_tol, double lam);

/* The smoothness prox (src/cpd/constraints/smooth.c): each column x of P
 * becomes rho (lam K + rho I)^{-1} x, K the reference's banded
 * second-difference matrix. */
void or_smooth_prox(double * P, uint64_t I, uint64_t R, double lam, double rho);

/* The proximal operators (row-major I x R, in place); thr is the lasso
 * threshold lam / rho for prox 4 (lasso.c) and 9 (nonneg_lasso.c), the
 * scale rho / (lam + rho) for prox 8 (nonneg_frob.c), unused by the others. */
void or_prox(int prox, double * P, uint64_t I, uint64_t R, double thr);

/* ---- streaming (the files under src/cpd/stream) --------------------- */

/* Sort in place into lexicographic order of perm (src/sort.c:912-956). */
void or_coo_sort(or_coo * tt, uint64_t const * perm);

/* COO MTTKRP in FP64, src/mttkrp.c:1699-1769. */
void or_mttkrp_stream(or_coo const * tt, double * const * mats, uint64_t rank,
    uint64_t mode, double * out);

/* src/cpd/cpd.c:533-577 and :580-656 (unsquared relative error). */
double or_kruskal_norm(double * const * factors, uint64_t const * dims,
    uint64_t nmodes, uint64_t rank, double const * lambda);
double or_cpd_error(or_coo const * tt, double * const * factors,
    uint64_t const * fdims, uint64_t rank, double const * lambda);

/* ParserSimple.cc:13-80: sort by (t, t+1, ...) cyclically and relabel the
 * other modes by first appearance; iperm[m] (optional) gets new -> old. */
void or_stream_prepare(or_coo * tt, uint64_t stream_mode, uint64_t ** iperm);

typedef struct {
  uint64_t stream_mode;
  uint64_t rank;
  double forget;
  uint64_t max_its;                 /* outer ALS iterations per slice */
  or_con cons[OR_MAX_NMODES];       /* per mode, after the stream's rules */
  uint64_t check_interval;          /* StreamCPD.cc:23-25 (100); 0: none */
  int refit;                        /* full CPD at checkpoints (draws rand()) */
} or_stream_cfg;

typedef struct {
  uint64_t max_batches;             /* capacity of time / local_err / outer_its */
  double * time;                    /* max_batches x rank */
  double * factors[OR_MAX_NMODES];  /* non-stream modes, caller sized */
  double * local_err;               /* per slice, squared (may be NULL) */
  uint64_t * outer_its;             /* per slice (may be NULL) */
  uint64_t nbatches;
  uint64_t dims[OR_MAX_NMODES];
  double final_err;                 /* squared, as printed by the reference */
  /* per slice, -1 off checkpoints (may be NULL): squared errors of all
   * slices so far and of the last 10, the refit's 1 - fit */
  double * global_err;
  double * local10_err;
  double * cpd_err;
} or_stream_out;

/* StreamCPD::compute (StreamCPD.cc:311-571) on a tensor already prepared
 * by or_stream_prepare; the caller seeds rand(). */
int or_stream_cpd(or_coo const * tt, or_stream_cfg const * cfg,
    or_stream_out * out);

/* Dense helpers exposed for pinning tests (src/matrix.c). */
void or_mat_aTa(double const * A, uint64_t I, uint64_t R, double * out);
int or_cholesky(double * G, uint64_t R);            /* in place, lower */
/* Threads of the oracle's row-independent dense loops (results do not
 * depend on the count); default 1. */
void or_set_threads(int n);
int or_get_threads(void);

void or_cholesky_solve(double const * L, uint64_t R, double * B,
    uint64_t nrows);                                /* rows of B */
void or_mat_normalize(double * A, uint64_t I, uint64_t R, double * lambda);

#endif
