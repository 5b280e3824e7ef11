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

/* Dense helpers exposed for pinning tests (src/matrix.c). */
void or_mat_aTa(double const * A, uint64_t I, uint64_t R, double * out);
int or_cholesky(double * G, uint64_t R);            /* in place, lower */
void or_cholesky_solve(double const * L, uint64_t R, double * B,
    uint64_t nrows);                                /* rows of B */
void or_mat_normalize(double * A, uint64_t I, uint64_t R, double * lambda);

#endif
