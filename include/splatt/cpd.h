/*
 * splatt/cpd.h -- CPD factorisation (ALS / AO-ADMM), constraints.
 */
#ifndef SPLATT_INCLUDE_CPD_H
#define SPLATT_INCLUDE_CPD_H

typedef struct _splatt_kruskal splatt_kruskal;

/* Layouts below follow include/splatt/cpd.h:24-149 field for field. */
typedef struct
{
  bool row_separable;
  bool sparsity_inducing;
} splatt_con_hints;

typedef enum
{
  SPLATT_CON_CLOSEDFORM,
  SPLATT_CON_ADMM
} splatt_con_solve_type;

typedef struct
{
  splatt_con_solve_type solve_type;
  splatt_con_hints hints;
  void * data;
  char description[256];
  void (* init_func) (splatt_val_t * vals, splatt_idx_t const nrows,
                      splatt_idx_t const ncols, void * data);
  void (* gram_func) (splatt_val_t * gram, splatt_idx_t const N, void * data);
  void (* prox_func) (splatt_val_t * primal, splatt_idx_t const nrows,
                      splatt_idx_t const ncols, splatt_idx_t const offset,
                      void * data, splatt_val_t const rho,
                      bool const should_parallelize);
  void (* clsd_func) (splatt_val_t * primal, splatt_idx_t const nrows,
                      splatt_idx_t const ncols, void * data);
  void (* post_func) (splatt_val_t * vals, splatt_idx_t const nrows,
                      splatt_idx_t const ncols, void * data);
  void (* free_func) (void * data);
} splatt_cpd_constraint;

typedef struct
{
  splatt_val_t tolerance;
  splatt_idx_t max_iterations;
  splatt_val_t inner_tolerance;
  splatt_idx_t max_inner_iterations;
  splatt_cpd_constraint * constraints[SPLATT_MAX_NMODES];
  splatt_idx_t chunk_sizes[SPLATT_MAX_NMODES];
} splatt_cpd_opts;

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/splatt/cpd.h:173 (src/cpd/cpd.c:116): factors filled
 * with the reference's glibc rand() sequence, lambda = 1. */
splatt_kruskal * splatt_alloc_cpd(
    splatt_csf const * const tensor,
    splatt_idx_t rank);

/* Replaces include/splatt/cpd.h:183 (src/cpd/cpd.c:144). */
void splatt_free_cpd(splatt_kruskal * factored);

/* Replaces include/splatt/cpd.h:204 (src/cpd/cpd.c:42). The ALS sweep runs
 * on the GPU: MTTKRP, Gram/Hadamard, Cholesky solve, normalisation and fit. */
splatt_error_type splatt_cpd(
    splatt_csf const * const tensor,
    splatt_idx_t rank,
    splatt_cpd_opts const * const cpd_opts,
    splatt_global_opts const * const global_opts,
    splatt_kruskal * factored);

/* Constraint registration: include/splatt/cpd.h:229-336. */
splatt_error_type splatt_register_maxcolnorm(
    splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

splatt_error_type splatt_register_maxcolnorm_nonneg(
    splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

splatt_error_type splatt_register_nonneg(
    splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

splatt_error_type splatt_register_frob(
    splatt_cpd_opts * opts,
    splatt_val_t const multiplier,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

/* include/splatt/cpd.h:297-300 (src/cpd/constraints/simplex.cc:194-216):
 * every column of the factor on the probability simplex (ADMM, one step at
 * a time; the prox couples all rows of a column). */
splatt_error_type splatt_register_colsimp(
    splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

/* include/splatt/cpd.h:349-353 (src/cpd/constraints/smooth.c:220-247):
 * smooth columns (a second-difference penalty with this multiplier), solved
 * by ADMM whose prox is a banded solve along each column. */
splatt_error_type splatt_register_smooth(
    splatt_cpd_opts * opts,
    splatt_val_t const multiplier,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

/* include/splatt/cpd.h:279-282 (src/cpd/constraints/simplex.cc:167-191):
 * every row of the factor on the probability simplex; row-separable, solved
 * by chunked ADMM (this library needs chunk_sizes > 0 for it). */
splatt_error_type splatt_register_rowsimp(
    splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

/* include/splatt/cpd.h:332-336 (src/cpd/constraints/lasso.c:80-107): L1
 * regularisation, solved by ADMM with soft thresholding at multiplier/rho;
 * row-separable, so rows are solved in independent chunks by default. */
splatt_error_type splatt_register_lasso(
    splatt_cpd_opts * opts,
    splatt_val_t const multiplier,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

/* Non-negative factor with a Frobenius penalty (src/cpd/constraints/
 * nonneg_frob.c:74-108) and with an L1 penalty (nonneg_lasso.c:67-99):
 * ADMM, entry-wise proxes max(x rho/(multiplier+rho), 0) and
 * max(x - multiplier/rho, 0); row-separable. The reference registers them
 * from its command line (src/cmds/cmd_cpd_constraint.h:89-90, "ntf-frob"
 * and "ntf-l1") without a declaration in its public header. */
splatt_error_type splatt_register_ntf_frob(
    splatt_cpd_opts * opts,
    splatt_val_t const multiplier,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);
splatt_error_type splatt_register_ntf_lasso(
    splatt_cpd_opts * opts,
    splatt_val_t const multiplier,
    splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes);

/* Options and constraint objects: include/splatt/cpd.h:372-418. */
splatt_cpd_opts * splatt_alloc_cpd_opts(void);
void splatt_free_cpd_opts(splatt_cpd_opts * const opts);
splatt_cpd_constraint * splatt_alloc_constraint(
    splatt_con_solve_type const solve_type);
void splatt_free_constraint(splatt_cpd_constraint * const con);
void splatt_register_constraint(
    splatt_cpd_opts * const opts,
    splatt_idx_t const mode,
    splatt_cpd_constraint * const con);

#ifdef __cplusplus
}
#endif
#endif
