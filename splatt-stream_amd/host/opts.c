/*
 * opts.c -- option arrays, global/CPD options, constraint objects and the
 * constraints this library runs on the GPU (frob, max-col-norm, non-neg).
 *
 * Defaults are the reference's (src/opts.c:10-74, src/base.h:40-44,
 * src/cpd/cpd.c:83-100); constraint semantics follow
 * src/cpd/constraints/{constraints,frob,maxcolnorm,maxcolnorm-nonneg,
 * nonneg}.c. The host bodies of the prox/gram functions below are the
 * definitions the GPU kernels implement; they also serve callers that invoke
 * them directly.
 */
#include "host.h"
#include <math.h>
#include <string.h>
#include <time.h>
#include <omp.h>

double * splatt_default_opts(void)
{
  double * opts = splatt_malloc(SPLATT_OPTION_NOPTIONS * sizeof(*opts));
  if(opts == NULL) {
    return NULL;
  }
  for(int i=0; i < SPLATT_OPTION_NOPTIONS; ++i) {
    opts[i] = SPLATT_VAL_OFF;
  }
  opts[SPLATT_OPTION_TOLERANCE]  = 1e-5;
  opts[SPLATT_OPTION_REGULARIZE] = 0.;
  opts[SPLATT_OPTION_NITER]      = 50;
  opts[SPLATT_OPTION_VERBOSITY]  = SPLATT_VERBOSITY_LOW;
  opts[SPLATT_OPTION_CSF_ALLOC]  = SPLATT_CSF_TWOMODE;
  opts[SPLATT_OPTION_TILE]       = SPLATT_NOTILE;
  opts[SPLATT_OPTION_PRIVTHRESH] = 0.02;
  opts[SPLATT_OPTION_TILELEVEL]  = 1;
  opts[SPLATT_OPTION_DECOMP]     = SPLATT_DECOMP_MEDIUM;
  opts[SPLATT_OPTION_COMM]       = SPLATT_COMM_ALL2ALL;
  opts[SPLATT_OPTION_RANDSEED]   = (double) time(NULL);
  opts[SPLATT_OPTION_NTHREADS]   = omp_in_parallel() ? 1 : omp_get_max_threads();
  return opts;
}

void splatt_free_opts(double * opts)
{
  free(opts);
}

splatt_global_opts * splatt_alloc_global_opts(void)
{
  splatt_global_opts * g = splatt_malloc(sizeof(*g));
  g->num_threads = omp_in_parallel() ? 1 : omp_get_max_threads();
  g->random_seed = (int) time(NULL);
  g->verbosity = SPLATT_VERBOSITY_NONE;
  return g;
}

void splatt_free_global_opts(splatt_global_opts * opts)
{
  splatt_free(opts);
}

/* ---- constraint objects (src/cpd/constraints/constraints.c:13-140) ---- */

static void p_clear_constraint(splatt_cpd_constraint * con)
{
  memset(&con->hints, 0, sizeof(con->hints));
  con->data = NULL;
  con->init_func = NULL;
  con->gram_func = NULL;
  con->prox_func = NULL;
  con->clsd_func = NULL;
  con->post_func = NULL;
  con->free_func = NULL;
  snprintf(con->description, sizeof(con->description), "UNCONSTRAINED");
}

splatt_cpd_constraint * splatt_alloc_constraint(
    splatt_con_solve_type const solve_type)
{
  splatt_cpd_constraint * con = splatt_malloc(sizeof(*con));
  con->solve_type = solve_type;
  p_clear_constraint(con);
  return con;
}

void splatt_free_constraint(splatt_cpd_constraint * const con)
{
  if(con == NULL) {
    return;
  }
  if(con->free_func != NULL) {
    con->free_func(con->data);
  }
  splatt_free(con);
}

void splatt_register_constraint(splatt_cpd_opts * const opts,
    splatt_idx_t const mode, splatt_cpd_constraint * const con)
{
  splatt_free_constraint(opts->constraints[mode]);
  opts->constraints[mode] = con;
}

splatt_cpd_opts * splatt_alloc_cpd_opts(void)
{
  splatt_cpd_opts * opts = splatt_malloc(sizeof(*opts));
  opts->tolerance = 1e-5;
  opts->max_iterations = 200;
  opts->inner_tolerance = 1e-2;
  opts->max_inner_iterations = 50;
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    opts->chunk_sizes[m] = 50;
    opts->constraints[m] = splatt_alloc_constraint(SPLATT_CON_CLOSEDFORM);
  }
  return opts;
}

void splatt_free_cpd_opts(splatt_cpd_opts * const opts)
{
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    splatt_free_constraint(opts->constraints[m]);
  }
  splatt_free(opts);
}

/* ---- Frobenius (Tikhonov): lambda added to the Gram diagonal ---------- */

void splatt_frob_gram(val_t * gram, idx_t const N, void * data)
{
  val_t const mult = *((val_t *) data);
  for(idx_t n=0; n < N; ++n) {
    gram[n + n * N] += mult;
  }
}

static void p_free_data(void * data)
{
  splatt_free(data);
}

splatt_error_type splatt_register_frob(splatt_cpd_opts * opts,
    splatt_val_t const multiplier, splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes)
{
  for(idx_t m=0; m < num_modes; ++m) {
    splatt_cpd_constraint * con = splatt_alloc_constraint(SPLATT_CON_CLOSEDFORM);
    con->gram_func = splatt_frob_gram;
    con->free_func = p_free_data;
    con->hints.row_separable = true;
    snprintf(con->description, sizeof(con->description),
        "FROBENIUS-REG (%0.1e)", multiplier);
    val_t * mult = splatt_malloc(sizeof(*mult));
    *mult = multiplier;
    con->data = mult;
    splatt_register_constraint(opts, modes_included[m], con);
  }
  return SPLATT_SUCCESS;
}

/* ---- column norms <= 1 -------------------------------------------------- */

void splatt_maxcolnorm_prox(val_t * primal, idx_t const nrows,
    idx_t const ncols, idx_t const offset, void * data, val_t const rho,
    bool const should_parallelize)
{
  (void) offset; (void) data; (void) rho; (void) should_parallelize;
  for(idx_t j=0; j < ncols; ++j) {
    double norm = 0.;
    for(idx_t i=0; i < nrows; ++i) {
      norm += primal[j + i * ncols] * primal[j + i * ncols];
    }
    norm = sqrt(norm);
    if(norm <= 1.) {
      continue;
    }
    for(idx_t i=0; i < nrows; ++i) {
      primal[j + i * ncols] /= norm;
    }
  }
}

void splatt_maxcolnorm_nonneg_prox(val_t * primal, idx_t const nrows,
    idx_t const ncols, idx_t const offset, void * data, val_t const rho,
    bool const should_parallelize)
{
  (void) offset; (void) data; (void) rho; (void) should_parallelize;
  for(idx_t j=0; j < ncols; ++j) {
    double norm = 0.;
    for(idx_t i=0; i < nrows; ++i) {
      val_t * p = primal + j + i * ncols;
      if(*p < 0.) {
        *p = 0.;
      } else {
        norm += (*p) * (*p);
      }
    }
    norm = sqrt(norm);
    if(norm <= 1.) {
      continue;
    }
    for(idx_t i=0; i < nrows; ++i) {
      primal[j + i * ncols] /= norm;
    }
  }
}

void splatt_nonneg_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset; (void) data; (void) rho; (void) should_parallelize;
  for(idx_t x=0; x < nrows * ncols; ++x) {
    primal[x] = (primal[x] > 0.) ? primal[x] : 0.;
  }
}

/* src/cpd/constraints/lasso.c:30-52: soft thresholding at lambda / rho. */
void splatt_lasso_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset; (void) should_parallelize;
  val_t const mult = *((val_t const *) data) / rho;
  for(idx_t x=0; x < nrows * ncols; ++x) {
    val_t const v = primal[x];
    if(v > mult) {
      primal[x] = v - mult;
    } else if(v < -mult) {
      primal[x] = v + mult;
    } else {
      primal[x] = 0.;
    }
  }
}

static int p_cmp_desc(void const * a, void const * b)
{
  val_t const x = *(val_t const *) a;
  val_t const y = *(val_t const *) b;
  return (x < y) ? 1 : ((x > y) ? -1 : 0);
}

/* src/cpd/constraints/simplex.cc:32-104: every row onto the probability
 * simplex (sort a copy in non-increasing order, threshold from the running
 * sum, clamp). */
void splatt_rowsimp_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset; (void) data; (void) rho; (void) should_parallelize;
  val_t * buf = (val_t *) malloc((ncols ? ncols : 1) * sizeof(*buf));
  if(buf == NULL) {
    return;
  }
  for(idx_t i=0; i < nrows; ++i) {
    val_t * const row = primal + i * ncols;
    memcpy(buf, row, ncols * sizeof(*buf));
    qsort(buf, ncols, sizeof(*buf), p_cmp_desc);
    val_t run = -1.;
    idx_t piv;
    for(piv=0; piv < ncols; ++piv) {
      run += buf[piv];
      val_t const v = buf[piv] - (run / (val_t) (piv + 1));
      if(v <= 0.) {
        run -= buf[piv];
        break;
      }
    }
    val_t const theta = run / (val_t) piv;
    for(idx_t j=0; j < ncols; ++j) {
      row[j] -= theta;
      row[j] = (row[j] > 0.) ? row[j] : 0.;
    }
  }
  free(buf);
}

/* src/cpd/constraints/smooth.c:30-184: every column x of the factor
 * becomes rho (lambda K + rho I)^{-1} x, K the banded second-difference
 * matrix of p_form_banded; symmetric positive definite, solved here by a
 * banded L D L^T factorisation. */
void splatt_smooth_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset; (void) should_parallelize;
  val_t const lam = *((val_t const *) data);
  if(nrows == 0) {
    return;
  }
  val_t * f = (val_t *) malloc(3 * nrows * sizeof(*f));
  if(f == NULL) {
    return;
  }
  val_t dm2 = 0., dm1 = 0., l1m1 = 0.;
  for(idx_t i=0; i < nrows; ++i) {
    val_t const a = ((i == 0 || i == nrows - 1) ? 5. * lam : 6. * lam) + rho;
    val_t const l2 = (i >= 2) ? lam / dm2 : 0.;
    val_t const l1 = (i >= 1) ? (-4. * lam - l2 * dm2 * l1m1) / dm1 : 0.;
    val_t const d = a - l2 * l2 * dm2 - l1 * l1 * dm1;
    f[3 * i] = d;
    f[3 * i + 1] = l1;
    f[3 * i + 2] = l2;
    dm2 = dm1;
    dm1 = d;
    l1m1 = l1;
  }
  for(idx_t c=0; c < ncols; ++c) {
    val_t y1 = 0., y2 = 0.;
    for(idx_t i=0; i < nrows; ++i) {
      val_t const y = primal[i * ncols + c] - f[3 * i + 1] * y1 - f[3 * i + 2] * y2;
      primal[i * ncols + c] = y / f[3 * i];
      y2 = y1;
      y1 = y;
    }
    val_t x1 = 0., x2 = 0.;
    for(idx_t i=nrows; i-- > 0;) {
      val_t const l1n = (i + 1 < nrows) ? f[3 * (i + 1) + 1] : 0.;
      val_t const l2n = (i + 2 < nrows) ? f[3 * (i + 2) + 2] : 0.;
      val_t const x = primal[i * ncols + c] - l1n * x1 - l2n * x2;
      primal[i * ncols + c] = x * rho;
      x2 = x1;
      x1 = x;
    }
  }
  free(f);
}

/* src/cpd/constraints/simplex.cc:120-158: every column onto the probability
 * simplex (the row projection applied to a copy of the column). */
void splatt_colsimp_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset; (void) data; (void) rho; (void) should_parallelize;
  val_t * col = (val_t *) malloc((nrows ? nrows : 1) * sizeof(*col));
  if(col == NULL) {
    return;
  }
  for(idx_t j=0; j < ncols; ++j) {
    for(idx_t i=0; i < nrows; ++i) {
      col[i] = primal[j + i * ncols];
    }
    splatt_rowsimp_prox(col, 1, nrows, 0, NULL, rho, false);
    for(idx_t i=0; i < nrows; ++i) {
      primal[j + i * ncols] = col[i];
    }
  }
  free(col);
}

static void p_lasso_free(void * data)
{
  free(data);
}

/* src/cpd/constraints/nonneg_frob.c:32-55: every entry scaled by
 * rho / (lambda + rho), then clamped at zero. */
void splatt_ntf_frob_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset;
  (void) should_parallelize;
  val_t const mult = rho / (*((val_t const *) data) + rho);
  for(idx_t x=0; x < nrows * ncols; ++x) {
    val_t const v = primal[x] * mult;
    primal[x] = (v > 0.) ? v : 0.;
  }
}

/* src/cpd/constraints/nonneg_lasso.c:25-45: one-sided soft threshold at
 * lambda / rho (entries below it become zero). */
void splatt_ntf_lasso_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize)
{
  (void) offset;
  (void) should_parallelize;
  val_t const thr = *((val_t const *) data) / rho;
  for(idx_t x=0; x < nrows * ncols; ++x) {
    val_t const v = primal[x];
    primal[x] = (v > thr) ? v - thr : 0.;
  }
}

/* One row-separable, sparsity-inducing ADMM constraint per mode holding its
 * own copy of the multiplier (the shape of nonneg_frob.c:74-108 and
 * nonneg_lasso.c:67-99, which both describe themselves as NTF-L1-REG). */
static splatt_error_type p_register_mult(splatt_cpd_opts * opts, splatt_val_t multiplier,
    splatt_idx_t const * modes, splatt_idx_t num,
    void (*prox)(val_t *, idx_t const, idx_t const, idx_t const, void *,
        val_t const, bool const), char const * fmt)
{
  for(idx_t m=0; m < num; ++m) {
    splatt_cpd_constraint * con = splatt_alloc_constraint(SPLATT_CON_ADMM);
    con->prox_func = prox;
    con->free_func = p_lasso_free;
    con->hints.row_separable = true;
    con->hints.sparsity_inducing = true;
    snprintf(con->description, sizeof(con->description), fmt, multiplier);
    val_t * mult = (val_t *) malloc(sizeof(*mult));
    if(mult == NULL) {
      splatt_free_constraint(con);
      return SPLATT_ERROR_NOMEMORY;
    }
    *mult = multiplier;
    con->data = mult;
    splatt_register_constraint(opts, modes[m], con);
  }
  return SPLATT_SUCCESS;
}

splatt_error_type splatt_register_ntf_frob(splatt_cpd_opts * opts,
    splatt_val_t const multiplier, splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes)
{
  return p_register_mult(opts, multiplier, modes_included, num_modes,
      splatt_ntf_frob_prox, "NTF-L1-REG (%0.1e)");
}

splatt_error_type splatt_register_ntf_lasso(splatt_cpd_opts * opts,
    splatt_val_t const multiplier, splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes)
{
  return p_register_mult(opts, multiplier, modes_included, num_modes,
      splatt_ntf_lasso_prox, "NTF-L1-REG (%0.1e)");
}

static splatt_error_type p_register_prox(splatt_cpd_opts * opts,
    splatt_idx_t const * modes, splatt_idx_t num,
    void (*prox)(val_t *, idx_t const, idx_t const, idx_t const, void *,
        val_t const, bool const),
    bool row_sep, bool sparse, char const * desc)
{
  for(idx_t m=0; m < num; ++m) {
    splatt_cpd_constraint * con = splatt_alloc_constraint(SPLATT_CON_ADMM);
    con->prox_func = prox;
    con->hints.row_separable = row_sep;
    con->hints.sparsity_inducing = sparse;
    snprintf(con->description, sizeof(con->description), "%s", desc);
    splatt_register_constraint(opts, modes[m], con);
  }
  return SPLATT_SUCCESS;
}

splatt_error_type splatt_register_maxcolnorm(splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included, splatt_idx_t const num_modes)
{
  return p_register_prox(opts, modes_included, num_modes,
      splatt_maxcolnorm_prox, false, false, "MAX-COL-NORM");
}

splatt_error_type splatt_register_maxcolnorm_nonneg(splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included, splatt_idx_t const num_modes)
{
  return p_register_prox(opts, modes_included, num_modes,
      splatt_maxcolnorm_nonneg_prox, false, true, "NONNEG-MAX-COL-NORM");
}

splatt_error_type splatt_register_nonneg(splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included, splatt_idx_t const num_modes)
{
  return p_register_prox(opts, modes_included, num_modes,
      splatt_nonneg_prox, true, true, "NON-NEGATIVE");
}

/* src/cpd/constraints/lasso.c:80-107: one ADMM constraint per mode holding
 * its own copy of the multiplier; row-separable and sparsity-inducing. */
splatt_error_type splatt_register_lasso(splatt_cpd_opts * opts,
    splatt_val_t const multiplier, splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes)
{
  for(idx_t m=0; m < num_modes; ++m) {
    splatt_cpd_constraint * con = splatt_alloc_constraint(SPLATT_CON_ADMM);
    con->prox_func = splatt_lasso_prox;
    con->free_func = p_lasso_free;
    con->hints.row_separable = true;
    con->hints.sparsity_inducing = true;
    snprintf(con->description, sizeof(con->description), "L1-REG (%0.1e)", multiplier);
    val_t * mult = (val_t *) malloc(sizeof(*mult));
    if(mult == NULL) {
      splatt_free_constraint(con);
      return SPLATT_ERROR_NOMEMORY;
    }
    *mult = multiplier;
    con->data = mult;
    splatt_register_constraint(opts, modes_included[m], con);
  }
  return SPLATT_SUCCESS;
}

/* src/cpd/constraints/simplex.cc:167-191: row simplex, row-separable (its
 * ADMM runs in chunks of rows; chunk_sizes must be positive). */
splatt_error_type splatt_register_rowsimp(splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included, splatt_idx_t const num_modes)
{
  return p_register_prox(opts, modes_included, num_modes,
      splatt_rowsimp_prox, true, true, "ROW-SIMPLEX");
}

/* src/cpd/constraints/smooth.c:220-247: one ADMM constraint per mode holding
 * its multiplier; not row-separable (the prox couples the rows of a
 * column). */
splatt_error_type splatt_register_smooth(splatt_cpd_opts * opts,
    splatt_val_t const multiplier, splatt_idx_t const * const modes_included,
    splatt_idx_t const num_modes)
{
  for(idx_t m=0; m < num_modes; ++m) {
    splatt_cpd_constraint * con = splatt_alloc_constraint(SPLATT_CON_ADMM);
    con->prox_func = splatt_smooth_prox;
    con->free_func = p_lasso_free;
    snprintf(con->description, sizeof(con->description), "SMOOTH-COL (%0.1e)", multiplier);
    val_t * mult = (val_t *) malloc(sizeof(*mult));
    if(mult == NULL) {
      splatt_free_constraint(con);
      return SPLATT_ERROR_NOMEMORY;
    }
    *mult = multiplier;
    con->data = mult;
    splatt_register_constraint(opts, modes_included[m], con);
  }
  return SPLATT_SUCCESS;
}

/* src/cpd/constraints/simplex.cc:194-216: column simplex; not row-separable
 * (the prox couples the rows of a column). */
splatt_error_type splatt_register_colsimp(splatt_cpd_opts * opts,
    splatt_idx_t const * const modes_included, splatt_idx_t const num_modes)
{
  return p_register_prox(opts, modes_included, num_modes,
      splatt_colsimp_prox, false, true, "COL-SIMPLEX");
}
