/*
 * cpd.c -- splatt_alloc_cpd / splatt_free_cpd / splatt_cpd
 * (src/cpd/cpd.c:42-141) over the GPU ALS driver.
 */
#include "host.h"

/* src/cpd/cpd.c:116-141: factors in mode order, each filled with the
 * glibc rand() sequence of src/util.c:15-40; lambda = 1. */
splatt_kruskal * splatt_alloc_cpd(splatt_csf const * const csf,
    splatt_idx_t rank)
{
  splatt_kruskal * cpd = splatt_malloc(sizeof(*cpd));
  cpd->nmodes = csf->nmodes;
  cpd->rank = rank;
  cpd->fit = 0.;
  cpd->lambda = splatt_malloc(rank * sizeof(*cpd->lambda));
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    cpd->factors[m] = NULL;
    cpd->dims[m] = 0;
  }
  for(idx_t m=0; m < csf->nmodes; ++m) {
    cpd->dims[m] = csf->dims[m];
    cpd->factors[m] = splatt_malloc(csf->dims[m] * rank * sizeof(val_t));
    spl_fill_rand(cpd->factors[m], csf->dims[m] * rank);
  }
  for(idx_t r=0; r < rank; ++r) {
    cpd->lambda[r] = 1.;
  }
  return cpd;
}

void splatt_free_cpd(splatt_kruskal * factored)
{
  if(factored == NULL) {
    return;
  }
  splatt_free(factored->lambda);
  for(idx_t m=0; m < factored->nmodes; ++m) {
    splatt_free(factored->factors[m]);
  }
  splatt_free(factored);
}

/* Translate a registered constraint into what the device driver runs.
 * Returns 0, or -1 for a constraint the device path does not provide. */
int spl_map_constraint(splatt_cpd_constraint const * con, idx_t mode,
    idx_t chunk, spd_con * out)
{
  out->kind = SPD_CON_NONE;
  out->prox = SPD_PROX_NONE;
  out->diag_add = 0.;
  out->chunk = 0;
  out->lam = 0.;
  if(con->gram_func == splatt_frob_gram) {
    out->diag_add = *((val_t const *) con->data);
  } else if(con->gram_func != NULL) {
    fprintf(stderr, "SPLATT: constraint '%s' (mode %" SPLATT_PF_IDX ") has a Gram "
        "callback this library cannot run on the GPU.\n", con->description, mode);
    return -1;
  }
  if(con->solve_type == SPLATT_CON_CLOSEDFORM) {
    if(con->clsd_func != NULL) {
      fprintf(stderr, "SPLATT: closed-form callback of '%s' is not supported.\n",
          con->description);
      return -1;
    }
    out->kind = (strcmp(con->description, "UNCONSTRAINED") == 0) ?
        SPD_CON_NONE : SPD_CON_CLOSED;
    return 0;
  }
  out->kind = SPD_CON_ADMM;
  if(con->prox_func == splatt_maxcolnorm_prox) {
    out->prox = SPD_PROX_MAXCOLNORM;
  } else if(con->prox_func == splatt_maxcolnorm_nonneg_prox) {
    out->prox = SPD_PROX_MAXCOLNORM_NONNEG;
  } else if(con->prox_func == splatt_nonneg_prox) {
    out->prox = SPD_PROX_NONNEG;
  } else if(con->prox_func == splatt_lasso_prox) {
    out->prox = SPD_PROX_LASSO;
    out->lam = *((val_t const *) con->data);
  } else if(con->prox_func == splatt_ntf_frob_prox) {
    out->prox = SPD_PROX_NTF_FROB;
    out->lam = *((val_t const *) con->data);
  } else if(con->prox_func == splatt_ntf_lasso_prox) {
    out->prox = SPD_PROX_NTF_LASSO;
    out->lam = *((val_t const *) con->data);
  } else if(con->prox_func == splatt_colsimp_prox) {
    out->prox = SPD_PROX_COLSIMP;
  } else if(con->prox_func == splatt_smooth_prox) {
    out->prox = SPD_PROX_SMOOTH;
    out->lam = *((val_t const *) con->data);
  } else if(con->prox_func == splatt_rowsimp_prox) {
    out->prox = SPD_PROX_ROWSIMP;
    if(chunk == 0) {
      fprintf(stderr, "SPLATT: '%s' (mode %" SPLATT_PF_IDX ") needs chunked ADMM "
          "(chunk_sizes > 0) on the GPU.\n", con->description, mode);
      return -1;
    }
  } else {
    fprintf(stderr, "SPLATT: proximity operator of '%s' is not supported.\n",
        con->description);
    return -1;
  }
  /* src/cpd/admm.c:356-366: row-separable constraints are solved in
   * independent chunks of `chunk` rows */
  if(con->hints.row_separable && chunk > 0) {
    out->chunk = chunk;
  }
  return 0;
}

splatt_error_type splatt_cpd(splatt_csf const * const tensor,
    splatt_idx_t rank, splatt_cpd_opts const * const cpd_options,
    splatt_global_opts const * const global_options, splatt_kruskal * factored)
{
  splatt_global_opts * gopts = (splatt_global_opts *) global_options;
  if(gopts == NULL) {
    gopts = splatt_alloc_global_opts();
  }
  splatt_cpd_opts * copts = (splatt_cpd_opts *) cpd_options;
  if(copts == NULL) {
    copts = splatt_alloc_cpd_opts();
  }
  idx_t const N = tensor->nmodes;
  splatt_error_type ret = SPLATT_SUCCESS;

  spd_cpd_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.nmodes = (int) N;
  cfg.rank = rank;
  cfg.max_its = copts->max_iterations;
  cfg.tol = copts->tolerance;
  cfg.max_inner = copts->max_inner_iterations;
  cfg.inner_tol = copts->inner_tolerance;
  cfg.dtype = spl_mttkrp_dtype();
  cfg.verbosity = (int) gopts->verbosity;
  cfg.unconstrained = 1;
  for(idx_t m=0; m < N; ++m) {
    splatt_cpd_constraint const * con = copts->constraints[m];
    if(strcmp(con->description, "UNCONSTRAINED") != 0) {
      cfg.unconstrained = 0;
    }
    if(spl_map_constraint(con, m, copts->chunk_sizes[m], &cfg.con[m]) != 0) {
      ret = SPLATT_ERROR_BADINPUT;
    }
  }

  /* src/cpd/cpd.c:206-210: the MTTKRP workspace comes from the default
   * options (TWOMODE): the leaf mode of tensor[0] uses tensor[1]. A
   * single-CSF array uses tensor[0] for every mode. */
  spd_csf * devs[SPLATT_MAX_NMODES];
  int rows_pos[SPLATT_MAX_NMODES];
  int mode_csf[SPLATT_MAX_NMODES];
  idx_t const count = (tensor->alloc_count > 0) ? tensor->alloc_count : 1;
  for(idx_t k=0; k < count && k < SPLATT_MAX_NMODES; ++k) {
    devs[k] = (spd_csf *) tensor[k].dev;
    rows_pos[k] = (tensor[k].ntiles == 1);
    if(devs[k] == NULL) {
      fprintf(stderr, "SPLATT: CPD needs a tensor built by this library.\n");
      ret = SPLATT_ERROR_BADINPUT;
    }
  }
  for(idx_t m=0; m < N; ++m) {
    mode_csf[m] = (count >= 2 && splatt_csf_mode_to_depth(tensor, m) == N - 1) ? 1 : 0;
  }

  if(ret == SPLATT_SUCCESS) {
    uint64_t its = 0;
    double fit = 0.;
    int rc = spd_cpd_als(devs, mode_csf, rows_pos, &cfg, factored->factors,
        factored->lambda, NULL, &its, &fit);
    if(rc != SPD_OK) {
      ret = spl_report(rc, "splatt_cpd");
    } else {
      factored->fit = fit;
    }
  }

  if(global_options == NULL) {
    splatt_free_global_opts(gopts);
  }
  if(cpd_options == NULL) {
    splatt_free_cpd_opts(copts);
  }
  return ret;
}
