/*
 * stream.c -- streaming CPD entry points: splatt_cpd_stream
 * (src/cpd/stream/stream.cc:48-80, the entry behind `splatt cpd --stream`)
 * and the in-memory splatt_amd_cpd_stream_coo.
 *
 * The parser step (ParserSimple.cc:13-80) sorts the stream by the streamed
 * mode, then the following modes cyclically, and relabels every other mode
 * by first appearance; the sort runs on the GPU (spd_coo_lexsort), the
 * relabelling is one host pass. Everything per slice runs on the GPU
 * (spd_stream_cpd).
 */
#include "host.h"
#include <math.h>

/* src/cpd/stream/stream.cc:22-40 */
static void p_print_stream_stats(char const * fname, idx_t rank, idx_t stream_mode,
    double forget, splatt_cpd_opts const * copts, splatt_global_opts const * gopts)
{
  printf("Streaming------------------------------------------------------\n");
  printf("STREAM-FILE=%s\n", fname ? fname : "(memory)");
  printf("STREAM-MODE=%" SPLATT_PF_IDX " ", stream_mode + 1);
  printf("STREAM-FORGET=%f ", forget);
  printf("RANK=%" SPLATT_PF_IDX " ", rank);
  printf("MAXITS=%" SPLATT_PF_IDX " ", copts->max_iterations);
  printf("TOL=%0.1e ", copts->tolerance);
  printf("SEED=%d ", gopts->random_seed);
  printf("THREADS=%d\n", gopts->num_threads);
}

void splatt_amd_free_stream_result(splatt_amd_stream_result * r)
{
  if(r == NULL) {
    return;
  }
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    free(r->factors[m]);
    free(r->iperm[m]);
  }
  free(r->local_err);
  free(r->batch_seconds);
  free(r->outer_its);
  free(r->global_err);
  free(r->local10_err);
  free(r->cpd_err);
  free(r);
}

static void p_fill_rand(void * ctx, double * vals, uint64_t n)
{
  (void) ctx;
  spl_fill_rand(vals, n);
}

/* StreamCPD::compute_cpd_errorsq (StreamCPD.cc:89-106): a CSF with the
 * default options, factors from rand() (splatt_alloc_cpd), splatt_cpd with
 * the default options; the relative squared error 1 - fit. */
static double p_refit(void * ctx, int nmodes, uint64_t nnz, uint64_t const * dims,
    uint32_t const * const * inds, double const * vals, uint64_t rank)
{
  (void) ctx;
  sptensor_t * tt = splatt_tt_alloc(nnz, (idx_t) nmodes);
  if(tt == NULL) {
    return NAN;
  }
  for(int m=0; m < nmodes; ++m) {
    tt->dims[m] = dims[m];
    for(uint64_t x=0; x < nnz; ++x) {
      tt->ind[m][x] = inds[m][x];
    }
  }
  memcpy(tt->vals, vals, nnz * sizeof(val_t));
  double * opts = splatt_default_opts();
  splatt_csf * csf = splatt_csf_alloc(tt, opts);
  splatt_tt_free(tt);
  double err = NAN;
  if(csf != NULL) {
    splatt_kruskal * k = splatt_alloc_cpd(csf, rank);
    if(k != NULL && splatt_cpd(csf, rank, NULL, NULL, k) == SPLATT_SUCCESS) {
      err = 1. - k->fit;
    }
    if(k != NULL) {
      splatt_free_cpd(k);
    }
    splatt_free_csf(csf, opts);
  }
  splatt_free_opts(opts);
  return err;
}

splatt_amd_stream_result * splatt_amd_cpd_stream_coo(
    sptensor_t const * const tt, splatt_idx_t rank, splatt_idx_t stream_mode,
    double forget, splatt_idx_t check_interval, int refit,
    splatt_cpd_opts * const cpd_options,
    splatt_global_opts const * const global_options)
{
  idx_t const N = tt->nmodes;
  idx_t const nnz = tt->nnz;
  if(stream_mode >= N) {
    fprintf(stderr, "ERROR: streaming mode %" SPLATT_PF_IDX
        " is larger than # modes.\n", stream_mode + 1);
    return NULL;
  }
  if(nnz == 0 || nnz >= (1ull << 32)) {
    fprintf(stderr, "SPLATT: stream needs 1 .. 2^32-1 nonzeros.\n");
    return NULL;
  }
  splatt_global_opts * gopts = (splatt_global_opts *) global_options;
  if(gopts == NULL) {
    gopts = splatt_alloc_global_opts();
  }
  splatt_cpd_opts * copts = cpd_options;
  if(copts == NULL) {
    copts = splatt_alloc_cpd_opts();
  }

  /* StreamCPD.cc:326-342: non-negative -> non-negative max-column-norm,
   * unconstrained -> max-column-norm, for every mode but the streamed one */
  for(idx_t m=0; m < N; ++m) {
    if(m == stream_mode) continue;
    if(strcmp(copts->constraints[m]->description, "NON-NEGATIVE") == 0) {
      splatt_register_maxcolnorm_nonneg(copts, &m, 1);
    } else if(strcmp(copts->constraints[m]->description, "UNCONSTRAINED") == 0) {
      splatt_register_maxcolnorm(copts, &m, 1);
    }
  }
  /* StreamCPD.cc:366-368 */
  copts->tolerance = 5e-2;
  copts->max_inner_iterations = 25;
  copts->inner_tolerance = 7e-2;

  splatt_amd_stream_result * res = NULL;
  spd_stream_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.stream_mode = (int) stream_mode;
  cfg.rank = rank;
  cfg.forget = forget;
  cfg.max_its = copts->max_iterations;
  cfg.tol = copts->tolerance;
  cfg.max_inner = copts->max_inner_iterations;
  cfg.inner_tol = copts->inner_tolerance;
  cfg.verbosity = (int) gopts->verbosity;
  cfg.fill_rand = p_fill_rand;
  cfg.check_interval = check_interval;
  cfg.refit = refit ? p_refit : NULL;
  int ok = 1;
  for(idx_t m=0; m < N; ++m) {
    if(spl_map_constraint(copts->constraints[m], m, 0, &cfg.con[m]) != 0) {
      ok = 0;
    }
  }
  if(cfg.con[stream_mode].kind == SPD_CON_ADMM) {
    fprintf(stderr, "SPLATT: the streamed mode takes closed-form constraints only.\n");
    ok = 0;
  }

  /* ---- parser: order on the GPU, relabel on the host ---- */
  uint32_t * order = NULL;
  uint32_t * sind[SPLATT_MAX_NMODES] = {0};
  double * svals = NULL;
  spd_coo * dcoo = NULL;
  if(ok) {
    int perm[SPLATT_MAX_NMODES];
    for(idx_t m=0; m < N; ++m) {
      perm[m] = (int) ((stream_mode + m) % N);
    }
    order = malloc(nnz * sizeof(uint32_t));
    int rc = spd_coo_upload((int) N, nnz, tt->dims, (uint64_t const * const *) tt->ind,
        tt->vals, &dcoo);
    if(rc == SPD_OK) rc = spd_coo_lexsort(dcoo, perm, order);
    spd_coo_free(dcoo);
    if(rc != SPD_OK) {
      spl_report(rc, "stream sort");
      ok = 0;
    }
  }
  if(ok) {
    res = calloc(1, sizeof(*res));
    res->nmodes = N;
    res->rank = rank;
    res->stream_mode = stream_mode;
    svals = malloc(nnz * sizeof(double));
    for(idx_t x=0; x < nnz; ++x) {
      svals[x] = tt->vals[order[x]];
    }
    for(idx_t m=0; m < N; ++m) {
      sind[m] = malloc(nnz * sizeof(uint32_t));
      idx_t const D = tt->dims[m];
      res->iperm[m] = malloc((D ? D : 1) * sizeof(idx_t));
      if(m == stream_mode) {
        for(idx_t x=0; x < nnz; ++x) sind[m][x] = (uint32_t) tt->ind[m][order[x]];
        for(idx_t i=0; i < D; ++i) res->iperm[m][i] = i;
        continue;
      }
      idx_t * lab = malloc((D ? D : 1) * sizeof(idx_t));
      for(idx_t i=0; i < D; ++i) {
        lab[i] = SPLATT_IDX_MAX;
        res->iperm[m][i] = SPLATT_IDX_MAX;
      }
      idx_t seen = 0;
      for(idx_t x=0; x < nnz; ++x) {
        idx_t const ind = tt->ind[m][order[x]];
        if(lab[ind] == SPLATT_IDX_MAX) {
          lab[ind] = seen;
          res->iperm[m][seen] = ind;
          ++seen;
        }
        sind[m][x] = (uint32_t) lab[ind];
      }
      free(lab);
    }
    for(idx_t m=0; m < N; ++m) {
      idx_t const rows = tt->dims[m];
      res->factors[m] = calloc((rows ? rows : 1) * rank, sizeof(val_t));
    }
    idx_t const T = tt->dims[stream_mode];
    res->local_err = calloc(T ? T : 1, sizeof(double));
    res->batch_seconds = calloc(T ? T : 1, sizeof(double));
    res->outer_its = calloc(T ? T : 1, sizeof(idx_t));
    res->global_err = calloc(T ? T : 1, sizeof(double));
    res->local10_err = calloc(T ? T : 1, sizeof(double));
    res->cpd_err = calloc(T ? T : 1, sizeof(double));

    spd_stream_result out;
    memset(&out, 0, sizeof(out));
    for(idx_t m=0; m < N; ++m) out.factors[m] = res->factors[m];
    out.local_err = res->local_err;
    out.batch_seconds = res->batch_seconds;
    out.outer_its = (uint64_t *) res->outer_its;
    out.final_err = &res->final_err;
    out.global_err = res->global_err;
    out.local10_err = res->local10_err;
    out.cpd_err = res->cpd_err;
    int rc = spd_stream_cpd((int) N, nnz, tt->dims, (uint32_t const * const *) sind,
        svals, &cfg, &out);
    if(rc != SPD_OK) {
      spl_report(rc, "splatt_cpd_stream");
      splatt_amd_free_stream_result(res);
      res = NULL;
    } else {
      res->nbatches = out.nbatches;
      for(idx_t m=0; m < N; ++m) res->dims[m] = out.dims[m];
    }
  }

  for(idx_t m=0; m < N; ++m) free(sind[m]);
  free(svals);
  free(order);
  if(global_options == NULL) {
    splatt_free_global_opts(gopts);
  }
  if(cpd_options == NULL) {
    splatt_free_cpd_opts(copts);
  }
  return res;
}

/* mat_write (src/io.c:692-738) of rows placed at their original labels
 * (perm_matrix with the parser's iperm, src/reorder.c:557-585). */
static void p_write_unpermuted(char const * fname, val_t const * vals, idx_t rows,
    idx_t outrows, idx_t rank, idx_t const * iperm)
{
  val_t * out = calloc((outrows ? outrows : 1) * rank, sizeof(val_t));
  for(idx_t i=0; i < rows; ++i) {
    idx_t const dst = iperm[i];
    if(dst < outrows) {
      memcpy(out + dst * rank, vals + i * rank, rank * sizeof(val_t));
    }
  }
  FILE * fout = fopen(fname, "w");
  if(fout == NULL) {
    fprintf(stderr, "SPLATT ERROR: failed to open '%s'\n.", fname);
    free(out);
    return;
  }
  for(idx_t i=0; i < outrows; ++i) {
    for(idx_t j=0; j < rank; ++j) {
      fprintf(fout, "%+0.8le ", out[i * rank + j]);
    }
    fprintf(fout, "\n");
  }
  fclose(fout);
  free(out);
}

splatt_error_type splatt_cpd_stream(char const * const filename,
    splatt_idx_t rank, splatt_idx_t const stream_mode, double const forget,
    splatt_cpd_opts * const cpd_options,
    splatt_global_opts const * const global_options)
{
  splatt_global_opts * gopts = (splatt_global_opts *) global_options;
  if(gopts == NULL) {
    gopts = splatt_alloc_global_opts();
  }
  splatt_cpd_opts * copts = cpd_options;
  if(copts == NULL) {
    copts = splatt_alloc_cpd_opts();
  }
  p_print_stream_stats(filename, rank, stream_mode, forget, copts, gopts);
  splatt_error_type ret = SPLATT_SUCCESS;
  sptensor_t * tt = splatt_tt_read(filename);
  splatt_amd_stream_result * res = NULL;
  if(tt == NULL) {
    ret = SPLATT_ERROR_BADINPUT;
  } else {
    /* the reference prints every batch line and the outer iterations'
     * deltas (StreamCPD.cc:472-477, :530-534) whatever the verbosity;
     * CHECK_ERR_INTERVAL is 100 (StreamCPD.cc:23-25) */
    splatt_global_opts loud = *gopts;
    if(loud.verbosity < SPLATT_VERBOSITY_HIGH) {
      loud.verbosity = SPLATT_VERBOSITY_HIGH;
    }
    res = splatt_amd_cpd_stream_coo(tt, rank, stream_mode, forget, 100, 1, copts, &loud);
    if(res == NULL) {
      ret = SPLATT_ERROR_BADINPUT;
    }
  }
  if(res != NULL) {
    double total = 0.;
    for(idx_t b=0; b < res->nbatches; ++b) total += res->batch_seconds[b];
    printf("stream-time: %0.3fs\n", total);
    printf("\n");
    printf("final-err: %0.5f\n", res->final_err);
    for(idx_t m=0; m < res->nmodes; ++m) {
      char fname[64];
      snprintf(fname, sizeof(fname), "mode%" SPLATT_PF_IDX ".mat", m + 1);
      idx_t const outrows = (m == stream_mode) ? res->dims[m] : tt->dims[m];
      p_write_unpermuted(fname, res->factors[m], res->dims[m], outrows, rank,
          res->iperm[m]);
    }
    splatt_amd_free_stream_result(res);
  }
  if(tt != NULL) {
    splatt_tt_free(tt);
  }
  if(global_options == NULL) {
    splatt_free_global_opts(gopts);
  }
  if(cpd_options == NULL) {
    splatt_free_cpd_opts(copts);
  }
  return ret;
}
