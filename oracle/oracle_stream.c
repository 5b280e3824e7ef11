/*
 * oracle_stream.c -- TEST INFRASTRUCTURE (the oracle; see oracle.h).
 * Streaming CPD restated on the CPU in FP64: the parser's ordering and
 * relabelling (src/cpd/stream/ParserSimple.cc:13-80, :89-149), the growth of
 * the factor matrices (StreamMatrix.cc:27-83, StreamCPD.cc:192-220), the
 * per-slice ALS with ADMM (StreamCPD.cc:311-571, add_historical :223-294),
 * the COO MTTKRP (src/mttkrp.c:1699-1769), the error measures
 * (src/cpd/cpd.c:533-656, StreamCPD.cc:77-89) and the checkpoints
 * (StreamCPD.cc:89-106, :517-528).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/mttkrp.c:1699-1769 in FP64: out (dims[mode] x R) zeroed, then per
 * nonzero acc = v, acc *= U_m[ind] for m != mode in ascending order, added
 * to the output row. */
void or_mttkrp_stream(or_coo const * tt, double * const * mats, uint64_t rank,
    uint64_t mode, double * out)
{
  uint64_t const I = tt->dims[mode];
  for(uint64_t x=0; x < I * rank; ++x) {
    out[x] = 0.;
  }
  double * acc = malloc(rank * sizeof(double));
  for(uint64_t n=0; n < tt->nnz; ++n) {
    for(uint64_t f=0; f < rank; ++f) {
      acc[f] = tt->vals[n];
    }
    for(uint64_t m=0; m < tt->nmodes; ++m) {
      if(m == mode) continue;
      double const * row = mats[m] + tt->ind[m][n] * rank;
      for(uint64_t f=0; f < rank; ++f) {
        acc[f] *= row[f];
      }
    }
    double * orow = out + tt->ind[mode][n] * rank;
    for(uint64_t f=0; f < rank; ++f) {
      orow[f] += acc[f];
    }
  }
  free(acc);
}

/* src/cpd/cpd.c:533-577: |lambda^T (Hadamard of U_m^T U_m) lambda| over the
 * upper triangle. */
double or_kruskal_norm(double * const * factors, uint64_t const * dims,
    uint64_t nmodes, uint64_t rank, double const * lambda)
{
  double * scratch = malloc(rank * rank * sizeof(double));
  double * ata = malloc(rank * rank * sizeof(double));
  for(uint64_t x=0; x < rank * rank; ++x) {
    scratch[x] = 1.;
  }
  for(uint64_t m=0; m < nmodes; ++m) {
    or_mat_aTa(factors[m], dims[m], rank, ata);
    for(uint64_t x=0; x < rank * rank; ++x) {
      scratch[x] *= ata[x];
    }
  }
  double norm = 0.;
  for(uint64_t i=0; i < rank; ++i) {
    norm += scratch[i * rank + i] * lambda[i] * lambda[i];
    for(uint64_t j=i+1; j < rank; ++j) {
      norm += scratch[i * rank + j] * lambda[i] * lambda[j] * 2;
    }
  }
  free(scratch);
  free(ata);
  return fabs(norm);
}

/* src/cpd/cpd.c:580-656: relative error (not squared) of the model against
 * the tensor, via an MTTKRP of the smallest mode. factor dims are dims[]. */
double or_cpd_error(or_coo const * tt, double * const * factors,
    uint64_t const * fdims, uint64_t rank, double const * lambda)
{
  uint64_t small = 0;
  for(uint64_t m=1; m < tt->nmodes; ++m) {
    if(tt->dims[m] < tt->dims[small]) small = m;
  }
  uint64_t const nrows = tt->dims[small];
  double * M = malloc((nrows ? nrows : 1) * rank * sizeof(double));
  or_mttkrp_stream(tt, factors, rank, small, M);
  double inner = 0.;
  /* rows beyond the factor's own rows cannot carry nonzeros */
  uint64_t const frows = (fdims[small] < nrows) ? fdims[small] : nrows;
  for(uint64_t r=0; r < rank; ++r) {
    double acc = 0.;
    for(uint64_t i=0; i < frows; ++i) {
      acc += factors[small][i * rank + r] * M[i * rank + r];
    }
    inner += acc * lambda[r];
  }
  double xn = 0.;
  for(uint64_t n=0; n < tt->nnz; ++n) {
    xn += tt->vals[n] * tt->vals[n];
  }
  double const zn = or_kruskal_norm(factors, fdims, tt->nmodes, rank, lambda);
  double const residual = sqrt(xn + zn - 2 * inner);
  free(M);
  return residual / sqrt(xn);
}

/* ParserSimple's constructor (ParserSimple.cc:13-80): sort by the stream
 * mode, then (mode+1, mode+2, ...) cyclically (src/sort.c:921-935), and
 * relabel every other mode by first appearance in that order. iperm[m]
 * (host, dims[m] entries, may be NULL) receives new -> original labels. */
void or_stream_prepare(or_coo * tt, uint64_t stream_mode, uint64_t ** iperm)
{
  uint64_t const N = tt->nmodes;
  uint64_t perm[OR_MAX_NMODES] = {0};
  for(uint64_t m=0; m < N; ++m) {
    perm[m] = (stream_mode + m) % N;
  }
  or_coo_sort(tt, perm);
  for(uint64_t m=0; m < N; ++m) {
    uint64_t const D = tt->dims[m];
    if(m == stream_mode) {
      if(iperm && iperm[m]) {
        for(uint64_t i=0; i < D; ++i) iperm[m][i] = i;
      }
      continue;
    }
    uint64_t * p = malloc((D ? D : 1) * sizeof(uint64_t));
    for(uint64_t i=0; i < D; ++i) {
      p[i] = UINT64_MAX;
      if(iperm && iperm[m]) iperm[m][i] = UINT64_MAX;
    }
    uint64_t seen = 0;
    for(uint64_t x=0; x < tt->nnz; ++x) {
      uint64_t const ind = tt->ind[m][x];
      if(p[ind] == UINT64_MAX) {
        p[ind] = seen;
        if(iperm && iperm[m]) iperm[m][seen] = ind;
        ++seen;
      }
      tt->ind[m][x] = p[ind];
    }
    free(p);
  }
}

/* ---- the streaming factorisation ------------------------------------------ */

typedef struct {
  uint64_t rows;
  uint64_t cap;
  double * v;
} p_smat;

static void p_smat_grow(p_smat * s, uint64_t rows, uint64_t R, int rnd)
{
  if(rows < s->rows) return;
  if(rows > s->cap) {
    uint64_t cap = s->cap ? s->cap : 128;
    while(rows > cap) cap *= 2;
    double * nv = calloc(cap * R, sizeof(double));
    if(s->v) memcpy(nv, s->v, s->rows * R * sizeof(double));
    free(s->v);
    s->v = nv;
    s->cap = cap;
  }
  if(rows > s->rows) {
    if(rnd) {
      or_fill_rand(s->v + s->rows * R, (rows - s->rows) * R);
    } else {
      memset(s->v + s->rows * R, 0, (rows - s->rows) * R * sizeof(double));
    }
  }
  s->rows = rows;
}

/* symmetric copy of the upper triangle into the lower (StreamCPD.cc:33-48) */
static void p_copy_upper(double * A, uint64_t R)
{
  for(uint64_t i=1; i < R; ++i) {
    for(uint64_t j=0; j < i; ++j) {
      A[i * R + j] = A[j * R + i];
    }
  }
}

/* gram = Hadamard over m != mode of aTa[m] (upper triangle as the
 * reference uses it), plus the constraint's diagonal term */
static void p_form_gram(double * const * aTa, uint64_t N, uint64_t mode,
    uint64_t R, double diag, double * gram)
{
  for(uint64_t x=0; x < R * R; ++x) gram[x] = 1.;
  for(uint64_t m=0; m < N; ++m) {
    if(m == mode) continue;
    for(uint64_t i=0; i < R; ++i) {
      for(uint64_t j=i; j < R; ++j) {
        gram[i * R + j] *= aTa[m][i * R + j];
      }
    }
  }
  p_copy_upper(gram, R);
  for(uint64_t i=0; i < R; ++i) gram[i * R + i] += diag;
}

/* stream_prev(nprev) (ParserSimple.cc:185-226) after `done` slices: the
 * nonzeros of slices done-nprev .. done-1 (or from 0), time labels shifted
 * to start at 0, the other dims at their running maxima. The caller frees
 * ind[t]. */
static or_coo p_window(or_coo const * tt, uint64_t t, uint64_t const * bstart,
    uint64_t done, uint64_t nprev, uint64_t const * prev_dim)
{
  uint64_t const b0 = (nprev < done) ? done - nprev : 0;
  uint64_t const x0 = bstart[b0];
  uint64_t const x1 = bstart[done];
  or_coo w;
  memset(&w, 0, sizeof(w));
  w.nmodes = tt->nmodes;
  w.nnz = x1 - x0;
  w.vals = tt->vals + x0;
  for(uint64_t m=0; m < tt->nmodes; ++m) {
    if(m == t) {
      w.ind[m] = malloc((w.nnz ? w.nnz : 1) * sizeof(uint64_t));
      for(uint64_t x=0; x < w.nnz; ++x) w.ind[m][x] = tt->ind[m][x0 + x] - b0;
      w.dims[m] = done - b0;
    } else {
      w.ind[m] = tt->ind[m] + x0;
      w.dims[m] = prev_dim[m];
    }
  }
  return w;
}

/* compute_cpd_errorsq (StreamCPD.cc:89-106): factors drawn from rand() for
 * every mode in order (splatt_alloc_cpd, src/cpd/cpd.c:112-135), then
 * CPD-ALS with the default options (tolerance 1e-5, 200 iterations,
 * src/cpd/cpd.c:82-99); returns 1 - fit. One CSF serves every mode. */
static double p_refit(or_coo const * w, uint64_t R)
{
  uint64_t const N = w->nmodes;
  or_coo c = *w;
  for(uint64_t m=0; m < N; ++m) {
    c.ind[m] = malloc((w->nnz ? w->nnz : 1) * sizeof(uint64_t));
    memcpy(c.ind[m], w->ind[m], w->nnz * sizeof(uint64_t));
  }
  c.vals = malloc((w->nnz ? w->nnz : 1) * sizeof(double));
  memcpy(c.vals, w->vals, w->nnz * sizeof(double));
  uint64_t perm[OR_MAX_NMODES];
  or_mode_order(c.dims, N, 0, 0, perm);
  or_csf * csf = or_csf_build(&c, perm, 0, 0, 1, 0);
  double * f[OR_MAX_NMODES];
  uint64_t mode_csf[OR_MAX_NMODES];
  for(uint64_t m=0; m < N; ++m) {
    f[m] = malloc(c.dims[m] * R * sizeof(double));
    or_fill_rand(f[m], c.dims[m] * R);
    mode_csf[m] = 0;
  }
  double * lam = malloc(R * sizeof(double));
  double * fits = malloc(200 * sizeof(double));
  int const its = or_cpd_als(&csf, mode_csf, N, R, f, lam, 200, 1e-5, fits, 1);
  double const err = (its > 0) ? 1. - fits[its - 1] : NAN;
  for(uint64_t m=0; m < N; ++m) {
    free(f[m]);
    free(c.ind[m]);
  }
  free(c.vals);
  free(lam);
  free(fits);
  or_csf_free(csf);
  return err;
}

int or_stream_cpd(or_coo const * tt, or_stream_cfg const * cfg,
    or_stream_out * out)
{
  uint64_t const N = tt->nmodes;
  uint64_t const R = cfg->rank;
  uint64_t const t = cfg->stream_mode;
  double const tol = 5e-2;          /* StreamCPD.cc:366-368 */
  uint64_t const max_inner = 25;
  double const inner_tol = 7e-2;

  p_smat snew[OR_MAX_NMODES], sold[OR_MAX_NMODES], sdual[OR_MAX_NMODES];
  memset(snew, 0, sizeof(snew));
  memset(sold, 0, sizeof(sold));
  memset(sdual, 0, sizeof(sdual));
  double * aTa[OR_MAX_NMODES];
  for(uint64_t m=0; m < N; ++m) aTa[m] = calloc(R * R, sizeof(double));
  double * old_gram = calloc(R * R, sizeof(double));
  double * gram = malloc(R * R * sizeof(double));
  double * hist = malloc(R * R * sizeof(double));
  double * cross = malloc(R * R * sizeof(double));
  double * time_row = calloc(R, sizeof(double));
  double * tmp = malloc(R * sizeof(double));
  uint64_t maxdim = 1;
  for(uint64_t m=0; m < N; ++m) if(tt->dims[m] > maxdim) maxdim = tt->dims[m];
  double * mttkrp = malloc(maxdim * R * sizeof(double));

  uint64_t prev_dim[OR_MAX_NMODES] = {0};
  uint64_t nnz_ptr = 0;
  uint64_t it = 0;
  /* bstart[b]: first nonzero of slice b; bstart[b + 1] its end */
  uint64_t * bstart = calloc(tt->dims[t] + 2, sizeof(uint64_t));
  while(nnz_ptr < tt->nnz) {
    /* ParserSimple::next_batch (ParserSimple.cc:89-149) */
    uint64_t start = nnz_ptr;
    while(start < tt->nnz && tt->ind[t][start] < it) ++start;
    uint64_t end = start;
    while(end < tt->nnz && tt->ind[t][end] < it + 1) ++end;
    uint64_t const bnnz = end - start;
    if(bnnz == 0) {
      break;   /* an empty time slice ends the stream (the reference asserts) */
    }
    bstart[it] = start;
    bstart[it + 1] = end;
    or_coo batch;
    memset(&batch, 0, sizeof(batch));
    batch.nmodes = N;
    batch.nnz = bnnz;
    batch.vals = tt->vals + start;
    uint64_t * zeros = calloc(bnnz, sizeof(uint64_t));
    for(uint64_t m=0; m < N; ++m) {
      if(m == t) {
        batch.ind[m] = zeros;
        batch.dims[m] = 1;
        continue;
      }
      batch.ind[m] = tt->ind[m] + start;
      uint64_t dim = 0;
      for(uint64_t x=0; x < bnnz; ++x) {
        if(batch.ind[m][x] > dim) dim = batch.ind[m][x];
      }
      batch.dims[m] = (dim + 1 > prev_dim[m]) ? dim + 1 : prev_dim[m];
      prev_dim[m] = batch.dims[m];
    }

    /* grow_mats (StreamCPD.cc:192-220) */
    for(uint64_t m=0; m < N; ++m) {
      if(m == t) continue;
      p_smat_grow(&snew[m], batch.dims[m], R, 1);
      or_mat_aTa(snew[m].v, snew[m].rows, R, aTa[m]);
      p_smat_grow(&sold[m], batch.dims[m], R, 0);
      p_smat_grow(&sdual[m], batch.dims[m], R, 0);
    }
    if(it == 0) {
      for(uint64_t m=0; m < N; ++m) {
        if(m == t) continue;
        or_mat_normalize(snew[m].v, snew[m].rows, R, tmp);
      }
    }
    double * mats[OR_MAX_NMODES];
    for(uint64_t m=0; m < N; ++m) mats[m] = (m == t) ? time_row : snew[m].v;

    double prev_delta = 0.;
    uint64_t outer;
    for(outer=0; outer < cfg->max_its; ++outer) {
      double delta = 0.;
      /* time row: MTTKRP and the closed-form solve */
      or_mttkrp_stream(&batch, mats, R, t, mttkrp);
      p_form_gram(aTa, N, t, R, cfg->cons[t].diag_add, gram);
      or_cholesky(gram, R);
      memcpy(time_row, mttkrp, R * sizeof(double));
      or_cholesky_solve(gram, R, time_row, 1);
      /* time Gram, updated every outer iteration (StreamCPD.cc:420-435) */
      p_copy_upper(aTa[t], R);
      memcpy(old_gram, aTa[t], R * R * sizeof(double));
      for(uint64_t i=0; i < R; ++i) {
        for(uint64_t j=0; j < R; ++j) {
          aTa[t][i * R + j] += time_row[i] * time_row[j];
        }
      }
      for(uint64_t m=0; m < N; ++m) {
        if(m == t) continue;
        uint64_t const I = batch.dims[m];
        or_mttkrp_stream(&batch, mats, R, m, mttkrp);
        /* add_historical (StreamCPD.cc:223-294) */
        memcpy(hist, old_gram, R * R * sizeof(double));
        p_copy_upper(hist, R);
        for(uint64_t mm=0; mm < N; ++mm) {
          if(mm == m || mm == t) continue;
          for(uint64_t x=0; x < R * R; ++x) cross[x] = 0.;
          for(uint64_t i=0; i < snew[mm].rows; ++i) {
            double const * a = sold[mm].v + i * R;
            double const * b = snew[mm].v + i * R;
            for(uint64_t r=0; r < R; ++r) {
              for(uint64_t s=0; s < R; ++s) {
                cross[r * R + s] += a[r] * b[s];
              }
            }
          }
          for(uint64_t x=0; x < R * R; ++x) hist[x] *= cross[x];
        }
        int const nthr = or_get_threads();
        #pragma omp parallel for num_threads(nthr) if(nthr > 1 && I >= 256) schedule(static)
        for(uint64_t i=0; i < I; ++i) {
          double const * orow = sold[m].v + i * R;
          for(uint64_t s=0; s < R; ++s) {
            double acc = 0.;
            for(uint64_t r=0; r < R; ++r) acc += orow[r] * hist[r * R + s];
            mttkrp[i * R + s] += acc;
          }
        }
        memset(sdual[m].v, 0, sdual[m].rows * R * sizeof(double));
        p_form_gram(aTa, N, m, R, cfg->cons[m].diag_add, gram);
        if(cfg->cons[m].kind == 2) {
          or_admm(cfg->cons[m].prox, mttkrp, snew[m].v, sdual[m].v, gram, I, R,
              max_inner, inner_tol, NULL, cfg->cons[m].lam);
        } else {
          or_cholesky(gram, R);
          memcpy(snew[m].v, mttkrp, I * R * sizeof(double));
          or_cholesky_solve(gram, R, snew[m].v, I);
        }
        or_mat_aTa(snew[m].v, I, R, aTa[m]);
        double dn = 0., nn = 0.;
        for(uint64_t x=0; x < I * R; ++x) {
          double const d = sold[m].v[x] - snew[m].v[x];
          dn += d * d;
          nn += snew[m].v[x] * snew[m].v[x];
        }
        delta += sqrt(dn) / sqrt(nn);
      }
      if(outer > 0 && fabs(delta - prev_delta) < tol) {
        ++outer;
        break;
      }
      prev_delta = delta;
    }
    if(out->outer_its && it < out->max_batches) out->outer_its[it] = outer;

    for(uint64_t x=0; x < R * R; ++x) aTa[t][x] *= cfg->forget;
    for(uint64_t m=0; m < N; ++m) {
      if(m == t) continue;
      memcpy(sold[m].v, snew[m].v, snew[m].rows * R * sizeof(double));
    }
    if(it < out->max_batches) memcpy(out->time + it * R, time_row, R * sizeof(double));
    ++it;

    /* local error of this slice, compute_errorsq(1) (StreamCPD.cc:77-89) */
    if(out->local_err && it - 1 < out->max_batches) {
      double * f[OR_MAX_NMODES];
      uint64_t fd[OR_MAX_NMODES];
      for(uint64_t m=0; m < N; ++m) {
        f[m] = (m == t) ? time_row : snew[m].v;
        fd[m] = (m == t) ? 1 : snew[m].rows;
      }
      double * ones = malloc(R * sizeof(double));
      for(uint64_t r=0; r < R; ++r) ones[r] = 1.;
      /* stream_prev(1): this slice with its time index shifted to 0 and the
       * other dims at their running maxima */
      double const e = or_cpd_error(&batch, f, fd, R, ones);
      out->local_err[it - 1] = e * e;
      free(ones);
    }
    /* checkpoints (StreamCPD.cc:517-528): every check_interval slices and
     * after the slice that consumes the last nonzero (last_batch,
     * ParserSimple.cc:229-232) */
    double ge = -1., l10 = -1., ce = -1.;
    if(cfg->check_interval > 0 && it <= out->max_batches &&
        (it % cfg->check_interval == 0 || end == tt->nnz)) {
      double * ones = malloc(R * sizeof(double));
      for(uint64_t r=0; r < R; ++r) ones[r] = 1.;
      uint64_t const nps[2] = {it, 10};
      double errs[2];
      for(int k=0; k < 2; ++k) {
        or_coo w = p_window(tt, t, bstart, it, nps[k], prev_dim);
        uint64_t const b0 = (nps[k] < it) ? it - nps[k] : 0;
        double * f[OR_MAX_NMODES];
        uint64_t fd[OR_MAX_NMODES];
        for(uint64_t m=0; m < N; ++m) {
          f[m] = (m == t) ? out->time + b0 * R : snew[m].v;
          fd[m] = (m == t) ? it - b0 : snew[m].rows;
        }
        double const e = or_cpd_error(&w, f, fd, R, ones);
        errs[k] = e * e;
        if(k == 0 && cfg->refit) {
          ce = p_refit(&w, R);
          if(isnan(ce)) ce = -1.;
        }
        free(w.ind[t]);
      }
      ge = errs[0];
      l10 = errs[1];
      free(ones);
    }
    if(it - 1 < out->max_batches) {
      if(out->global_err) out->global_err[it - 1] = ge;
      if(out->local10_err) out->local10_err[it - 1] = l10;
      if(out->cpd_err) out->cpd_err[it - 1] = ce;
    }
    free(zeros);
    nnz_ptr = end;
  }
  free(bstart);
  out->nbatches = it;
  for(uint64_t m=0; m < N; ++m) {
    if(m == t) {
      out->dims[m] = it;
      continue;
    }
    out->dims[m] = snew[m].rows;
    if(out->factors[m]) memcpy(out->factors[m], snew[m].v, snew[m].rows * R * sizeof(double));
  }
  /* final-err: cpd_error of the whole stream (StreamCPD.cc:549-556) */
  {
    double * f[OR_MAX_NMODES];
    uint64_t fd[OR_MAX_NMODES];
    for(uint64_t m=0; m < N; ++m) {
      f[m] = (m == t) ? out->time : snew[m].v;
      fd[m] = (m == t) ? it : snew[m].rows;
    }
    double * ones = malloc(R * sizeof(double));
    for(uint64_t r=0; r < R; ++r) ones[r] = 1.;
    double const e = or_cpd_error(tt, f, fd, R, ones);
    out->final_err = e * e;
    free(ones);
  }

  for(uint64_t m=0; m < N; ++m) {
    free(aTa[m]);
    free(snew[m].v);
    free(sold[m].v);
    free(sdual[m].v);
  }
  free(old_gram);
  free(gram);
  free(hist);
  free(cross);
  free(time_row);
  free(tmp);
  free(mttkrp);
  return 0;
}
