/*
 * oracle_cpd.c -- TEST INFRASTRUCTURE (the oracle; see oracle.h).
 * CPD-ALS restated on the CPU in FP64, with textbook dense kernels in place
 * of BLAS/LAPACK (syrk, potrf, potrs).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* Threads of the row-independent dense loops (per-row triangular solves and
 * the streaming history product): every row is computed by one thread in
 * the same order, so results do not depend on the count. Reductions over
 * rows (Grams, norms, residuals) stay serial. Default 1. */
static int g_or_threads = 1;

void or_set_threads(int n)
{
  g_or_threads = (n > 1) ? n : 1;
}

int or_get_threads(void)
{
  return g_or_threads;
}

/* src/util.c:15-40: magnitude 3*rand()/RAND_MAX, then sign from a second
 * rand() (even -> negative). */
void or_fill_rand(double * vals, uint64_t n)
{
  for(uint64_t i=0; i < n; ++i) {
    double v = 3.0 * ((double) rand() / (double) RAND_MAX);
    if(rand() % 2 == 0) {
      v *= -1;
    }
    vals[i] = v;
  }
}

/* A^T A of a row-major I x R matrix (src/matrix.c:171-205 computes the
 * upper triangle with syrk); both triangles are filled here. */
void or_mat_aTa(double const * A, uint64_t I, uint64_t R, double * out)
{
  for(uint64_t x=0; x < R * R; ++x) {
    out[x] = 0;
  }
  for(uint64_t i=0; i < I; ++i) {
    double const * row = A + i * R;
    for(uint64_t a=0; a < R; ++a) {
      for(uint64_t b=a; b < R; ++b) {
        out[a * R + b] += row[a] * row[b];
      }
    }
  }
  for(uint64_t a=0; a < R; ++a) {
    for(uint64_t b=0; b < a; ++b) {
      out[a * R + b] = out[b * R + a];
    }
  }
}

/* Cholesky G = L L^T reading the upper triangle of row-major G and writing
 * L into its lower triangle (src/matrix.c:104-124 calls potrf('L') on the
 * column-major view, which is the same factor). Returns 0, or the 1-based
 * column where a non-positive pivot appeared (LAPACK's info). */
int or_cholesky(double * G, uint64_t R)
{
  for(uint64_t j=0; j < R; ++j) {
    double d = G[j * R + j];
    for(uint64_t k=0; k < j; ++k) {
      d -= G[j * R + k] * G[j * R + k];
    }
    if(!(d > 0)) {
      return (int) j + 1;
    }
    d = sqrt(d);
    G[j * R + j] = d;
    for(uint64_t i=j+1; i < R; ++i) {
      double s = G[j * R + i];   /* upper (j,i) == lower (i,j) of G */
      for(uint64_t k=0; k < j; ++k) {
        s -= G[i * R + k] * G[j * R + k];
      }
      G[i * R + j] = s / d;
    }
  }
  return 0;
}

/* potrs, src/matrix.c:127-151: each row b of B becomes G^{-1} b. */
void or_cholesky_solve(double const * L, uint64_t R, double * B,
    uint64_t nrows)
{
  #pragma omp parallel for num_threads(g_or_threads) if(g_or_threads > 1 && nrows >= 256) \
      schedule(static)
  for(uint64_t n=0; n < nrows; ++n) {
    double * b = B + n * R;
    for(uint64_t i=0; i < R; ++i) {
      double s = b[i];
      for(uint64_t k=0; k < i; ++k) {
        s -= L[i * R + k] * b[k];
      }
      b[i] = s / L[i * R + i];
    }
    for(uint64_t i=R; i-- != 0; ) {
      double s = b[i];
      for(uint64_t k=i+1; k < R; ++k) {
        s -= L[k * R + i] * b[k];
      }
      b[i] = s / L[i * R + i];
    }
  }
}

/* src/matrix.c:37-96: lambda_j = ||A(:,j)||_2, columns divided by it. */
void or_mat_normalize(double * A, uint64_t I, uint64_t R, double * lambda)
{
  for(uint64_t r=0; r < R; ++r) {
    lambda[r] = 0;
  }
  for(uint64_t i=0; i < I; ++i) {
    for(uint64_t r=0; r < R; ++r) {
      lambda[r] += A[i * R + r] * A[i * R + r];
    }
  }
  for(uint64_t r=0; r < R; ++r) {
    lambda[r] = sqrt(lambda[r]);
  }
  for(uint64_t i=0; i < I; ++i) {
    for(uint64_t r=0; r < R; ++r) {
      A[i * R + r] /= lambda[r];
    }
  }
}

static int p_cmp_desc(void const * a, void const * b)
{
  double const x = *(double const *) a;
  double const y = *(double const *) b;
  return (x < y) ? 1 : ((x > y) ? -1 : 0);
}

/* src/cpd/constraints/simplex.cc:32-66 (p_project_vector): sort a copy of
 * the row into non-increasing order, accumulate from -1 until an entry no
 * longer exceeds the running mean, theta = sum / count, row = max(row -
 * theta, 0). */
static void p_simplex_row(double * row, double * buf, uint64_t n)
{
  memcpy(buf, row, n * sizeof(double));
  qsort(buf, n, sizeof(double), p_cmp_desc);
  double run = -1.;
  uint64_t piv;
  for(piv=0; piv < n; ++piv) {
    run += buf[piv];
    double const v = buf[piv] - (run / (double) (piv + 1));
    if(v <= 0.) {
      run -= buf[piv];
      break;
    }
  }
  double const theta = run / (double) piv;
  for(uint64_t j=0; j < n; ++j) {
    row[j] -= theta;
    row[j] = (row[j] > 0.) ? row[j] : 0.;
  }
}

/* src/cpd/constraints/maxcolnorm.c:32-61, maxcolnorm-nonneg.c:32-66,
 * nonneg.c: column norms capped at 1 (after clamping negatives for the
 * non-negative variant); plain clamping for nonneg; lasso.c:30-52: entries
 * above thr lose thr, entries below -thr gain thr, the rest become 0;
 * prox 5 projects every row, prox 7 every column, onto the probability
 * simplex. */
void or_prox(int prox, double * P, uint64_t I, uint64_t R, double thr)
{
  if(prox == 7) {
    /* simplex.cc:120-158 (splatt_colsimp_prox): every column on its own */
    double * col = malloc((I ? I : 1) * sizeof(double));
    double * buf = malloc((I ? I : 1) * sizeof(double));
    for(uint64_t j=0; j < R; ++j) {
      for(uint64_t i=0; i < I; ++i) {
        col[i] = P[j + i * R];
      }
      p_simplex_row(col, buf, I);
      for(uint64_t i=0; i < I; ++i) {
        P[j + i * R] = col[i];
      }
    }
    free(buf);
    free(col);
    return;
  }
  if(prox == 5) {
    double * buf = malloc((R ? R : 1) * sizeof(double));
    for(uint64_t i=0; i < I; ++i) {
      p_simplex_row(P + i * R, buf, R);
    }
    free(buf);
    return;
  }
  if(prox == 8) {
    /* nonneg_frob.c:32-55: scaled by thr = rho / (lambda + rho), clamped */
    for(uint64_t x=0; x < I * R; ++x) {
      double const v = P[x] * thr;
      P[x] = (v > 0.) ? v : 0.;
    }
    return;
  }
  if(prox == 9) {
    /* nonneg_lasso.c:25-45: one-sided soft threshold at thr = lambda / rho */
    for(uint64_t x=0; x < I * R; ++x) {
      double const v = P[x];
      P[x] = (v > thr) ? v - thr : 0.;
    }
    return;
  }
  if(prox == 4) {
    for(uint64_t x=0; x < I * R; ++x) {
      double const v = P[x];
      if(v > thr) {
        P[x] = v - thr;
      } else if(v < -thr) {
        P[x] = v + thr;
      } else {
        P[x] = 0.;
      }
    }
    return;
  }
  if(prox == 3) {
    for(uint64_t x=0; x < I * R; ++x) {
      P[x] = (P[x] > 0.) ? P[x] : 0.;
    }
    return;
  }
  if(prox != 1 && prox != 2) {
    return;
  }
  for(uint64_t j=0; j < R; ++j) {
    double norm = 0.;
    for(uint64_t i=0; i < I; ++i) {
      double * p = P + j + i * R;
      if(prox == 2 && *p < 0.) {
        *p = 0.;
      } else {
        norm += (*p) * (*p);
      }
    }
    norm = sqrt(norm);
    if(norm <= 1.) {
      continue;
    }
    for(uint64_t i=0; i < I; ++i) {
      P[j + i * R] /= norm;
    }
  }
}

/* src/cpd/constraints/smooth.c:30-184: every column x of P (length I) is
 * replaced by rho A^{-1} x, A = lam K + rho I with K the pentadiagonal
 * second-difference stencil [1 -4 6 -4 1] whose first and last diagonal
 * entries are 5 (p_form_banded). A is symmetric positive definite; it is
 * factored A = L D L^T (L unit lower with two subdiagonals) and solved per
 * column. */
void or_smooth_prox(double * P, uint64_t I, uint64_t R, double lam, double rho)
{
  if(I == 0) {
    return;
  }
  double * f = malloc(3 * I * sizeof(double));
  double dm2 = 0., dm1 = 0., l1m1 = 0.;
  for(uint64_t i=0; i < I; ++i) {
    double const a = ((i == 0 || i == I - 1) ? 5. * lam : 6. * lam) + rho;
    double const l2 = (i >= 2) ? lam / dm2 : 0.;
    double const l1 = (i >= 1) ? (-4. * lam - l2 * dm2 * l1m1) / dm1 : 0.;
    double const d = a - l2 * l2 * dm2 - l1 * l1 * dm1;
    f[3 * i] = d;
    f[3 * i + 1] = l1;
    f[3 * i + 2] = l2;
    dm2 = dm1;
    dm1 = d;
    l1m1 = l1;
  }
  for(uint64_t c=0; c < R; ++c) {
    double y1 = 0., y2 = 0.;
    for(uint64_t i=0; i < I; ++i) {
      double const y = P[i * R + c] - f[3 * i + 1] * y1 - f[3 * i + 2] * y2;
      P[i * R + c] = y / f[3 * i];
      y2 = y1;
      y1 = y;
    }
    double x1 = 0., x2 = 0.;
    for(uint64_t i=I; i-- > 0;) {
      double const l1n = (i + 1 < I) ? f[3 * (i + 1) + 1] : 0.;
      double const l2n = (i + 2 < I) ? f[3 * (i + 2) + 2] : 0.;
      double const x = P[i * R + c] - l1n * x1 - l2n * x2;
      P[i * R + c] = x * rho;
      x2 = x1;
      x1 = x;
    }
  }
  free(f);
}

/* p_admm_iterate_chunk (src/cpd/admm.c:240-308) on rows of one chunk: L is
 * the Cholesky factor of G + rho I. */
static uint64_t p_admm_iterate(int prox, double const * M, double * P, double * D,
    double const * L, double rho, uint64_t I, uint64_t R, uint64_t max_inner,
    double inner_tol, double lam)
{
  double * aux = malloc((I ? I : 1) * R * sizeof(double));
  double * init = malloc((I ? I : 1) * R * sizeof(double));
  uint64_t it;
  for(it=0; it < max_inner; ++it) {
    memcpy(init, P, I * R * sizeof(double));
    for(uint64_t x=0; x < I * R; ++x) {
      aux[x] = M[x] + rho * (P[x] + D[x]);
    }
    or_cholesky_solve(L, R, aux, I);
    for(uint64_t x=0; x < I * R; ++x) {
      P[x] = aux[x] - D[x];
    }
    if(prox == 6) {
      or_smooth_prox(P, I, R, lam, rho);
    } else {
      or_prox(prox, P, I, R, (prox == 8) ? rho / (lam + rho) : lam / rho);
    }
    double dual_norm = 0.;
    for(uint64_t x=0; x < I * R; ++x) {
      D[x] += P[x] - aux[x];
      dual_norm += D[x] * D[x];
    }
    double p_norm = 0., p_res = 0., d_res = 0.;
    for(uint64_t x=0; x < I * R; ++x) {
      double const pd = P[x] - aux[x];
      double const dd = P[x] - init[x];
      p_norm += P[x] * P[x];
      p_res += pd * pd;
      d_res += dd * dd;
    }
    if(p_res <= inner_tol * p_norm && d_res <= inner_tol * dual_norm) {
      ++it;
      break;
    }
  }
  free(aux);
  free(init);
  return it;
}

/* src/cpd/admm.c:346-354: rho = trace(G)/R added to the diagonal, then the
 * Cholesky factorisation, once per mode. */
static double p_admm_setup(double * G, uint64_t R)
{
  double tr = 0.;
  for(uint64_t i=0; i < R; ++i) {
    tr += G[i * R + i];
  }
  double const rho = tr / (double) R;
  for(uint64_t i=0; i < R; ++i) {
    G[i * R + i] += rho;
  }
  or_cholesky(G, R);
  return rho;
}

uint64_t or_admm(int prox, double const * M, double * P, double * D,
    double * G, uint64_t I, uint64_t R, uint64_t max_inner, double inner_tol,
    double * rho_out, double lam)
{
  double const rho = p_admm_setup(G, R);
  if(rho_out) {
    *rho_out = rho;
  }
  return p_admm_iterate(prox, M, P, D, G, rho, I, R, max_inner, inner_tol, lam);
}

uint64_t or_admm_chunked(int prox, double const * M, double * P, double * D,
    double * G, uint64_t I, uint64_t R, uint64_t chunk, uint64_t max_inner,
    double inner_tol, double lam)
{
  double const rho = p_admm_setup(G, R);
  /* src/cpd/admm.c:355-400: chunks of `chunk` rows, the last one ragged,
   * each iterated to its own convergence; the count is per row */
  uint64_t nchunks = I / chunk;
  if(I % chunk > 0) {
    ++nchunks;
  }
  uint64_t total = 0;
  for(uint64_t c=0; c < nchunks; ++c) {
    uint64_t const start = c * chunk;
    uint64_t const stop = (c == nchunks - 1) ? I : (c + 1) * chunk;
    uint64_t const off = start * R;
    uint64_t const its = p_admm_iterate(prox, M + off, P + off, D + off, G, rho,
        stop - start, R, max_inner, inner_tol, lam);
    total += its * (stop - start);
  }
  return total;
}

int or_cpd_als(or_csf * const * csfs, uint64_t const * mode_csf,
    uint64_t nmodes, uint64_t rank, double ** factors, double * lambda,
    uint64_t max_its, double tol, double * fits, int nthreads)
{
  or_con cons[OR_MAX_NMODES];
  for(uint64_t m=0; m < OR_MAX_NMODES; ++m) {
    cons[m].kind = 0;
    cons[m].prox = 0;
    cons[m].diag_add = 0.;
    cons[m].chunk = 0;
  }
  return or_cpd_als_con(csfs, mode_csf, nmodes, rank, factors, lambda, max_its,
      tol, 50, 1e-2, cons, 1, fits, nthreads);
}

int or_cpd_als_con(or_csf * const * csfs, uint64_t const * mode_csf,
    uint64_t nmodes, uint64_t rank, double ** factors, double * lambda,
    uint64_t max_its, double tol, uint64_t max_inner, double inner_tol,
    or_con const * cons, int unconstrained, double * fits, int nthreads)
{
  uint64_t const R = rank;
  or_csf const * c0 = csfs[0];
  uint64_t maxdim = 0;
  for(uint64_t m=0; m < nmodes; ++m) {
    if(c0->dims[m] > maxdim) maxdim = c0->dims[m];
  }
  double * aTa[OR_MAX_NMODES];
  double * duals[OR_MAX_NMODES];
  for(uint64_t m=0; m < nmodes; ++m) {
    aTa[m] = calloc(R * R, sizeof(double));
    duals[m] = (cons[m].kind == 2) ? calloc(c0->dims[m] * R, sizeof(double)) : NULL;
  }
  double * gram = malloc(R * R * sizeof(double));
  double * mttkrp = malloc(maxdim * R * sizeof(double));
  double * tmp = malloc(R * sizeof(double));

  /* src/cpd/cpd.c:179-204 */
  for(uint64_t m=0; m < nmodes; ++m) {
    or_mat_normalize(factors[m], c0->dims[m], R, lambda);
  }
  for(uint64_t r=0; r < R; ++r) {
    lambda[r] = 1.;
  }
  for(uint64_t m=1; m < nmodes; ++m) {
    or_mat_aTa(factors[m], c0->dims[m], R, aTa[m]);
  }
  double const ttnormsq = or_csf_frobsq(c0);
  double olderr = 1.;
  int its = 0;
  for(uint64_t it=0; it < max_its; ++it) {
    for(uint64_t m=0; m < nmodes; ++m) {
      uint64_t const I = c0->dims[m];
      or_mttkrp_csf(csfs[mode_csf[m]], factors, R, m, mttkrp, nthreads);
      /* src/matrix.c:311-345, the constraint's gram_func, then
       * src/cpd/admm.c:217-235 (closed form) or :346-406 (ADMM) */
      for(uint64_t x=0; x < R * R; ++x) {
        gram[x] = 1.;
      }
      for(uint64_t mm=0; mm < nmodes; ++mm) {
        if(mm == m) continue;
        for(uint64_t x=0; x < R * R; ++x) {
          gram[x] *= aTa[mm][x];
        }
      }
      for(uint64_t i=0; i < R; ++i) {
        gram[i * R + i] += cons[m].diag_add;
      }
      if(cons[m].kind == 2 && cons[m].chunk > 0) {
        or_admm_chunked(cons[m].prox, mttkrp, factors[m], duals[m], gram, I, R,
            cons[m].chunk, max_inner, inner_tol, cons[m].lam);
      } else if(cons[m].kind == 2) {
        or_admm(cons[m].prox, mttkrp, factors[m], duals[m], gram, I, R,
            max_inner, inner_tol, NULL, cons[m].lam);
      } else {
        or_cholesky(gram, R);
        memcpy(factors[m], mttkrp, I * R * sizeof(double));
        or_cholesky_solve(gram, R, factors[m], I);
        if(unconstrained) {
          or_mat_normalize(factors[m], I, R, lambda);
        }
      }
      or_mat_aTa(factors[m], I, R, aTa[m]);
    }
    /* src/cpd/cpd.c:453-487 */
    double norm = 0;
    for(uint64_t i=0; i < R; ++i) {
      for(uint64_t j=i; j < R; ++j) {
        double h = 1.;
        for(uint64_t m=0; m < nmodes; ++m) {
          h *= aTa[m][i * R + j];
        }
        norm += (i == j) ? h * lambda[i] * lambda[i]
                         : 2. * h * lambda[i] * lambda[j];
      }
    }
    norm = fabs(norm);
    /* src/cpd/cpd.c:490-531 with the last mode's MTTKRP */
    uint64_t const last = nmodes - 1;
    double inner = 0;
    for(uint64_t r=0; r < R; ++r) {
      double s = 0;
      for(uint64_t i=0; i < c0->dims[last]; ++i) {
        s += factors[last][i * R + r] * mttkrp[i * R + r];
      }
      inner += s * lambda[r];
    }
    double const err = (ttnormsq + norm - 2 * inner) / ttnormsq;
    fits[it] = 1. - err;
    its = (int) it + 1;
    if(it > 0 && fabs(olderr - err) < tol) {
      break;
    }
    olderr = err;
  }
  /* src/cpd/cpd.c:296-301 and :313-332 */
  if(unconstrained) {
    for(uint64_t m=0; m < nmodes; ++m) {
      or_mat_normalize(factors[m], c0->dims[m], R, tmp);
      for(uint64_t r=0; r < R; ++r) {
        lambda[r] *= tmp[r];
      }
    }
  }
  for(uint64_t m=0; m < nmodes; ++m) {
    free(aTa[m]);
    free(duals[m]);
  }
  free(gram);
  free(mttkrp);
  free(tmp);
  return its;
}
