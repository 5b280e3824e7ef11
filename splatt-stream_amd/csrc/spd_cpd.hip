// spd_cpd.hip -- CPD-ALS sweeps on the GPU (the reference's cpd_iterate,
// src/cpd/cpd.c:161-309, with admm() of src/cpd/admm.c:316-406 per mode).
//
// Everything between the initial upload of the factors and the final
// download stays on the device: MTTKRP (spd_mttkrp), the Hadamard of Grams
// and its Cholesky inverse, the solve, normalisation, Grams (MFMA) and the
// fit terms. The host reads two scalars per iteration (the fit test); the
// ADMM inner loop tests its convergence on the device (ranks up to 64).
#include "spd_internal.hpp"
#include "spd_csf_dev.hpp"
#include "spd_dense.hpp"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <vector>

namespace {

struct DevAllocs {
  std::vector<void *> ptrs;
  ~DevAllocs() { for(void * p : ptrs) spd_free(p); }
  int get(void ** out, size_t bytes) {
    *out = nullptr;
    if(bytes == 0) bytes = 8;
    SPD_TRY(spd_malloc(out, bytes));
    ptrs.push_back(*out);
    return SPD_OK;
  }
};

double wall_s()
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

}  // namespace

extern "C" int spd_cpd_als(spd_csf * const * csfs, int const * mode_csf,
    int const * rows_pos, spd_cpd_cfg const * cfg, double * const * factors,
    double * lambda, double * fits, uint64_t * its_out, double * fit_out)
{
  SPD_TRY(spd_init());
  hipStream_t const st = spd_hs(nullptr);
  int const N = cfg->nmodes;
  uint64_t const R64 = cfg->rank;
  if(N < 2 || N > SPD_MAX_MODES || R64 == 0 || R64 > 4096) {
    spd_set_error(SPD_EINPUT, "cpd: bad shape (nmodes %d, rank %llu)", N,
        (unsigned long long) R64);
    return SPD_EINPUT;
  }
  uint32_t const R = (uint32_t) R64;
  uint32_t const RP = dn_rpad(R);
  int const dt = cfg->dtype;
  size_t const sv = (dt == SPD_F32) ? sizeof(float) : sizeof(double);
  spd_csf const * c0 = csfs[0];
  uint64_t dims[SPD_MAX_MODES];
  uint64_t maxdim = 0;
  for(int m=0; m < N; ++m) {
    dims[m] = c0->dims[m];
    if(dims[m] > maxdim) maxdim = dims[m];
  }
  bool any_admm = false;
  for(int m=0; m < N; ++m) any_admm |= (cfg->con[m].kind == SPD_CON_ADMM);

  DevAllocs mem;
  double * U[SPD_MAX_MODES] = {};
  float * U32[SPD_MAX_MODES] = {};
  double * aTa[SPD_MAX_MODES] = {};
  double * dual[SPD_MAX_MODES] = {};
  void * mats[SPD_MAX_MODES] = {};
  for(int m=0; m < N; ++m) {
    SPD_TRY(mem.get((void **) &U[m], dims[m] * R * sizeof(double)));
    SPD_TRY(spd_h2d(U[m], factors[m], dims[m] * R * sizeof(double)));
    if(dt == SPD_F32) SPD_TRY(mem.get((void **) &U32[m], dims[m] * R * sizeof(float)));
    mats[m] = (dt == SPD_F32) ? (void *) U32[m] : (void *) U[m];
    SPD_TRY(mem.get((void **) &aTa[m], (size_t) R * R * sizeof(double)));
    if(cfg->con[m].kind == SPD_CON_ADMM) {
      SPD_TRY(mem.get((void **) &dual[m], dims[m] * R * sizeof(double)));
      SPD_TRY(spd_zero_async(dual[m], dims[m] * R * sizeof(double), st));
    }
  }
  void * Mbuf = nullptr;
  SPD_TRY(mem.get(&Mbuf, maxdim * R * sv));
  double * aux = nullptr;
  double * init = nullptr;
  if(any_admm) {
    SPD_TRY(mem.get((void **) &aux, maxdim * R * sizeof(double)));
    SPD_TRY(mem.get((void **) &init, maxdim * R * sizeof(double)));
  }
  double * ginv = nullptr;
  SPD_TRY(mem.get((void **) &ginv, (size_t) RP * RP * sizeof(double)));
  double * lam = nullptr;
  double * tmpv = nullptr;
  SPD_TRY(mem.get((void **) &lam, R * sizeof(double)));
  SPD_TRY(mem.get((void **) &tmpv, R * sizeof(double)));
  // small values: [0] rho, [1] cpd norm, [2] inner product, [8 ..] ADMM sums
  double * sc = nullptr;
  SPD_TRY(mem.get((void **) &sc, (8 + 4 + R) * sizeof(double)));
  int * info = nullptr;
  SPD_TRY(mem.get((void **) &info, sizeof(int)));
  SPD_TRY(spd_zero_async(info, sizeof(int), st));
  int * ctl = nullptr;
  SPD_TRY(mem.get((void **) &ctl, 2 * sizeof(int)));
  unsigned long long * its_rows = nullptr;
  SPD_TRY(mem.get((void **) &its_rows, sizeof(unsigned long long)));
  double * const d_rho = sc + 0;
  double * const d_norm = sc + 1;
  double * const d_inner = sc + 2;
  double * const d_sums = sc + 8;

  // initial normalisation of every factor, then lambda = 1
  // (src/cpd/cpd.c:179-196)
  for(int m=0; m < N; ++m) {
    SPD_TRY(dn_colnorm(U[m], dims[m], R, tmpv, st));
    SPD_TRY(dn_div_cols(U[m], dims[m], R, tmpv, U32[m], st));
  }
  {
    std::vector<double> ones(R, 1.);
    SPD_TRY(spd_h2d(lam, ones.data(), R * sizeof(double)));
  }
  double const * aTa_c[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) aTa_c[m] = aTa[m];
  for(int m=1; m < N; ++m) {
    SPD_TRY(dn_gram(U[m], dims[m], R, nullptr, nullptr, aTa[m], st));
  }
  double ttnormsq = 0.;
  SPD_TRY(spd_csf_frobsq(csfs[0], &ttnormsq));

  double olderr = 1.;
  double err = 0.;
  uint64_t it = 0;
  uint64_t nits = 0;
  double inner_its[SPD_MAX_MODES] = {};
  for(it=0; it < cfg->max_its; ++it) {
    double const t0 = wall_s();
    for(int m=0; m < N; ++m) {
      spd_con const & con = cfg->con[m];
      SPD_TRY(spd_mttkrp(csfs[mode_csf[m]], m, R, mats, Mbuf, dt,
          rows_pos[mode_csf[m]], st));
      inner_its[m] = 0.;
      if(con.kind == SPD_CON_ADMM && con.chunk > 0) {
        // row-separable prox: independent chunks of rows (src/cpd/admm.c:355-400)
        SPD_TRY(dn_chol_inv(aTa_c, N, m, R, con.diag_add, 1, ginv, d_rho, info, st));
        SPD_TRY(spd_zero_async(its_rows, sizeof(unsigned long long), st));
        SPD_TRY(dn_admm_chunked(Mbuf, dt, ginv, d_rho, dims[m], R, con.chunk, U[m], dual[m],
            con.prox, con.lam, cfg->max_inner, cfg->inner_tol, its_rows, st));
        if(cfg->verbosity > 1) {
          unsigned long long hr = 0;
          SPD_TRY(spd_d2h(&hr, its_rows, sizeof(hr)));
          inner_its[m] = (double) hr / (double) dims[m];
        }
        SPD_TRY(dn_gram(U[m], dims[m], R, nullptr, U32[m], aTa[m], st));
      } else if(con.kind == SPD_CON_ADMM) {
        SPD_TRY(dn_chol_inv(aTa_c, N, m, R, con.diag_add, 1, ginv, d_rho, info, st));
        SPD_TRY(dn_admm_solve(Mbuf, dt, ginv, d_rho, dims[m], R, U[m], dual[m], aux, init,
            con.prox, con.lam, cfg->max_inner, cfg->inner_tol, d_sums, ctl, st));
        if(cfg->verbosity > 1) {
          int hc[2];
          SPD_TRY(spd_d2h(hc, ctl, sizeof(hc)));
          inner_its[m] = (double) hc[1];
        }
        SPD_TRY(dn_gram(U[m], dims[m], R, nullptr, U32[m], aTa[m], st));
      } else if(con.kind == SPD_CON_CLOSED || !cfg->unconstrained) {
        SPD_TRY(dn_chol_inv(aTa_c, N, m, R, con.diag_add, 0, ginv, d_rho, info, st));
        SPD_TRY(dn_apply(Mbuf, dt, ginv, dims[m], R, U[m], nullptr, st));
        SPD_TRY(dn_gram(U[m], dims[m], R, nullptr, U32[m], aTa[m], st));
      } else {
        // closed form, then absorb the column norms into lambda
        SPD_TRY(dn_chol_inv(aTa_c, N, m, R, 0., 0, ginv, d_rho, info, st));
        SPD_TRY(dn_apply(Mbuf, dt, ginv, dims[m], R, U[m], lam, st));
        SPD_TRY(dn_gram(U[m], dims[m], R, lam, U32[m], aTa[m], st));
      }
    }
    // fit (src/cpd/cpd.c:244-256)
    SPD_TRY(dn_cpd_norm(aTa_c, N, R, lam, d_norm, st));
    SPD_TRY(dn_innerprod(U[N - 1], Mbuf, dt, dims[N - 1], R, lam, d_inner, st));
    double ni[2];
    SPD_TRY(spd_d2h(ni, d_norm, sizeof(ni)));
    double const residual = ttnormsq + ni[0] - 2. * ni[1];
    err = residual / ttnormsq;
    if(fits) fits[it] = 1. - err;
    nits = it + 1;
    if(cfg->verbosity > 0) {
      printf("  its = %4llu (%0.3fs)  rel-errsq = %0.5f  delta = %+0.4e\n",
          (unsigned long long) (it + 1), wall_s() - t0, err, err - olderr);
      if(cfg->verbosity > 1) {
        for(int m=0; m < N; ++m) {
          printf("     mode = %1d", m + 1);
          if(inner_its[m] > 0) printf(" [%4.1f ADMM its per row]", inner_its[m]);
          printf("\n");
        }
      }
      fflush(stdout);
    }
    if(it > 0 && fabs(olderr - err) < cfg->tol) {
      break;
    }
    olderr = err;
  }
  int hinfo = 0;
  SPD_TRY(spd_d2h(&hinfo, info, sizeof(int)));
  if(hinfo != 0) {
    fprintf(stderr, "SPLATT: Cholesky factorisation of the Gram matrix failed "
        "(non-positive pivot at column %d).\n", hinfo);
  }

  // absorb into lambda (src/cpd/cpd.c:313-332)
  std::vector<double> hl(R, 1.);
  SPD_TRY(spd_d2h(hl.data(), lam, R * sizeof(double)));
  if(cfg->unconstrained) {
    std::vector<double> tmp(R);
    for(int m=0; m < N; ++m) {
      SPD_TRY(dn_colnorm(U[m], dims[m], R, tmpv, st));
      SPD_TRY(dn_div_cols(U[m], dims[m], R, tmpv, nullptr, st));
      SPD_TRY(spd_d2h(tmp.data(), tmpv, R * sizeof(double)));
      for(uint32_t r=0; r < R; ++r) hl[r] *= tmp[r];
    }
  }
  for(int m=0; m < N; ++m) {
    SPD_TRY(spd_d2h(factors[m], U[m], dims[m] * R * sizeof(double)));
  }
  memcpy(lambda, hl.data(), R * sizeof(double));
  *its_out = nits;
  *fit_out = 1. - err;
  return SPD_OK;
}
