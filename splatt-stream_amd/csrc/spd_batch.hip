// spd_batch.hip -- the batched streaming-CPD step: one time slice of each of
// up to BT_MAXK streams, run in lockstep by kernels that each cover the
// whole batch (spd_batch.hpp). The per-slice algorithm is the one of
// StreamState::slice (the reference's StreamCPD::compute,
// src/cpd/stream/StreamCPD.cc:406-506): per outer iteration the time row
// (COO MTTKRP to one row, closed-form solve, time Gram update), then per
// non-streamed mode the COO MTTKRP, the history term, the dual reset, ADMM
// (or the closed form), the Gram and the change measure; the outer loop and
// ADMM test convergence on the device, so the host queues the whole step
// and synchronises once per batch.
#include "spd_internal.hpp"
#include "spd_dense.hpp"
#include "spd_batch.hpp"
#include "spd_batch_host.hpp"
#include <string.h>
#include <vector>

// batched launchers (spd_dense.hip, spd_stream.hip)
int dnb_gram(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, int what, int k, int gate, hipStream_t st);
int dnb_colnorm_div(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st);
int dnb_innerprod(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st);
int dnb_gj_inv(uint32_t K, BSlot const * dS, BHead const * dH, uint32_t TT, int mode, int gate,
    hipStream_t st);
int dnb_apply(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, int what, int gate, hipStream_t st);
int dnb_admm(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, uint32_t max_inner, hipStream_t st);
int dnb_cpd_norm(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st);
int sstb_coo(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, int copy, int gate, int t, hipStream_t st);
int sstb_zero_dual(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st);
int sstb_time_gram(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st);
int sstb_hist(uint32_t K, BSlot const * dS, BHead const * dH, int ncross, hipStream_t st);
int sstb_delta(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st);
int sstb_ctl_reset(BSlot const * dS, BHead const * dH, hipStream_t st);
int sstb_begin(BSlot const * dS, BHead const * dH, hipStream_t st);
int sstb_outer_check(BSlot const * dS, BHead const * dH, int o, hipStream_t st);
int sstb_finish(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st);
int sstb_grow(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st);
int sstb_collect(uint32_t K, BSlot const * dS, BHead const * dH, double * out, uint32_t ld,
    hipStream_t st);

#define BG_ALL 0
#define BG_LIVE 1
#define BG_CLOSED 4
#define BG_SMALL 5
#define SB_ALL 0
#define SB_LIVE 1
#define SB_SMALL 2

size_t bt_part_doubles(uint32_t R, uint64_t cap)
{
  uint64_t const RP = dn_rpad(R);
  uint64_t g = ((cap + 3) / 4 + 63) / 64;
  if(g < 1) g = 1;
  if(g > 1024) g = 1024;
  uint64_t a = (cap + 63) / 64;
  if(a < 1) a = 1;
  if(a > 1536) a = 1536;
  size_t n = g * RP * RP;
  if(a * (RP + 4) > n) n = a * (RP + 4);
  if(2048ull * R > n) n = 2048ull * R;
  if(1024ull * (R > 2 ? R : 2) > n) n = 1024ull * (R > 2 ? R : 2);
  return n;
}

struct BtEngine {
  hipStream_t st = nullptr;
  BSlot * dS = nullptr;           // device descriptors
  BHead * dH = nullptr;
  BCtl * dC = nullptr;            // per batch position
  double * dRes = nullptr;        // per batch position results
  uint32_t * dZero = nullptr;     // zero labels of the streamed mode
  uint64_t zero_cap = 0;
  BSlot * hS = nullptr;           // pinned host copies of the descriptors
  BHead * hH = nullptr;
  double * hRes = nullptr;
};

static uint32_t const RES_LD = 64 + 4;

int bt_engine_create(BtEngine ** out)
{
  *out = nullptr;
  BtEngine * e = new BtEngine();
  int rc = SPD_OK;
  if(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) rc = SPD_EHIP;
  if(rc == SPD_OK) rc = spd_malloc((void **) &e->dS, BT_MAXK * sizeof(BSlot));
  if(rc == SPD_OK) rc = spd_malloc((void **) &e->dH, sizeof(BHead));
  if(rc == SPD_OK) rc = spd_malloc((void **) &e->dC, BT_MAXK * sizeof(BCtl));
  if(rc == SPD_OK) rc = spd_malloc((void **) &e->dRes, BT_MAXK * RES_LD * sizeof(double));
  if(rc == SPD_OK && hipHostMalloc((void **) &e->hS, BT_MAXK * sizeof(BSlot),
      hipHostMallocDefault) != hipSuccess) rc = SPD_ENOMEM;
  if(rc == SPD_OK && hipHostMalloc((void **) &e->hH, sizeof(BHead), hipHostMallocDefault) !=
      hipSuccess) rc = SPD_ENOMEM;
  if(rc == SPD_OK && hipHostMalloc((void **) &e->hRes, BT_MAXK * RES_LD * sizeof(double),
      hipHostMallocDefault) != hipSuccess) rc = SPD_ENOMEM;
  if(rc != SPD_OK) {
    bt_engine_free(e);
    spd_set_error(rc, "batch engine: device or pinned allocation failed");
    return rc;
  }
  *out = e;
  return SPD_OK;
}

void bt_engine_free(BtEngine * e)
{
  if(e == nullptr) return;
  if(e->st) (void) hipStreamSynchronize(e->st);
  spd_free(e->dS);
  spd_free(e->dH);
  spd_free(e->dC);
  spd_free(e->dRes);
  spd_free(e->dZero);
  if(e->hS) (void) hipHostFree(e->hS);
  if(e->hH) (void) hipHostFree(e->hH);
  if(e->hRes) (void) hipHostFree(e->hRes);
  if(e->st) (void) hipStreamDestroy(e->st);
  delete e;
}

hipStream_t bt_engine_stream(BtEngine * e)
{
  return e->st;
}

// One step of every slice in[0..K): the slices' device arrays are already
// on the device (uploaded on the engine stream by the caller).
int bt_engine_run(BtEngine * e, BtSlice * const * in, uint32_t K)
{
  if(K == 0) return SPD_OK;
  if(K > BT_MAXK) {
    spd_set_error(SPD_EINPUT, "batch of %u slices (at most %d)", K, BT_MAXK);
    return SPD_EINPUT;
  }
  hipStream_t const st = e->st;
  StreamState const & s0 = *in[0]->S;
  int const N = s0.N;
  int const t = s0.t;
  int const lead = (t + 1) % N;
  uint32_t const TT = s0.RP / 16;
  uint32_t max_its = 0, max_inner = 0;
  uint64_t maxnnz = 1;
  for(uint32_t q=0; q < K; ++q) {
    StreamState const & S = *in[q]->S;
    if(S.N != N || S.t != t || S.RP != s0.RP || S.R > 64) {
      spd_set_error(SPD_EINPUT, "batch mixes stream shapes");
      return SPD_EINPUT;
    }
    if(S.cfg.max_its > max_its) max_its = (uint32_t) S.cfg.max_its;
    if(S.cfg.max_inner > max_inner) max_inner = (uint32_t) S.cfg.max_inner;
    if(in[q]->nnz > maxnnz) maxnnz = in[q]->nnz;
  }
  if(maxnnz > e->zero_cap) {
    uint64_t nc = e->zero_cap ? e->zero_cap : 4096;
    while(nc < maxnnz) nc *= 2;
    SPD_HIP(hipStreamSynchronize(st));
    spd_free(e->dZero);
    e->dZero = nullptr;
    e->zero_cap = 0;
    SPD_TRY(spd_malloc((void **) &e->dZero, nc * sizeof(uint32_t)));
    SPD_TRY(spd_zero_async(e->dZero, nc * sizeof(uint32_t), st));
    e->zero_cap = nc;
  }
  // descriptors
  BSlot * const hS = e->hS;
  for(uint32_t q=0; q < K; ++q) {
    BtSlice const & b = *in[q];
    StreamState const & S = *b.S;
    BSlot & d = hS[q];
    memset(&d, 0, sizeof(d));
    for(int m=0; m < N; ++m) {
      d.Unew[m] = S.Unew[m];
      d.Uold[m] = S.Uold[m];
      d.Dual[m] = S.Dual[m];
      d.aTa[m] = S.aTa[m];
      d.rows[m] = (m == t) ? 1 : b.rows[m];
      d.grow0[m] = (m == t) ? 0 : b.grow0[m];
      d.grow[m] = (m == t || b.rows[m] == b.grow0[m]) ? nullptr : b.d_grow[m];
      d.kind[m] = S.cfg.con[m].kind;
      d.prox[m] = S.cfg.con[m].prox;
      d.diag_add[m] = S.cfg.con[m].diag_add;
      d.lam[m] = S.cfg.con[m].lam;
      for(int c=0; c < N; ++c) {
        // copy c: the slice's own order for the lead and streamed modes
        int const src = (c == t || c == lead) ? lead : c;
        d.ind[c][m] = (m == t) ? e->dZero : b.d_ind[src][m];
      }
    }
    for(int c=0; c < N; ++c) d.vals[c] = b.d_vals[(c == t || c == lead) ? lead : c];
    d.time_row = S.time_row;
    d.Mbuf = S.Mbuf;
    d.aux = S.aux;
    d.init = S.init_buf;
    d.old_gram = S.old_gram;
    d.ginv = S.ginv;
    d.H = S.H;
    d.cross = S.cross;
    d.tgram = S.tgram;
    d.lam1 = S.lam1;
    d.tmpv = S.tmpv;
    d.part = b.d_part;
    d.c = e->dC + q;
    d.nnz = b.nnz;
    d.first = b.first;
    d.small = b.small;
    d.forget = S.cfg.forget;
    d.tol = S.cfg.tol;
    d.inner_tol = S.cfg.inner_tol;
    d.max_inner = (uint32_t) S.cfg.max_inner;
    d.max_its = (uint32_t) S.cfg.max_its;
    d.R = S.R;
  }
  e->hH->K = K;
  e->hH->N = N;
  e->hH->t = t;
  e->hH->lead = lead;
  SPD_HIP(hipMemcpyAsync(e->dS, hS, K * sizeof(BSlot), hipMemcpyHostToDevice, st));
  SPD_HIP(hipMemcpyAsync(e->dH, e->hH, sizeof(BHead), hipMemcpyHostToDevice, st));
  BSlot const * const dS = e->dS;
  BHead const * const dH = e->dH;

  // grow_mats (StreamCPD.cc:192-220): new rows, Grams; normalise the first
  // slice's factors after their Grams
  SPD_TRY(sstb_begin(dS, dH, st));
  SPD_TRY(sstb_grow(K, dS, dH, st));
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    SPD_TRY(dnb_gram(hS, dS, dH, K, TT, m, 0, 0, BG_ALL, st));
  }
  bool any_first = false;
  for(uint32_t q=0; q < K; ++q) any_first = any_first || hS[q].first;
  if(any_first) {
    for(int m=0; m < N; ++m) {
      if(m != t) SPD_TRY(dnb_colnorm_div(hS, dS, dH, K, m, st));
    }
  }
  for(uint32_t o=0; o < max_its; ++o) {
    // time row: MTTKRP to one row, closed-form solve, time Gram
    SPD_TRY(sstb_coo(hS, dS, dH, K, TT, t, lead, SB_LIVE, t, st));
    SPD_TRY(dnb_gj_inv(K, dS, dH, TT, t, BG_LIVE, st));
    SPD_TRY(dnb_apply(hS, dS, dH, K, TT, t, 0, BG_LIVE, st));
    SPD_TRY(sstb_time_gram(K, dS, dH, st));
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      SPD_TRY(sstb_coo(hS, dS, dH, K, TT, m, m, SB_LIVE, t, st));
      // add_historical (StreamCPD.cc:223-294)
      int nc = 0;
      for(int mm=0; mm < N; ++mm) {
        if(mm == m || mm == t) continue;
        SPD_TRY(dnb_gram(hS, dS, dH, K, TT, mm, 1, nc, BG_LIVE, st));
        ++nc;
      }
      SPD_TRY(sstb_hist(K, dS, dH, nc, st));
      SPD_TRY(dnb_apply(hS, dS, dH, K, TT, m, 2, BG_LIVE, st));
      SPD_TRY(sstb_zero_dual(hS, dS, dH, K, m, st));
      SPD_TRY(dnb_gj_inv(K, dS, dH, TT, m, BG_LIVE, st));
      bool any_admm = false, any_closed = false;
      for(uint32_t q=0; q < K; ++q) {
        if(hS[q].kind[m] == SPD_CON_ADMM) any_admm = true;
        else any_closed = true;
      }
      if(any_admm) {
        SPD_TRY(sstb_ctl_reset(dS, dH, st));
        SPD_TRY(dnb_admm(hS, dS, dH, K, TT, m, max_inner, st));
      }
      if(any_closed) SPD_TRY(dnb_apply(hS, dS, dH, K, TT, m, 1, BG_CLOSED, st));
      SPD_TRY(dnb_gram(hS, dS, dH, K, TT, m, 0, 0, BG_LIVE, st));
      SPD_TRY(sstb_delta(hS, dS, dH, K, m, st));
    }
    SPD_TRY(sstb_outer_check(dS, dH, (int) o, st));
  }
  // forget, old <- new
  SPD_TRY(sstb_finish(K, dS, dH, st));
  // local error of each slice (StreamCPD.cc:77-89) on the mode of smallest
  // dimension (the streamed mode unless a mode has a single row)
  for(int m=0; m < N; ++m) {
    bool used = false;
    for(uint32_t q=0; q < K; ++q) used = used || hS[q].small == m;
    if(!used) continue;
    SPD_TRY(sstb_coo(hS, dS, dH, K, TT, m, lead, SB_SMALL, t, st));
    SPD_TRY(dnb_innerprod(hS, dS, dH, K, m, st));
  }
  SPD_TRY(dnb_gram(hS, dS, dH, K, TT, t, 2, 0, BG_ALL, st));
  SPD_TRY(dnb_cpd_norm(K, dS, dH, st));
  SPD_TRY(sstb_collect(K, dS, dH, e->dRes, RES_LD, st));
  SPD_HIP(hipMemcpyAsync(e->hRes, e->dRes, K * RES_LD * sizeof(double), hipMemcpyDeviceToHost,
      st));
  SPD_HIP(hipStreamSynchronize(st));
  for(uint32_t q=0; q < K; ++q) {
    BtSlice & b = *in[q];
    uint32_t const R = b.S->R;
    double const * const r = e->hRes + (uint64_t) q * RES_LD;
    b.time_row.assign(r, r + R);
    b.outer = (uint64_t) r[R];
    b.local_err = (b.xnorm + r[R + 2] - 2. * r[R + 1]) / b.xnorm;
    b.info = (int) r[R + 3];
  }
  return SPD_OK;
}
