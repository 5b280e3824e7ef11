/*
 * mttkrp.c -- the MTTKRP entry points (src/mttkrp.c:1287-1343, :1775-1945)
 * over the GPU kernels.
 */
#include "host.h"
#include <pthread.h>

/* Mode -> CSF map, src/mttkrp.c:1847-1881. */
static int p_mode_csf_map(splatt_csf const * tensors, splatt_csf_type which,
    idx_t * map, idx_t * num_csf)
{
  idx_t const N = tensors->nmodes;
  for(idx_t m=0; m < N; ++m) {
    switch(which) {
    case SPLATT_CSF_ONEMODE:
      map[m] = 0;
      *num_csf = 1;
      break;
    case SPLATT_CSF_TWOMODE:
      map[m] = (splatt_csf_mode_to_depth(tensors, m) == N - 1) ? 1 : 0;
      *num_csf = 2;
      break;
    case SPLATT_CSF_ALLMODE:
      map[m] = m;
      *num_csf = N;
      break;
    default:
      return -1;
    }
  }
  return 0;
}

splatt_mttkrp_ws * splatt_mttkrp_alloc_ws(splatt_csf const * const tensors,
    splatt_idx_t const ncolumns, double const * const opts)
{
  (void) ncolumns;
  splatt_mttkrp_ws * ws = calloc(1, sizeof(*ws));
  idx_t const nthreads = (idx_t) opts[SPLATT_OPTION_NTHREADS];
  ws->num_threads = nthreads;
  if(p_mode_csf_map(tensors, (splatt_csf_type) opts[SPLATT_OPTION_CSF_ALLOC],
      ws->mode_csf_map, &ws->num_csf) != 0) {
    fprintf(stderr, "SPLATT: CSF type '%d' not recognized.\n",
        (int) opts[SPLATT_OPTION_CSF_ALLOC]);
    free(ws);
    return NULL;
  }
  /* privatisation flags as the reference computes them (src/mttkrp.c:221);
   * the GPU kernels do not need the buffers */
  double const thresh = opts[SPLATT_OPTION_PRIVTHRESH];
  for(idx_t m=0; m < tensors->nmodes; ++m) {
    ws->is_privatized[m] = (nthreads > 1) &&
        ((double) (tensors->dims[m] * nthreads) <= thresh * (double) tensors->nnz);
  }
  ws->privatize_buffer = NULL;
  ws->reduction_time = 0;
  return ws;
}

void splatt_mttkrp_free_ws(splatt_mttkrp_ws * const ws)
{
  free(ws);
}

/* Device buffers of the host-array MTTKRP entry points, kept between calls
 * and grown on demand (hipMalloc and hipFree synchronise the whole device):
 * one per input mode, the output, and an FP64 staging buffer for FP32 runs.
 * Calls are serialised on a mutex (the reference's splatt_mttkrp is not
 * re-entrant either: src/mttkrp.c:15, :1298-1300). */
static struct {
  pthread_mutex_t mu;
  void * buf[SPLATT_MAX_NMODES + 2];
  size_t cap[SPLATT_MAX_NMODES + 2];
} g_hws = { PTHREAD_MUTEX_INITIALIZER, {0}, {0} };

static int p_hws_get(int slot, size_t bytes, void ** out)
{
  if(g_hws.cap[slot] < bytes) {
    spd_free(g_hws.buf[slot]);
    g_hws.buf[slot] = NULL;
    g_hws.cap[slot] = 0;
    size_t want = bytes + bytes / 8;
    int rc = spd_malloc(&g_hws.buf[slot], want);
    if(rc != SPD_OK) {
      return rc;
    }
    g_hws.cap[slot] = want;
  }
  *out = g_hws.buf[slot];
  return SPD_OK;
}

/* MTTKRP of host factors through the GPU: upload the inputs, run, download
 * the output. mats[m] (host, dims[m] x R) for m != mode. */
static int p_mttkrp_host_locked(splatt_csf const * ct, idx_t mode, idx_t R,
    val_t * const * mats, val_t * out)
{
  idx_t const N = ct->nmodes;
  int const dtype = spl_mttkrp_dtype();
  size_t const sv = (dtype == SPD_F32) ? sizeof(float) : sizeof(double);
  void * dmats[SPLATT_MAX_NMODES] = {0};
  int const STAGE = SPLATT_MAX_NMODES;
  int const OUT = SPLATT_MAX_NMODES + 1;
  void * stage = NULL;
  void * dout = NULL;
  int rc = SPD_OK;
  for(idx_t m=0; m < N && rc == SPD_OK; ++m) {
    if(m == mode) {
      continue;
    }
    size_t const n = ct->dims[m] * R;
    rc = p_hws_get((int) m, n * sv, &dmats[m]);
    if(rc != SPD_OK) break;
    if(dtype == SPD_F32) {
      rc = p_hws_get(STAGE, n * sizeof(double), &stage);
      if(rc == SPD_OK) rc = spd_h2d(stage, mats[m], n * sizeof(double));
      if(rc == SPD_OK) rc = spd_f64_to_f32((double const *) stage, (float *) dmats[m], n, NULL);
    } else {
      rc = spd_h2d(dmats[m], mats[m], n * sizeof(double));
    }
  }
  size_t const nout = ct->dims[mode] * R;
  if(rc == SPD_OK) rc = p_hws_get(OUT, nout * sv, &dout);
  if(rc == SPD_OK) {
    rc = spd_mttkrp((spd_csf *) ct->dev, (int) mode, R, dmats, dout, dtype,
        ct->ntiles == 1, NULL);
  }
  if(rc == SPD_OK) {
    if(dtype == SPD_F32) {
      rc = p_hws_get(STAGE, nout * sizeof(double), &stage);
      if(rc == SPD_OK) rc = spd_f32_to_f64((float const *) dout, (double *) stage, nout, NULL);
      if(rc == SPD_OK) rc = spd_d2h(out, stage, nout * sizeof(double));
    } else {
      rc = spd_d2h(out, dout, nout * sizeof(double));
    }
  }
  return rc;
}

static int p_mttkrp_host(splatt_csf const * ct, idx_t mode, idx_t R,
    val_t * const * mats, val_t * out)
{
  if(ct->dev == NULL) {
    fprintf(stderr, "SPLATT: MTTKRP needs a tensor built by this library.\n");
    return SPD_EINPUT;
  }
  pthread_mutex_lock(&g_hws.mu);
  int const rc = p_mttkrp_host_locked(ct, mode, R, mats, out);
  pthread_mutex_unlock(&g_hws.mu);
  return rc;
}

int splatt_mttkrp(splatt_idx_t const mode, splatt_idx_t const ncolumns,
    splatt_csf const * const tensors, splatt_val_t ** matrices,
    splatt_val_t * const matout, double const * const options)
{
  idx_t map[SPLATT_MAX_NMODES];
  idx_t ncsf = 0;
  if(mode >= tensors->nmodes || p_mode_csf_map(tensors,
        (splatt_csf_type) options[SPLATT_OPTION_CSF_ALLOC], map, &ncsf) != 0) {
    return SPLATT_ERROR_BADINPUT;
  }
  int rc = p_mttkrp_host(tensors + map[mode], mode, ncolumns, matrices, matout);
  return (int) spl_report(rc, "splatt_mttkrp");
}

void splatt_mttkrp_csf(splatt_csf const * const tensors, matrix_t ** mats,
    splatt_idx_t const mode, void * const thds, splatt_mttkrp_ws * const ws,
    splatt_global_opts const * const global_opts)
{
  (void) thds;
  (void) global_opts;
  matrix_t * M = mats[SPLATT_MAX_NMODES];
  M->I = tensors[0].dims[mode];
  val_t * host[SPLATT_MAX_NMODES] = {0};
  for(idx_t m=0; m < tensors->nmodes; ++m) {
    host[m] = mats[m]->vals;
  }
  int rc = p_mttkrp_host(tensors + ws->mode_csf_map[mode], mode, M->J, host,
      M->vals);
  spl_report(rc, "mttkrp_csf");
}

splatt_error_type splatt_amd_mttkrp_dev(splatt_idx_t const mode,
    splatt_idx_t const ncolumns, splatt_csf const * const tensors,
    void ** d_mats, void * d_out, int precision, double const * const options,
    void * stream)
{
  idx_t map[SPLATT_MAX_NMODES];
  idx_t ncsf = 0;
  if(mode >= tensors->nmodes || p_mode_csf_map(tensors,
        (splatt_csf_type) options[SPLATT_OPTION_CSF_ALLOC], map, &ncsf) != 0) {
    return SPLATT_ERROR_BADINPUT;
  }
  splatt_csf const * ct = tensors + map[mode];
  if(ct->dev == NULL) {
    return SPLATT_ERROR_BADINPUT;
  }
  int const dtype = (precision == SPLATT_AMD_F32) ? SPD_F32 : SPD_F64;
  int rc = spd_mttkrp((spd_csf *) ct->dev, (int) mode, ncolumns, d_mats, d_out,
      dtype, ct->ntiles == 1, stream);
  return spl_report(rc, "splatt_amd_mttkrp_dev");
}

splatt_error_type splatt_amd_mttkrp_dev_noalloc(splatt_idx_t const mode,
    splatt_idx_t const ncolumns, splatt_csf const * const tensors,
    void ** d_mats, void * d_out, int precision, double const * const options,
    void * stream)
{
  idx_t map[SPLATT_MAX_NMODES];
  idx_t ncsf = 0;
  if(mode >= tensors->nmodes || p_mode_csf_map(tensors,
        (splatt_csf_type) options[SPLATT_OPTION_CSF_ALLOC], map, &ncsf) != 0) {
    return SPLATT_ERROR_BADINPUT;
  }
  splatt_csf const * ct = tensors + map[mode];
  if(ct->dev == NULL) {
    return SPLATT_ERROR_BADINPUT;
  }
  int const dtype = (precision == SPLATT_AMD_F32) ? SPD_F32 : SPD_F64;
  int rc = spd_mttkrp_noalloc((spd_csf *) ct->dev, (int) mode, ncolumns, d_mats, d_out,
      dtype, ct->ntiles == 1, stream);
  return spl_report(rc, "splatt_amd_mttkrp_dev_noalloc");
}

size_t splatt_amd_mttkrp_noalloc_bytes(void)
{
  return spd_mttkrp_noalloc_bytes();
}

splatt_error_type splatt_amd_mttkrp_cost(splatt_idx_t const mode,
    splatt_idx_t const ncolumns, splatt_csf const * const tensors,
    int precision, double const * const options, double * bytes, double * flops)
{
  idx_t map[SPLATT_MAX_NMODES];
  idx_t ncsf = 0;
  if(mode >= tensors->nmodes || p_mode_csf_map(tensors,
        (splatt_csf_type) options[SPLATT_OPTION_CSF_ALLOC], map, &ncsf) != 0) {
    return SPLATT_ERROR_BADINPUT;
  }
  splatt_csf const * ct = tensors + map[mode];
  if(ct->dev == NULL) {
    return SPLATT_ERROR_BADINPUT;
  }
  int rc = spd_mttkrp_model((spd_csf const *) ct->dev, (int) mode, ncolumns,
      precision == SPLATT_AMD_F32 ? SPD_F32 : SPD_F64, bytes, flops);
  return spl_report(rc, "splatt_amd_mttkrp_cost");
}

int spd_mttkrp_timing(int enable);

void splatt_amd_mttkrp_timing(int enable)
{
  spd_mttkrp_timing(enable);
}

double splatt_amd_mttkrp_last_ms(void)
{
  return spd_mttkrp_last_kernel_ms();
}

int splatt_amd_mttkrp_kernel_times(double * ms, int max)
{
  return spd_mttkrp_kernel_times(ms, max);
}

void * splatt_amd_dev_alloc(size_t bytes)
{
  void * p = NULL;
  if(spd_init() != SPD_OK || spd_malloc(&p, bytes) != SPD_OK) {
    return NULL;
  }
  return p;
}

void splatt_amd_dev_free(void * ptr)
{
  spd_free(ptr);
}

splatt_error_type splatt_amd_h2d(void * dst, void const * src, size_t bytes)
{
  return spl_report(spd_h2d(dst, src, bytes), "h2d");
}

splatt_error_type splatt_amd_d2h(void * dst, void const * src, size_t bytes)
{
  return spl_report(spd_d2h(dst, src, bytes), "d2h");
}

splatt_error_type splatt_amd_sync(void)
{
  return spl_report(spd_sync(), "sync");
}
