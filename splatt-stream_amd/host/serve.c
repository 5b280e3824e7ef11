/*
 * serve.c -- the multi-stream server's C API (splatt_amd.h). No counterpart
 * in the reference, which factors one stream per process; each stream here
 * follows the reference's streaming rules (StreamCPD.cc:326-368).
 */
#include "host.h"

splatt_amd_server * splatt_amd_server_create(size_t hbm_budget_bytes, int max_inflight)
{
  spd_server * sv = NULL;
  int rc = spd_server_create(hbm_budget_bytes, max_inflight, &sv);
  if(rc != SPD_OK) {
    spl_report(rc, "splatt_amd_server_create");
    return NULL;
  }
  return (splatt_amd_server *) sv;
}

int64_t splatt_amd_server_open(splatt_amd_server * srv, splatt_idx_t nmodes,
    splatt_idx_t const * dims_cap, splatt_idx_t rank, splatt_idx_t stream_mode,
    double forget, splatt_cpd_opts const * cpd_options, unsigned seed)
{
  if(nmodes < 2 || nmodes > SPLATT_MAX_NMODES || stream_mode >= nmodes) {
    fprintf(stderr, "SPLATT: bad stream shape.\n");
    return -1;
  }
  splatt_cpd_opts * copts = (splatt_cpd_opts *) cpd_options;
  if(copts == NULL) {
    copts = splatt_alloc_cpd_opts();
  }
  spd_stream_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.stream_mode = (int) stream_mode;
  cfg.rank = rank;
  cfg.forget = forget;
  cfg.max_its = copts->max_iterations;
  cfg.tol = 5e-2;                  /* StreamCPD.cc:366-368 */
  cfg.max_inner = 25;
  cfg.inner_tol = 7e-2;
  int ok = 1;
  for(idx_t m=0; m < nmodes; ++m) {
    splatt_cpd_constraint const * con = copts->constraints[m];
    if(m != stream_mode && strcmp(con->description, "NON-NEGATIVE") == 0) {
      cfg.con[m].kind = SPD_CON_ADMM;
      cfg.con[m].prox = SPD_PROX_MAXCOLNORM_NONNEG;
    } else if(m != stream_mode && strcmp(con->description, "UNCONSTRAINED") == 0) {
      cfg.con[m].kind = SPD_CON_ADMM;
      cfg.con[m].prox = SPD_PROX_MAXCOLNORM;
    } else if(spl_map_constraint(con, m, 0, &cfg.con[m]) != 0) {
      ok = 0;
    }
  }
  if(cpd_options == NULL) {
    splatt_free_cpd_opts(copts);
  }
  if(!ok) {
    return -1;
  }
  uint64_t cap[SPLATT_MAX_NMODES];
  for(idx_t m=0; m < nmodes; ++m) cap[m] = dims_cap[m];
  int64_t id = -1;
  int rc = spd_server_open((spd_server *) srv, (int) nmodes, cap, &cfg, seed, &id);
  if(rc != SPD_OK) {
    spl_report(rc, "splatt_amd_server_open");
    return -1;
  }
  return id;
}

int64_t splatt_amd_server_submit(splatt_amd_server * srv, int64_t stream,
    splatt_idx_t nnz, splatt_idx_t const * const * inds, splatt_val_t const * vals)
{
  int64_t ticket = -1;
  int rc = spd_server_submit((spd_server *) srv, stream, nnz,
      (uint64_t const * const *) inds, vals, &ticket);
  if(rc != SPD_OK) {
    spl_report(rc, "splatt_amd_server_submit");
    return -1;
  }
  return ticket;
}

splatt_error_type splatt_amd_server_drain(splatt_amd_server * srv)
{
  return spl_report(spd_server_drain((spd_server *) srv), "splatt_amd_server_drain");
}

int splatt_amd_server_slice(splatt_amd_server * srv, int64_t ticket, double * latency_s,
    double * gpu_s, double * local_err, splatt_idx_t * outer_its)
{
  uint64_t o = 0;
  int rc = spd_server_slice((spd_server *) srv, ticket, latency_s, gpu_s, local_err, &o);
  if(outer_its) *outer_its = o;
  return rc;
}

splatt_error_type splatt_amd_server_factors(splatt_amd_server * srv, int64_t stream,
    splatt_idx_t mode, splatt_val_t * out, splatt_idx_t * rows)
{
  uint64_t r = 0;
  int rc = spd_server_factors((spd_server *) srv, stream, (int) mode, out, &r);
  if(rows) *rows = r;
  return spl_report(rc, "splatt_amd_server_factors");
}

splatt_error_type splatt_amd_server_close(splatt_amd_server * srv, int64_t stream)
{
  return spl_report(spd_server_close((spd_server *) srv, stream), "splatt_amd_server_close");
}

void splatt_amd_server_stats(splatt_amd_server * srv, size_t * used, size_t * budget,
    int * admitted, int * waiting)
{
  spd_server_stats((spd_server *) srv, used, budget, admitted, waiting);
}

void splatt_amd_server_free(splatt_amd_server * srv)
{
  spd_server_free((spd_server *) srv);
}
