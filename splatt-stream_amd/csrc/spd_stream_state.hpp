// spd_stream_state.hpp -- the per-stream device state of streaming CPD and
// its per-slice step, shared by the whole-stream driver (spd_stream_cpd)
// and the multi-stream server.
#pragma once
#include "spd_internal.hpp"
#include <vector>

// Device allocations freed together; stream-ordered on st when it is set
// (allocation and free then never wait for the rest of the device).
struct SpdMem {
  std::vector<void *> ptrs;
  hipStream_t st = nullptr;
  ~SpdMem() {
    for(void * p : ptrs) {
      if(st) (void) spd_free_on(p, st);
      else spd_free(p);
    }
  }
  int get(void ** out, size_t bytes) {
    *out = nullptr;
    if(bytes == 0) bytes = 8;
    if(st) SPD_TRY(spd_malloc_on(out, bytes, st));
    else SPD_TRY(spd_malloc(out, bytes));
    ptrs.push_back(*out);
    return SPD_OK;
  }
};

// One stream's factors, Grams and work buffers, allocated for fixed row
// capacities; everything runs on `st`.
struct StreamState {
  int N = 0;
  int t = 0;
  uint32_t R = 0;
  uint32_t RP = 0;
  spd_stream_cfg cfg = {};
  hipStream_t st = nullptr;
  uint64_t cap[SPD_MAX_MODES] = {};
  uint64_t rows[SPD_MAX_MODES] = {};       // active factor rows
  uint64_t prev_dim[SPD_MAX_MODES] = {};   // running dims of the slices
  uint64_t batch_max[SPD_MAX_MODES] = {};  // largest label of the next slice (set by the caller)
  uint64_t nslices = 0;
  SpdMem mem;
  double * Unew[SPD_MAX_MODES] = {};
  double * Uold[SPD_MAX_MODES] = {};
  double * Dual[SPD_MAX_MODES] = {};
  double * aTa[SPD_MAX_MODES] = {};
  double * time_row = nullptr;
  double * Mbuf = nullptr;
  double * aux = nullptr;
  double * init_buf = nullptr;
  double * old_gram = nullptr;
  double * ginv = nullptr;
  double * H = nullptr;
  double * cross = nullptr;
  double * tgram = nullptr;
  double * sc = nullptr;       // [0] rho, [4..5] error terms, [8..] ADMM, [20 + R ..] change sums per mode
  double * lam1 = nullptr;
  double * tmpv = nullptr;
  int * info = nullptr;
  int * ctl = nullptr;
  uint32_t * d_zero_ind = nullptr;
  uint64_t zero_cap = 0;
  // the slice re-sorted by each mode whose rows do not come in runs in the
  // slice's own order (every mode but the streamed one and the one after
  // it): by_ind[m][k] is mode k's index array of that copy, by_vals[m] its
  // values; sort keys and payloads; sized for sort_cap nonzeros
  uint32_t * by_ind[SPD_MAX_MODES][SPD_MAX_MODES] = {};
  double * by_vals[SPD_MAX_MODES] = {};
  uint64_t * sk[2] = {};
  uint32_t * sv[2] = {};
  uint64_t sort_cap = 0;
  SpdMem sort_mem;
  std::vector<double> hrand;

  // device bytes init() allocates for these capacities
  static size_t footprint(int N, int t, uint64_t R, uint64_t const * cap);
  // ordered != 0: allocate and free in stream order on the stream in use
  int init(int nmodes, spd_stream_cfg const & c, uint64_t const * caps, hipStream_t stream,
      int ordered = 0);
  // Run later slices on another stream (the work queued so far on the old
  // one must be complete).
  void set_stream(hipStream_t stream);
  // One time slice: ind[m] (device, the slice's nonzeros, labels of every
  // mode but the streamed one) and vals (device); xnorm = sum of squared
  // values. batch_max[] must hold the slice's largest label per mode.
  int slice(uint32_t const * const * ind, double const * vals, uint64_t nnz, double xnorm,
      double * local_err, uint64_t * outer_its, double * seconds);
  int cholesky_info(int * out);
};

int sst_check_cfg(int N, spd_stream_cfg const & cfg);
int sst_coo_mttkrp(uint32_t const * const * ind, double const * vals, uint64_t nnz,
    int N, int mode, double const * const * mats, uint32_t R, uint64_t rows, double * out,
    hipStream_t st);
double sst_wall_s();
