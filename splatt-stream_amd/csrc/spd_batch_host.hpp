// spd_batch_host.hpp -- host side of the batched streaming step: a prepared
// slice and the engine that runs a batch of them (spd_batch.hip).
#pragma once
#include "spd_internal.hpp"
#include "spd_stream_state.hpp"
#include <vector>

// One slice ready to run: the stream's state and the slice's device arrays
// (labels of every non-streamed mode per copy, values per copy), the rows
// it activates and where their values are, and its outputs.
struct BtSlice {
  StreamState * S = nullptr;
  uint64_t nnz = 0;
  uint64_t rows[SPD_MAX_MODES] = {};        // active rows after the slice's growth
  uint64_t grow0[SPD_MAX_MODES] = {};       // active rows before it
  double const * d_grow[SPD_MAX_MODES] = {};    // new rows' values (device)
  // copy c of the nonzeros is ordered for output mode c (the first mode
  // after the streamed one: the slice's own order)
  uint32_t const * d_ind[SPD_MAX_MODES][SPD_MAX_MODES] = {};
  double const * d_vals[SPD_MAX_MODES] = {};
  double * d_part = nullptr;                // partial sums, bt_part_doubles()
  int first = 0;
  int small = 0;
  double xnorm = 0.;
  // outputs
  std::vector<double> time_row;
  uint64_t outer = 0;
  double local_err = 0.;
  int info = 0;
};

#define BT_ENGINES 2         // batch engines of a server (each its own HIP stream)
#define BT_PREP_THREADS 8    // host threads preparing slices ahead of the device

struct BtEngine;
int bt_engine_create(BtEngine ** out);
void bt_engine_free(BtEngine * e);
hipStream_t bt_engine_stream(BtEngine * e);
int bt_engine_run(BtEngine * e, BtSlice * const * in, uint32_t K);
// doubles of partial sums a stream of rank R and row capacity cap needs
size_t bt_part_doubles(uint32_t R, uint64_t cap);
