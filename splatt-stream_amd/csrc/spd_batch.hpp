// spd_batch.hpp -- one time slice of many independent streams, stepped in
// lockstep by kernels that each cover every slice of the batch.
//
// A batch holds up to BT_MAXK slices of different streams with the same
// mode count, streamed mode and padded rank (R rounded up to 16). Every
// batched kernel receives the slices' descriptors (BSlot, device memory)
// and the batch header (BHead). A kernel that the per-slice path launches
// with nb(slice) blocks runs here as work items (slice s, block b < nb(s)):
// each block of the batched launch computes nb(s) of every slice from the
// slice's sizes, prefix-sums them in LDS, and strides over the items. Item
// (s, b) executes exactly what block b of the per-slice launch executes
// (same rows, same partial-sum slots, same reduction order), so a slice's
// result does not depend on the other slices of its batch nor on the grid
// of the batched launch.
//
// The outer ALS loop and the ADMM inner loop test their convergence on the
// device: per slice flags in BCtl make later kernels skip a converged
// slice (its items are not counted).
#pragma once
#include "spd_internal.hpp"

#define BT_MAXK 64           // slices per batch, at most
#define BT_GRID 2048         // blocks of an item-strided batched launch, at most

struct BCtl {
  int done;                 // outer loop finished
  int outer;                // outer iterations run
  int admm[2];              // ADMM: [0] converged, [1] index of the final state
  int info;                 // Cholesky / Gauss-Jordan status
  int pad;
  double delta;             // change measure of this outer iteration
  double prev_delta;
  double rho;
  double diff[2];           // sum (old - new)^2, sum new^2
  double sums[4];           // ADMM test terms
  double inner[2];          // local error: <X, Z>, ||Z||^2
  double nrm[3][64];        // prox column norms of the three ADMM states
};

struct BSlot {
  // stream state (allocated by the stream, sized for its capacities)
  double * Unew[SPD_MAX_MODES];
  double * Uold[SPD_MAX_MODES];
  double * Dual[SPD_MAX_MODES];
  double * aTa[SPD_MAX_MODES];
  double * time_row;
  double * Mbuf;
  double * aux;
  double * init;
  double * old_gram;
  double * ginv;
  double * H;
  double * cross;           // SPD_MAX_MODES x R x R
  double * tgram;
  double * lam1;            // R ones
  double * tmpv;            // R
  double * part;            // partial sums of this slice's reductions
  BCtl * c;
  // the slice: copy q of the nonzeros is ordered for output mode q (the
  // copy of the first mode after the streamed one is the slice's own
  // order); ind[q][t] points at zeros (the streamed mode's label is 0)
  uint32_t const * ind[SPD_MAX_MODES][SPD_MAX_MODES];
  double const * vals[SPD_MAX_MODES];
  uint64_t nnz;
  uint64_t rows[SPD_MAX_MODES];     // active rows this slice (time: 1)
  uint64_t grow0[SPD_MAX_MODES];    // rows before this slice's growth
  double const * grow[SPD_MAX_MODES];   // values of the new rows
  int first;                        // the stream's first slice
  int small;                        // mode of the local-error MTTKRP
  // configuration
  int kind[SPD_MAX_MODES];          // SPD_CON_*
  int prox[SPD_MAX_MODES];
  double diag_add[SPD_MAX_MODES];
  double lam[SPD_MAX_MODES];
  double forget;
  double tol;
  double inner_tol;
  uint32_t max_inner;
  uint32_t max_its;
  uint32_t R;
};

struct BHead {
  uint32_t K;               // slices in the batch
  int N;                    // modes
  int t;                    // streamed mode
  int lead;                 // first mode after the streamed one
};

// Work items of a batched launch. Every block calls bt_items() (all threads)
// and then strides over [0, total): bt_find() maps an item to its slice.
struct BItems {
  uint32_t pref[BT_MAXK + 1];
  uint32_t total;
};

// s_it.pref[s + 1] - s_it.pref[s] = nb (slice s), from the per-lane counts
// computed by the first wave.
__device__ __forceinline__ void bt_items(BItems & s_it, uint32_t K, uint32_t my_nb)
{
  if(threadIdx.x < 64) {
    uint32_t v = ((uint32_t) threadIdx.x < K) ? my_nb : 0u;
    #pragma unroll
    for(int o=1; o < 64; o <<= 1) {
      uint32_t const u = __shfl_up(v, o, 64);
      if((int) threadIdx.x >= o) v += u;
    }
    s_it.pref[threadIdx.x + 1] = v;
    if(threadIdx.x == 0) s_it.pref[0] = 0;
    if(threadIdx.x == 63) s_it.total = v;
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t bt_find(BItems const & s_it, uint32_t K, uint32_t w)
{
  uint32_t lo = 0, hi = K;      // pref[lo] <= w < pref[hi]
  while(hi - lo > 1) {
    uint32_t const mid = (lo + hi) >> 1;
    if(s_it.pref[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}
