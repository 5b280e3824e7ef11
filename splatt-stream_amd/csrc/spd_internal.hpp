// spd_internal.hpp -- shared helpers for the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "spd.h"
#include <functional>
#include <thread>
#include <vector>

// Records an error message for spd_last_error().
void spd_set_error(int code, const char * fmt, ...);

#define SPD_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if(e_ != hipSuccess) {                                                   \
      spd_set_error(SPD_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,     \
          hipGetErrorString(e_));                                            \
      return SPD_EHIP;                                                       \
    }                                                                        \
  } while(0)

#define SPD_TRY(call)                                                        \
  do {                                                                       \
    int r_ = (call);                                                         \
    if(r_ != SPD_OK) {                                                       \
      return r_;                                                             \
    }                                                                        \
  } while(0)

#define SPD_LAUNCH_CHECK()                                                   \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if(e_ != hipSuccess) {                                                   \
      spd_set_error(SPD_EHIP, "%s:%d launch: %s", __FILE__, __LINE__,        \
          hipGetErrorString(e_));                                            \
      return SPD_EHIP;                                                       \
    }                                                                        \
  } while(0)

static inline hipStream_t spd_hs(void * s)
{
  return s ? (hipStream_t) s : (hipStream_t) spd_stream();
}

__host__ __device__ static inline uint64_t spd_cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Exclusive prefix sum of n uint32 values (in place allowed: in == out).
// Returns the total in *total (host) if total != NULL (synchronises).
int spd_scan_u32(uint32_t const * in, uint32_t * out, uint64_t n,
    uint32_t * total, hipStream_t st);

// Scratch device memory owned by the library, grown on demand. Slot k keeps
// its buffer until the next request for a larger size. Slots: 0 radix
// histograms, 1 norms, 2-3 root carries, 4-7 dense partials, 8-12 scan
// levels, 13-14 streaming partials.
#define SPD_NSCRATCH 16
struct SpdScratchCtx {
  void * p[SPD_NSCRATCH];
  size_t bytes[SPD_NSCRATCH];
  hipStream_t st;            // the stream that uses these buffers
};
int spd_scratch(int slot, size_t bytes, void ** ptr);
// Route this thread's scratch requests to ctx (nullptr: the library's own).
void spd_scratch_bind(SpdScratchCtx * ctx);
// Free a context's buffers (stream-ordered on its stream).
void spd_scratch_release(SpdScratchCtx * ctx);
// The context bound to this thread (nullptr: the library's own).
SpdScratchCtx * spd_scratch_bound(void);

// Scratch for one call on a caller's stream. Entry points that accept any
// hipStream_t and return without synchronising must not share the library's
// scratch: two calls on two streams would write the same buffers at once.
// When the thread has no context bound and st is not the library stream,
// this binds a context of its own for the call; its buffers are allocated
// and freed in stream order on st (hipMallocAsync / hipFreeAsync), so they
// stay alive for the kernels the call queued and no call waits for another.
struct SpdCallScratch {
  SpdScratchCtx ctx;
  SpdScratchCtx * prev;
  bool active;
  explicit SpdCallScratch(hipStream_t st) : ctx(), prev(spd_scratch_bound()), active(false)
  {
    if(prev == nullptr && st != (hipStream_t) spd_stream()) {
      ctx.st = st;
      spd_scratch_bind(&ctx);
      active = true;
    }
  }
  ~SpdCallScratch()
  {
    if(active) {
      spd_scratch_release(&ctx);
      spd_scratch_bind(prev);
    }
  }
  SpdCallScratch(SpdCallScratch const &) = delete;
  SpdCallScratch & operator=(SpdCallScratch const &) = delete;
};

// Stream-ordered device allocation from the device's memory pool: no
// device-wide synchronisation (hipFree waits for the whole device), the
// memory is usable by work queued on st after the call and is returned
// after the work queued before the free.
int spd_malloc_on(void ** ptr, size_t bytes, hipStream_t st);
int spd_free_on(void * ptr, hipStream_t st);

// Device-to-host copy of `bytes` in chunks through pinned staging on the
// library stream: fn(host_chunk, byte_offset, chunk_bytes) is called for
// each chunk in order (on the calling thread) while the next one copies.
int spd_d2h_chunks(void const * src, size_t bytes,
    std::function<void(void const *, size_t, size_t)> const & fn);

// Run f(lo, hi) over [0, n) split across up to 16 host threads (one range
// each; small n runs on the calling thread).
template<typename F>
void spd_par_range(uint64_t n, F const & f)
{
  unsigned hc = std::thread::hardware_concurrency();
  unsigned const T = (hc == 0) ? 1u : (hc < 16u ? hc : 16u);
  if(T == 1 || n < ((uint64_t) 1 << 18)) {
    f((uint64_t) 0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  for(unsigned k=0; k < T; ++k) {
    uint64_t const lo = n * k / T, hi = n * (k + 1) / T;
    th.emplace_back([&f, lo, hi] { f(lo, hi); });
  }
  for(std::thread & x : th) x.join();
}

// Zero fill with a kernel on stream st.
int spd_zero_async(void * ptr, size_t bytes, hipStream_t st);

// Sort (key, value) pairs by the low `bits` bits of key, stable. Ping-pongs
// between (k0,v0) and (k1,v1); on return *which tells where the result is
// (0: k0/v0, 1: k1/v1).
int spd_radix_sort(uint64_t * k0, uint32_t * v0, uint64_t * k1, uint32_t * v1,
    uint64_t n, int bits, int * which, hipStream_t st);

// Device-wide wavefront helpers (64 lanes on gfx950).
__device__ __forceinline__ int spd_lane() { return __lane_id(); }
