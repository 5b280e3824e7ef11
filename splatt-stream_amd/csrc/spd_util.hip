// spd_util.hip -- runtime plumbing for the device ABI: error text, the
// library stream, allocation, fills, conversions and a device-wide
// exclusive scan.
#include "spd_internal.hpp"
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <functional>

static char g_err[512] = "";
static hipStream_t g_stream = nullptr;
static int g_inited = 0;
static std::mutex g_init_mu;

void spd_set_error(int code, const char * fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  int off = snprintf(g_err, sizeof(g_err), "[spd %d] ", code);
  if(off < 0 || off >= (int) sizeof(g_err)) off = 0;
  vsnprintf(g_err + off, sizeof(g_err) - off, fmt, ap);
  va_end(ap);
}

extern "C" const char * spd_last_error(void) { return g_err; }

static int init_runtime(void)
{
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if(e != hipSuccess || ndev < 1) {
    spd_set_error(SPD_ENODEV, "no HIP device available (%s)",
        e == hipSuccess ? "count 0" : hipGetErrorString(e));
    return SPD_ENODEV;
  }
  SPD_HIP(hipSetDevice(0));
  // default priority, non-blocking w.r.t. the legacy null stream
  SPD_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return SPD_OK;
}

extern "C" int spd_init(void)
{
  std::lock_guard<std::mutex> lk(g_init_mu);
  if(g_inited) {
    return SPD_OK;
  }
  // The HIP runtime's start-up draws from (or reseeds) the C library's
  // rand() generator. Callers seed rand() to fix the factor initialisation
  // (the reference draws factors from rand(), src/util.c:15-40), and the
  // first device call may come after that seed, so the start-up runs on a
  // private generator state and the caller's state is put back untouched.
  static int32_t s_rand_state[32];
  char * const prev = initstate(1u, (char *) s_rand_state, sizeof(s_rand_state));
  int const rc = init_runtime();
  setstate(prev);
  if(rc != SPD_OK) {
    return rc;
  }
  g_inited = 1;
  return SPD_OK;
}

extern "C" void * spd_stream(void) { return (void *) g_stream; }

extern "C" int spd_sync(void)
{
  SPD_HIP(hipStreamSynchronize(g_stream));
  return SPD_OK;
}

extern "C" int spd_malloc(void ** ptr, size_t bytes)
{
  *ptr = nullptr;
  if(bytes == 0) {
    bytes = 256;
  }
  hipError_t e = hipMalloc(ptr, bytes);
  if(e != hipSuccess) {
    *ptr = nullptr;
    (void) hipGetLastError();
    spd_set_error(SPD_ENOMEM, "hipMalloc(%zu) failed: %s", bytes,
        hipGetErrorString(e));
    return SPD_ENOMEM;
  }
  return SPD_OK;
}

int spd_malloc_on(void ** ptr, size_t bytes, hipStream_t st)
{
  *ptr = nullptr;
  if(bytes == 0) {
    bytes = 256;
  }
  hipError_t e = hipMallocAsync(ptr, bytes, st);
  if(e != hipSuccess) {
    *ptr = nullptr;
    (void) hipGetLastError();
    spd_set_error(SPD_ENOMEM, "hipMallocAsync(%zu) failed: %s", bytes,
        hipGetErrorString(e));
    return SPD_ENOMEM;
  }
  return SPD_OK;
}

int spd_free_on(void * ptr, hipStream_t st)
{
  if(ptr != nullptr) {
    SPD_HIP(hipFreeAsync(ptr, st));
  }
  return SPD_OK;
}

extern "C" int spd_free(void * ptr)
{
  if(ptr != nullptr) {
    SPD_HIP(hipFree(ptr));
  }
  return SPD_OK;
}

extern "C" int spd_h2d(void * dst, void const * src, size_t bytes)
{
  if(bytes == 0) return SPD_OK;
  SPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  SPD_HIP(hipStreamSynchronize(g_stream));
  return SPD_OK;
}

extern "C" int spd_d2h(void * dst, void const * src, size_t bytes)
{
  if(bytes == 0) return SPD_OK;
  SPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  SPD_HIP(hipStreamSynchronize(g_stream));
  return SPD_OK;
}

// Chunked device-to-host copy through two pinned staging buffers: chunk
// i + 1 is copied while the caller's function consumes chunk i (a DMA from
// pinned memory instead of the runtime's pageable path, and the consumer's
// stores into its own memory run beside the copy). One copy at a time per
// process (the buffers are shared).
#define SPD_STAGE_BYTES ((size_t) 64 << 20)
static std::mutex g_stage_mu;
static void * g_stage[2] = {nullptr, nullptr};
static hipEvent_t g_stage_ev[2];

int spd_d2h_chunks(void const * src, size_t bytes,
    std::function<void(void const *, size_t, size_t)> const & fn)
{
  if(bytes == 0) return SPD_OK;
  std::lock_guard<std::mutex> lk(g_stage_mu);
  if(g_stage[0] == nullptr) {
    for(int k=0; k < 2; ++k) {
      SPD_HIP(hipHostMalloc(&g_stage[k], SPD_STAGE_BYTES, hipHostMallocDefault));
      SPD_HIP(hipEventCreateWithFlags(&g_stage_ev[k], hipEventDisableTiming));
    }
  }
  hipStream_t const st = g_stream;
  size_t const nch = (bytes + SPD_STAGE_BYTES - 1) / SPD_STAGE_BYTES;
  auto issue = [&](size_t i) -> int {
    size_t const off = i * SPD_STAGE_BYTES;
    size_t const n = (off + SPD_STAGE_BYTES < bytes) ? SPD_STAGE_BYTES : bytes - off;
    SPD_HIP(hipMemcpyAsync(g_stage[i % 2], (char const *) src + off, n, hipMemcpyDeviceToHost,
        st));
    SPD_HIP(hipEventRecord(g_stage_ev[i % 2], st));
    return SPD_OK;
  };
  SPD_TRY(issue(0));
  for(size_t i=0; i < nch; ++i) {
    if(i + 1 < nch) SPD_TRY(issue(i + 1));
    SPD_HIP(hipEventSynchronize(g_stage_ev[i % 2]));
    size_t const off = i * SPD_STAGE_BYTES;
    size_t const n = (off + SPD_STAGE_BYTES < bytes) ? SPD_STAGE_BYTES : bytes - off;
    fn(g_stage[i % 2], off, n);
  }
  return SPD_OK;
}

extern "C" int spd_d2d(void * dst, void const * src, size_t bytes)
{
  if(bytes == 0) return SPD_OK;
  SPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
  return SPD_OK;
}

__global__ void k_fill_u32(uint32_t * p, uint64_t n, uint32_t v)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    p[i] = v;
  }
}

__global__ void k_fill_bytes(unsigned char * p, uint64_t n)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    p[i] = 0;
  }
}

// Zero fill with a kernel (never a memset node, so it stays correct when a
// caller captures the stream into a graph).
int spd_zero_async(void * ptr, size_t bytes, hipStream_t st)
{
  if(bytes == 0) return SPD_OK;
  uintptr_t const a = (uintptr_t) ptr;
  if((a & 3) == 0 && (bytes & 3) == 0) {
    uint64_t const n = bytes / 4;
    unsigned grid = (unsigned) (spd_cdiv(n, 256) < 8192 ? spd_cdiv(n, 256) : 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, st,
        (uint32_t *) ptr, n, 0u);
  } else {
    unsigned grid = (unsigned) (spd_cdiv(bytes, 256) < 8192 ? spd_cdiv(bytes, 256) : 8192);
    hipLaunchKernelGGL(k_fill_bytes, dim3(grid), dim3(256), 0, st,
        (unsigned char *) ptr, (uint64_t) bytes);
  }
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

extern "C" int spd_zero(void * ptr, size_t bytes)
{
  return spd_zero_async(ptr, bytes, g_stream);
}

__global__ void k_f64_to_f32(double const * __restrict__ s, float * __restrict__ d, uint64_t n)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) d[i] = (float) s[i];
}

__global__ void k_f32_to_f64(float const * __restrict__ s, double * __restrict__ d, uint64_t n)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) d[i] = (double) s[i];
}

extern "C" int spd_f64_to_f32(double const * src, float * dst, uint64_t n, void * stream)
{
  if(n == 0) return SPD_OK;
  unsigned grid = (unsigned) (spd_cdiv(n, 256) < 16384 ? spd_cdiv(n, 256) : 16384);
  hipLaunchKernelGGL(k_f64_to_f32, dim3(grid), dim3(256), 0, spd_hs(stream), src, dst, n);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

extern "C" int spd_f32_to_f64(float const * src, double * dst, uint64_t n, void * stream)
{
  if(n == 0) return SPD_OK;
  unsigned grid = (unsigned) (spd_cdiv(n, 256) < 16384 ? spd_cdiv(n, 256) : 16384);
  hipLaunchKernelGGL(k_f32_to_f64, dim3(grid), dim3(256), 0, spd_hs(stream), src, dst, n);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// ---- scratch cache -------------------------------------------------------

// Each host thread draws scratch from its own context: the library's
// context (on the library stream) unless the thread bound another one with
// spd_scratch_bind (worker threads that run on their own stream).
static SpdScratchCtx g_ctx_lib;
static thread_local SpdScratchCtx * t_ctx = nullptr;

void spd_scratch_bind(SpdScratchCtx * ctx)
{
  t_ctx = ctx;
}

SpdScratchCtx * spd_scratch_bound(void)
{
  return t_ctx;
}

void spd_scratch_release(SpdScratchCtx * ctx)
{
  if(ctx == nullptr) return;
  for(int k=0; k < SPD_NSCRATCH; ++k) {
    if(ctx->p[k]) {
      if(ctx->st) (void) spd_free_on(ctx->p[k], ctx->st);
      else (void) hipFree(ctx->p[k]);
    }
    ctx->p[k] = nullptr;
    ctx->bytes[k] = 0;
  }
}

int spd_scratch(int slot, size_t bytes, void ** ptr)
{
  if(slot < 0 || slot >= SPD_NSCRATCH) {
    spd_set_error(SPD_EINPUT, "bad scratch slot %d", slot);
    return SPD_EINPUT;
  }
  SpdScratchCtx * c = t_ctx ? t_ctx : &g_ctx_lib;
  hipStream_t const st = (c == &g_ctx_lib) ? g_stream : c->st;
  if(c->bytes[slot] < bytes || c->p[slot] == nullptr) {
    bool const own = (c != &g_ctx_lib);
    if(c->p[slot] != nullptr) {
      // the stream may still use the old buffer: a context with its own
      // stream frees in stream order, the library's waits for its stream
      if(own) {
        SPD_TRY(spd_free_on(c->p[slot], st));
      } else {
        SPD_HIP(hipStreamSynchronize(st));
        SPD_HIP(hipFree(c->p[slot]));
      }
      c->p[slot] = nullptr;
      c->bytes[slot] = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    want += want / 8;
    if(own) SPD_TRY(spd_malloc_on(&c->p[slot], want, st));
    else SPD_TRY(spd_malloc(&c->p[slot], want));
    c->bytes[slot] = want;
  }
  *ptr = c->p[slot];
  return SPD_OK;
}

// ---- exclusive scan ------------------------------------------------------

#define SCAN_THREADS 256
#define SCAN_ITEMS 16
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t * wtot,
    uint32_t * total)
{
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  uint32_t inc = v;
  #pragma unroll
  for(int o=1; o < 64; o <<= 1) {
    uint32_t const t = __shfl_up(inc, o, 64);
    if(lane >= o) inc += t;
  }
  if(lane == 63) wtot[wave] = inc;
  __syncthreads();
  uint32_t woff = 0;
  uint32_t tot = 0;
  #pragma unroll
  for(int w=0; w < SCAN_THREADS / 64; ++w) {
    uint32_t const x = wtot[w];
    if(w < wave) woff += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return woff + inc - v;
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_sums(uint32_t const * __restrict__ in, uint64_t n, uint32_t * __restrict__ sums)
{
  __shared__ uint32_t wtot[SCAN_THREADS / 64];
  uint64_t const base = (uint64_t) blockIdx.x * SCAN_TILE + (uint64_t) threadIdx.x * SCAN_ITEMS;
  uint32_t s = 0;
  #pragma unroll
  for(int k=0; k < SCAN_ITEMS; ++k) {
    uint64_t const i = base + k;
    if(i < n) s += in[i];
  }
  uint32_t tot;
  (void) block_excl_scan(s, wtot, &tot);
  if(threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_apply(uint32_t const * in, uint32_t * out, uint64_t n,
    uint32_t const * __restrict__ offs)
{
  __shared__ uint32_t wtot[SCAN_THREADS / 64];
  uint64_t const base = (uint64_t) blockIdx.x * SCAN_TILE + (uint64_t) threadIdx.x * SCAN_ITEMS;
  uint32_t vals[SCAN_ITEMS];
  uint32_t s = 0;
  #pragma unroll
  for(int k=0; k < SCAN_ITEMS; ++k) {
    uint64_t const i = base + k;
    vals[k] = (i < n) ? in[i] : 0u;
    s += vals[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, wtot, &tot) + offs[blockIdx.x];
  #pragma unroll
  for(int k=0; k < SCAN_ITEMS; ++k) {
    uint64_t const i = base + k;
    if(i < n) out[i] = run;
    run += vals[k];
  }
}

static int scan_rec(uint32_t const * in, uint32_t * out, uint64_t n,
    hipStream_t st, int depth)
{
  uint64_t const nb = spd_cdiv(n, SCAN_TILE);
  void * sp = nullptr;
  SPD_TRY(spd_scratch(8 + depth, (nb + 1) * sizeof(uint32_t), &sp));
  uint32_t * sums = (uint32_t *) sp;
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned) nb), dim3(SCAN_THREADS), 0, st,
      in, n, sums);
  SPD_LAUNCH_CHECK();
  if(nb > 1) {
    SPD_TRY(scan_rec(sums, sums, nb, st, depth + 1));
  } else {
    // a single block: its offset is zero
    SPD_TRY(spd_zero_async(sums, sizeof(uint32_t), st));
  }
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned) nb), dim3(SCAN_THREADS), 0, st,
      in, out, n, sums);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// The last element's inclusive sum equals out[n-1] + in[n-1]; read both to
// report the total (before the in-place overwrite would lose in[n-1]).
int spd_scan_u32(uint32_t const * in, uint32_t * out, uint64_t n,
    uint32_t * total, hipStream_t st)
{
  if(n == 0) {
    if(total) *total = 0;
    return SPD_OK;
  }
  if(n > ((uint64_t) SCAN_TILE) * SCAN_TILE * SCAN_TILE) {
    spd_set_error(SPD_EINPUT, "scan too long (%llu)", (unsigned long long) n);
    return SPD_EINPUT;
  }
  uint32_t last_in = 0;
  if(total) {
    SPD_HIP(hipMemcpyAsync(&last_in, in + (n - 1), sizeof(uint32_t),
        hipMemcpyDeviceToHost, st));
    SPD_HIP(hipStreamSynchronize(st));
  }
  SPD_TRY(scan_rec(in, out, n, st, 0));
  if(total) {
    uint32_t last_out = 0;
    SPD_HIP(hipMemcpyAsync(&last_out, out + (n - 1), sizeof(uint32_t),
        hipMemcpyDeviceToHost, st));
    SPD_HIP(hipStreamSynchronize(st));
    *total = last_out + last_in;
  }
  return SPD_OK;
}
