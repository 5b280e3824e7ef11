// spd_sort.hip -- stable LSD radix sort of (uint64 key, uint32 value) pairs,
// 8-bit digits, written for 64-lane wavefronts.
//
// One block = 256 threads = 4 waves sorts a tile of 4096 keys per pass.
// Wave w owns keys [w*1024, (w+1)*1024) of the tile and walks them 64 at a
// time, so the order (wave, step, lane) is the input order and a per-wave,
// per-digit running count gives each key a stable rank: lanes holding the
// same digit in one step are found with eight ballots (one per digit bit).
#include "spd_internal.hpp"

#define RS_THREADS 256
#define RS_ITEMS 16
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_WAVES (RS_THREADS / 64)
#define RS_BINS 256

__global__ void __launch_bounds__(RS_THREADS)
k_rs_hist(uint64_t const * __restrict__ keys, uint64_t n, int shift,
    uint32_t * __restrict__ hist, uint32_t nblocks)
{
  __shared__ uint32_t h[RS_BINS];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t const base = (uint64_t) blockIdx.x * RS_TILE;
  #pragma unroll 4
  for(int k=0; k < RS_ITEMS; ++k) {
    uint64_t const i = base + (uint64_t) k * RS_THREADS + threadIdx.x;
    if(i < n) {
      atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
  }
  __syncthreads();
  hist[(uint64_t) threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_THREADS)
k_rs_scatter(uint64_t const * __restrict__ kin, uint32_t const * __restrict__ vin,
    uint64_t * __restrict__ kout, uint32_t * __restrict__ vout,
    uint64_t n, int shift, uint32_t const * __restrict__ offs, uint32_t nblocks)
{
  __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
  __shared__ uint32_t goff[RS_BINS];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  #pragma unroll
  for(int w=0; w < RS_WAVES; ++w) {
    wcnt[w][threadIdx.x] = 0;
  }
  goff[threadIdx.x] = offs[(uint64_t) threadIdx.x * nblocks + blockIdx.x];
  __syncthreads();

  uint64_t const lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint64_t const base = (uint64_t) blockIdx.x * RS_TILE + (uint64_t) wave * (RS_TILE / RS_WAVES);
  uint64_t key[RS_ITEMS];
  uint32_t val[RS_ITEMS];
  uint32_t rank[RS_ITEMS];
  #pragma unroll
  for(int j=0; j < RS_ITEMS; ++j) {
    uint64_t const i = base + (uint64_t) j * 64 + lane;
    bool const ok = i < n;
    key[j] = ok ? kin[i] : 0ull;
    val[j] = ok ? vin[i] : 0u;
  }
  #pragma unroll
  for(int j=0; j < RS_ITEMS; ++j) {
    uint64_t const i = base + (uint64_t) j * 64 + lane;
    bool const ok = i < n;
    uint32_t const d = (uint32_t) (key[j] >> shift) & 255u;
    uint64_t peers = __ballot(ok);
    #pragma unroll
    for(int b=0; b < 8; ++b) {
      bool const bit = (d >> b) & 1u;
      uint64_t const bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    uint32_t const r = (uint32_t) __popcll(peers & lt);
    uint32_t prior = 0;
    if(ok) {
      prior = wcnt[wave][d];
    }
    rank[j] = prior + r;
    if(ok && r == 0) {
      wcnt[wave][d] = prior + (uint32_t) __popcll(peers);
    }
  }
  __syncthreads();
  {
    uint32_t run = 0;
    #pragma unroll
    for(int w=0; w < RS_WAVES; ++w) {
      uint32_t const c = wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = run;
      run += c;
    }
  }
  __syncthreads();
  #pragma unroll
  for(int j=0; j < RS_ITEMS; ++j) {
    uint64_t const i = base + (uint64_t) j * 64 + lane;
    if(i < n) {
      uint32_t const d = (uint32_t) (key[j] >> shift) & 255u;
      uint32_t const dst = goff[d] + wcnt[wave][d] + rank[j];
      if(dst < n) {   // always true when the offsets are consistent
        kout[dst] = key[j];
        vout[dst] = val[j];
      }
    }
  }
}

int spd_radix_sort(uint64_t * k0, uint32_t * v0, uint64_t * k1, uint32_t * v1,
    uint64_t n, int bits, int * which, hipStream_t st)
{
  *which = 0;
  if(n <= 1 || bits <= 0) {
    return SPD_OK;
  }
  if(n >= (1ull << 32)) {
    spd_set_error(SPD_EINPUT, "radix sort supports < 2^32 keys");
    return SPD_EINPUT;
  }
  uint64_t const nb = spd_cdiv(n, RS_TILE);
  void * hp = nullptr;
  SPD_TRY(spd_scratch(0, nb * RS_BINS * sizeof(uint32_t), &hp));
  uint32_t * hist = (uint32_t *) hp;
  uint64_t * kin = k0;
  uint32_t * vin = v0;
  uint64_t * kout = k1;
  uint32_t * vout = v1;
  for(int shift=0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned) nb), dim3(RS_THREADS), 0, st,
        kin, n, shift, hist, (uint32_t) nb);
    SPD_LAUNCH_CHECK();
    SPD_TRY(spd_scan_u32(hist, hist, nb * RS_BINS, nullptr, st));
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned) nb), dim3(RS_THREADS), 0, st,
        kin, vin, kout, vout, n, shift, hist, (uint32_t) nb);
    SPD_LAUNCH_CHECK();
    uint64_t * tk = kin; kin = kout; kout = tk;
    uint32_t * tv = vin; vin = vout; vout = tv;
    *which ^= 1;
  }
  return SPD_OK;
}
