// spd_stream.hip -- streaming CPD on the GPU: the reference's
// StreamCPD::compute (src/cpd/stream/StreamCPD.cc:311-571) over a tensor
// that is resident in HBM for the whole stream.
//
// The whole stream (sorted by time, other modes relabelled by first
// appearance) is uploaded once; a time slice is a contiguous range of it, so
// a batch costs no copy. Factor matrices are allocated at their final row
// counts up front; "growing" a factor only activates rows (filled with the
// reference's rand() sequence on the host) and zero rows of the history and
// dual matrices stay zero. Per slice everything runs on the device: the COO
// MTTKRP (src/mttkrp.c:1699-1769: atomics for scattered rows, fixed-order
// block partials for the single time row), the closed-form time solve, the
// time Gram update, the history term (cross Grams on MFMA and a product
// with the old factor), ADMM with the max-column-norm prox, the Grams and
// the change measure. ADMM tests its convergence on the device; the host
// reads the two scalars of the outer change measure per mode.
#include "spd_internal.hpp"
#include "spd_dense.hpp"
#include "spd_batch.hpp"
#include "spd_stream_state.hpp"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <vector>

#define CO_THREADS 256

struct CooArgs {
  uint32_t const * ind[SPD_MAX_MODES];
  double const * vals;
  double const * mats[SPD_MAX_MODES];
  double * out;        // rows > 1: atomic target
  double * part;       // single row: per-block partial rows
  uint64_t nnz;
  uint32_t R;
  int N;
  int mode;
  int reduce;
};

// COO MTTKRP: per nonzero v * prod_{m != mode} U_m[ind_m] (modes ascending,
// as the reference multiplies), added to the output row. L lanes per
// nonzero, NC column chunks of L. A group takes CO_CHUNK consecutive
// nonzeros: its first lanes load their values and indices once (coalesced),
// the products of CO_BATCH nonzeros are formed with every row load issued
// before any is used, then runs of equal output rows are added up and
// written with one atomic update per run (a slice is sorted by the mode
// after the streamed one, so that mode's rows come in runs).
#define CO_CHUNK 16
#define CO_BATCH 8
template<int L, int NC>
__device__ __forceinline__ void
b_coo_mttkrp(CooArgs a,
    uint32_t blk, uint32_t nblk)
{
  constexpr int G = 64 / L;
  constexpr int W = CO_THREADS / 64;
  __shared__ double s_red[W * G][L * NC];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const c = lane % L;
  int const g = lane / L;
  int const base = g * L;
  uint32_t const R = a.R;
  int const N = a.N;
  int const mode = a.mode;
  uint64_t const group = ((uint64_t) blk * W + wave) * G + g;
  uint64_t const ngroups = (uint64_t) nblk * W * G;
  double acc[NC];
  #pragma unroll
  for(int k=0; k < NC; ++k) acc[k] = 0.;
  uint32_t cur = 0xFFFFFFFFu;
  for(uint64_t n0 = group * CO_CHUNK; n0 < a.nnz; n0 += ngroups * CO_CHUNK) {
    uint32_t const cnt = (n0 + CO_CHUNK < a.nnz) ? CO_CHUNK : (uint32_t) (a.nnz - n0);
    // lane c < cnt of the group holds nonzero n0 + c
    double myv = 0.;
    uint32_t myo = 0;
    uint32_t myi[SPD_MAX_MODES];
    #pragma unroll
    for(int m=0; m < SPD_MAX_MODES; ++m) myi[m] = 0;
    if((uint32_t) c < cnt) {
      myv = a.vals[n0 + c];
      #pragma unroll
      for(int m=0; m < SPD_MAX_MODES; ++m) {
        if(m < N && m != mode) myi[m] = a.ind[m][n0 + c];
      }
      if(!a.reduce) myo = a.ind[mode][n0 + c];
    }
    for(uint32_t k0=0; k0 < cnt; k0 += CO_BATCH) {
      double p[CO_BATCH][NC];
      uint32_t orow[CO_BATCH];
      #pragma unroll
      for(int k=0; k < CO_BATCH; ++k) {
        int const src = base + (int) k0 + k;
        bool const live = (k0 + k) < cnt;
        double const v = __shfl(myv, src, 64);
        orow[k] = __shfl(myo, src, 64);
        #pragma unroll
        for(int q=0; q < NC; ++q) p[k][q] = v;
        #pragma unroll
        for(int m=0; m < SPD_MAX_MODES; ++m) {
          if(m >= N) break;
          if(m == mode) continue;
          uint32_t const im = __shfl(myi[m], src, 64);
          double const * const row = a.mats[m] + (uint64_t) im * R;
          #pragma unroll
          for(int q=0; q < NC; ++q) {
            uint32_t const col = c + q * L;
            if(live && col < R) p[k][q] *= row[col];
          }
        }
      }
      #pragma unroll
      for(int k=0; k < CO_BATCH; ++k) {
        if(k0 + k >= cnt) break;
        if(!a.reduce) {
          if(orow[k] != cur) {
            if(cur != 0xFFFFFFFFu) {
              double * const o = a.out + (uint64_t) cur * R;
              #pragma unroll
              for(int q=0; q < NC; ++q) {
                uint32_t const col = c + q * L;
                if(col < R) unsafeAtomicAdd(o + col, acc[q]);
              }
            }
            #pragma unroll
            for(int q=0; q < NC; ++q) acc[q] = 0.;
            cur = orow[k];
          }
        }
        #pragma unroll
        for(int q=0; q < NC; ++q) acc[q] += p[k][q];
      }
    }
    if(!a.reduce && cur != 0xFFFFFFFFu) {
      double * const o = a.out + (uint64_t) cur * R;
      #pragma unroll
      for(int q=0; q < NC; ++q) {
        uint32_t const col = c + q * L;
        if(col < R) unsafeAtomicAdd(o + col, acc[q]);
        acc[q] = 0.;
      }
      cur = 0xFFFFFFFFu;
    }
  }
  if(a.reduce) {
    #pragma unroll
    for(int q=0; q < NC; ++q) s_red[wave * G + g][c + q * L] = acc[q];
    __syncthreads();
    for(uint32_t col=threadIdx.x; col < R; col += CO_THREADS) {
      double sum = 0.;
      for(int w=0; w < W * G; ++w) sum += s_red[w][col];
      a.part[(uint64_t) blk * R + col] = sum;
    }
  }
}

template<int L, int NC>
__global__ void __launch_bounds__(CO_THREADS)
k_coo_mttkrp(CooArgs a)
{
  b_coo_mttkrp<L, NC>(a, blockIdx.x, gridDim.x);
}

// out[c] = sum_b part[b * R + c]: one block per 32 columns, the partials
// split over 32 thread groups and added in a fixed order.
__device__ __forceinline__ void
b_rowsum(double const * __restrict__ part, uint32_t nb, uint32_t R,
    double * __restrict__ out,
    uint32_t blk, uint32_t nblk)
{
  __shared__ double s[32][33];
  int const t = threadIdx.x;
  uint32_t const c = blk * 32 + t % 32;
  int const q = t / 32;
  double v = 0.;
  if(c < R) {
    for(uint32_t b=q; b < nb; b += 32) v += part[(uint64_t) b * R + c];
  }
  s[q][t % 32] = v;
  __syncthreads();
  if(t < 32 && c < R) {
    double tot = 0.;
    for(int k=0; k < 32; ++k) tot += s[k][t];
    out[c] = tot;
  }
}

__global__ void __launch_bounds__(1024)
k_rowsum(double const * __restrict__ part, uint32_t nb, uint32_t R,
    double * __restrict__ out)
{
  b_rowsum(part, nb, R, out, blockIdx.x, gridDim.x);
}

// Time Gram: symmetric copy of the upper triangle into old, then
// A = old + t t^T (StreamCPD.cc:420-435). One block.
__device__ __forceinline__ void
b_time_gram(double * __restrict__ A, double const * __restrict__ t,
    double * __restrict__ old, uint32_t R)
{
  for(uint32_t e=threadIdx.x; e < R * R; e += blockDim.x) {
    uint32_t const i = e / R;
    uint32_t const j = e - i * R;
    old[e] = (i > j) ? A[j * R + i] : A[e];
  }
  __syncthreads();
  for(uint32_t e=threadIdx.x; e < R * R; e += blockDim.x) {
    uint32_t const i = e / R;
    uint32_t const j = e - i * R;
    A[e] = old[e] + t[i] * t[j];
  }
}

__global__ void
k_time_gram(double * __restrict__ A, double const * __restrict__ t,
    double * __restrict__ old, uint32_t R)
{
  b_time_gram(A, t, old, R);
}

struct HistSrc { double const * m[SPD_MAX_MODES]; };

// H (RP x RP, zero padded) = old_gram o cross_0 o cross_1 ...
__device__ __forceinline__ void
b_hist(double const * __restrict__ old, HistSrc cr, int ncross,
    uint32_t R, uint32_t RP, double * __restrict__ H)
{
  for(uint32_t e=threadIdx.x; e < RP * RP; e += blockDim.x) {
    uint32_t const i = e / RP;
    uint32_t const j = e - i * RP;
    double v = 0.;
    if(i < R && j < R) {
      v = old[i * R + j];
      for(int k=0; k < ncross; ++k) v *= cr.m[k][i * R + j];
    }
    H[e] = v;
  }
}

__global__ void
k_hist(double const * __restrict__ old, HistSrc cr, int ncross,
    uint32_t R, uint32_t RP, double * __restrict__ H)
{
  b_hist(old, cr, ncross, R, RP, H);
}

__global__ void k_scale(double * __restrict__ x, uint64_t n, double f)
{
  uint64_t const e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if(e < n) x[e] *= f;
}

// Block partials of sum (a-b)^2 and sum b^2.
__device__ __forceinline__ void
b_diff2(double const * __restrict__ a, double const * __restrict__ b, uint64_t n,
    double * __restrict__ part,
    uint32_t blk, uint32_t nblk)
{
  __shared__ double s[2][256];
  uint64_t e = (uint64_t) blk * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) nblk * blockDim.x;
  double d2 = 0., b2 = 0.;
  for(; e < n; e += stride) {
    double const x = a[e] - b[e];
    d2 += x * x;
    b2 += b[e] * b[e];
  }
  int const t = threadIdx.x;
  s[0][t] = d2;
  s[1][t] = b2;
  __syncthreads();
  for(int h=128; h > 0; h >>= 1) {
    if(t < h) {
      s[0][t] += s[0][t + h];
      s[1][t] += s[1][t + h];
    }
    __syncthreads();
  }
  if(t < 2) part[(uint64_t) blk * 2 + t] = s[t][0];
}

__global__ void __launch_bounds__(256)
k_diff2(double const * __restrict__ a, double const * __restrict__ b, uint64_t n,
    double * __restrict__ part)
{
  b_diff2(a, b, n, part, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void
b_sum2(double const * __restrict__ part, uint32_t nb, double * __restrict__ out)
{
  __shared__ double s[128][2];
  int const t = threadIdx.x;
  int const k = t % 2;
  int const q = t / 2;
  double v = 0.;
  for(uint32_t b=q; b < nb; b += 128) v += part[(uint64_t) b * 2 + k];
  s[q][k] = v;
  __syncthreads();
  if(t < 2) {
    double tot = 0.;
    for(int j=0; j < 128; ++j) tot += s[j][t];
    out[t] = tot;
  }
}

__global__ void __launch_bounds__(256)
k_sum2(double const * __restrict__ part, uint32_t nb, double * __restrict__ out)
{
  b_sum2(part, nb, out);
}

__global__ void k_fill_f64(double * __restrict__ x, uint64_t n, double v)
{
  uint64_t const e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if(e < n) x[e] = v;
}

// Sort keys (one mode's labels) and positions for a stable re-sort.
__global__ void k_key_pos(uint32_t const * __restrict__ ind, uint64_t n,
    uint64_t * __restrict__ key, uint32_t * __restrict__ pos)
{
  uint64_t x = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; x < n; x += stride) {
    key[x] = ind[x];
    pos[x] = (uint32_t) x;
  }
}

struct SliceCopy { uint32_t const * in[SPD_MAX_MODES]; uint32_t * out[SPD_MAX_MODES]; };

// out[m][x] = in[m][perm[x]] for the modes given, vals likewise.
__global__ void k_gather_slice(SliceCopy c, int N, double const * __restrict__ vals,
    double * __restrict__ ovals, uint32_t const * __restrict__ perm, uint64_t n)
{
  uint64_t x = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; x < n; x += stride) {
    uint32_t const p = perm[x];
    #pragma unroll
    for(int m=0; m < SPD_MAX_MODES; ++m) {
      if(m < N && c.out[m] != nullptr) c.out[m][x] = c.in[m][p];
    }
    ovals[x] = vals[p];
  }
}

namespace {

template<int L, int NC>
int launch_coo(CooArgs const & a, uint32_t nb, hipStream_t st)
{
  hipLaunchKernelGGL((k_coo_mttkrp<L, NC>), dim3(nb), dim3(CO_THREADS), 0, st, a);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// COO MTTKRP of nnz nonzeros (ind[m][x], vals[x]) into out (rows x R,
// overwritten). With rows == 1 the output-mode indices are not read.
int coo_mttkrp(uint32_t const * const * ind, double const * vals,
    uint64_t nnz, int N, int mode, double const * const * mats, uint32_t R,
    uint64_t rows, double * out, hipStream_t st)
{
  CooArgs a = {};
  for(int m=0; m < N; ++m) {
    a.ind[m] = ind[m];
    a.mats[m] = mats[m];
  }
  a.vals = vals;
  a.nnz = nnz;
  a.R = R;
  a.N = N;
  a.mode = mode;
  a.reduce = (rows == 1);
  int L = 64, NC = 1;
  if(R <= 16) { L = 16; NC = 1; }
  else if(R <= 32) { L = 32; NC = 1; }
  else if(R <= 64) { L = 64; NC = 1; }
  else if(R <= 128) { L = 64; NC = 2; }
  else if(R <= 256) { L = 64; NC = 4; }
  else {
    spd_set_error(SPD_EINPUT, "streaming MTTKRP supports rank <= 256 (got %u)", R);
    return SPD_EINPUT;
  }
  uint64_t const per_block = (uint64_t) (CO_THREADS / L) * CO_CHUNK;
  uint64_t nb64 = spd_cdiv(nnz > 0 ? nnz : 1, per_block);
  uint32_t const cap = a.reduce ? 2048u : 8192u;
  uint32_t const nb = (uint32_t) (nb64 < cap ? nb64 : cap);
  void * part = nullptr;
  if(a.reduce) {
    SPD_TRY(spd_scratch(13, (size_t) nb * R * sizeof(double), &part));
    a.part = (double *) part;
  } else {
    SPD_TRY(spd_zero_async(out, rows * R * sizeof(double), st));
    a.out = out;
  }
  if(nnz > 0 || a.reduce) {
    if(L == 16) SPD_TRY((launch_coo<16, 1>(a, nb, st)));
    else if(L == 32) SPD_TRY((launch_coo<32, 1>(a, nb, st)));
    else if(NC == 1) SPD_TRY((launch_coo<64, 1>(a, nb, st)));
    else if(NC == 2) SPD_TRY((launch_coo<64, 2>(a, nb, st)));
    else SPD_TRY((launch_coo<64, 4>(a, nb, st)));
  }
  if(a.reduce) {
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned) spd_cdiv(R, 32)), dim3(1024), 0, st,
        (double const *) part, nb, R, out);
    SPD_LAUNCH_CHECK();
  }
  return SPD_OK;
}

// sum (a-b)^2 and sum b^2 over n entries -> out[0..1] (device)
int diff2(double const * a, double const * b, uint64_t n, double * out, hipStream_t st)
{
  uint64_t nb64 = spd_cdiv(n > 0 ? n : 1, 256 * 8);
  uint32_t const nb = (uint32_t) (nb64 < 1024 ? nb64 : 1024);
  void * part = nullptr;
  SPD_TRY(spd_scratch(14, (size_t) nb * 2 * sizeof(double), &part));
  hipLaunchKernelGGL(k_diff2, dim3(nb), dim3(256), 0, st, a, b, n, (double *) part);
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum2, dim3(1), dim3(256), 0, st, (double const *) part, nb, out);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

}  // namespace

int sst_coo_mttkrp(uint32_t const * const * ind, double const * vals, uint64_t nnz,
    int N, int mode, double const * const * mats, uint32_t R, uint64_t rows, double * out,
    hipStream_t st)
{
  return coo_mttkrp(ind, vals, nnz, N, mode, mats, R, rows, out, st);
}


double sst_wall_s()
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

static int st_h2d(void * dst, void const * src, size_t bytes, hipStream_t st)
{
  if(bytes == 0) return SPD_OK;
  SPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  SPD_HIP(hipStreamSynchronize(st));
  return SPD_OK;
}

static int st_d2h(void * dst, void const * src, size_t bytes, hipStream_t st)
{
  if(bytes == 0) return SPD_OK;
  SPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  SPD_HIP(hipStreamSynchronize(st));
  return SPD_OK;
}

static int st_d2d(void * dst, void const * src, size_t bytes, hipStream_t st)
{
  if(bytes == 0) return SPD_OK;
  SPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
  return SPD_OK;
}

size_t StreamState::footprint(int N, int t, uint64_t R, uint64_t const * cap)
{
  uint64_t maxdim = 1;
  size_t b = 0;
  for(int m=0; m < N; ++m) {
    b += R * R * sizeof(double);
    if(m == t) continue;
    b += 3 * cap[m] * R * sizeof(double);
    if(cap[m] > maxdim) maxdim = cap[m];
  }
  uint64_t const RP = dn_rpad((uint32_t) R);
  b += 3 * maxdim * R * sizeof(double);
  b += (4 + SPD_MAX_MODES) * R * R * sizeof(double) + 2 * RP * RP * sizeof(double);
  return b;
}

void StreamState::set_stream(hipStream_t stream)
{
  st = stream;
  if(mem.st) mem.st = stream;
  if(sort_mem.st) sort_mem.st = stream;
}

int StreamState::init(int nmodes, spd_stream_cfg const & c, uint64_t const * caps,
    hipStream_t stream, int ordered)
{
  if(ordered) {
    mem.st = stream;
    sort_mem.st = stream;
  }
  N = nmodes;
  cfg = c;
  t = c.stream_mode;
  R = (uint32_t) c.rank;
  RP = dn_rpad(R);
  st = stream;
  nslices = 0;
  zero_cap = 0;
  d_zero_ind = nullptr;
  for(int m=0; m < N; ++m) {
    cap[m] = (m == t) ? 1 : caps[m];
    rows[m] = 0;
    prev_dim[m] = 0;
  }
  uint64_t maxdim = 1;
  for(int m=0; m < N; ++m) if(m != t && cap[m] > maxdim) maxdim = cap[m];
  for(int m=0; m < N; ++m) {
    Unew[m] = Uold[m] = Dual[m] = nullptr;
    SPD_TRY(mem.get((void **) &aTa[m], (size_t) R * R * sizeof(double)));
    SPD_TRY(spd_zero_async(aTa[m], (size_t) R * R * sizeof(double), st));
    if(m == t) continue;
    size_t const bytes = cap[m] * R * sizeof(double);
    SPD_TRY(mem.get((void **) &Unew[m], bytes));
    SPD_TRY(mem.get((void **) &Uold[m], bytes));
    SPD_TRY(mem.get((void **) &Dual[m], bytes));
    SPD_TRY(spd_zero_async(Unew[m], bytes, st));
    SPD_TRY(spd_zero_async(Uold[m], bytes, st));
    SPD_TRY(spd_zero_async(Dual[m], bytes, st));
  }
  SPD_TRY(mem.get((void **) &time_row, R * sizeof(double)));
  SPD_TRY(spd_zero_async(time_row, R * sizeof(double), st));
  SPD_TRY(mem.get((void **) &Mbuf, maxdim * R * sizeof(double)));
  SPD_TRY(mem.get((void **) &aux, maxdim * R * sizeof(double)));
  SPD_TRY(mem.get((void **) &init_buf, maxdim * R * sizeof(double)));
  SPD_TRY(mem.get((void **) &old_gram, (size_t) R * R * sizeof(double)));
  SPD_TRY(spd_zero_async(old_gram, (size_t) R * R * sizeof(double), st));
  SPD_TRY(mem.get((void **) &ginv, (size_t) RP * RP * sizeof(double)));
  SPD_TRY(mem.get((void **) &H, (size_t) RP * RP * sizeof(double)));
  SPD_TRY(mem.get((void **) &cross, (size_t) SPD_MAX_MODES * R * R * sizeof(double)));
  SPD_TRY(mem.get((void **) &tgram, (size_t) R * R * sizeof(double)));
  SPD_TRY(mem.get((void **) &sc, (16 + 4 + R + 2 * SPD_MAX_MODES) * sizeof(double)));
  SPD_TRY(mem.get((void **) &info, sizeof(int)));
  SPD_TRY(spd_zero_async(info, sizeof(int), st));
  SPD_TRY(mem.get((void **) &ctl, 2 * sizeof(int)));
  SPD_TRY(mem.get((void **) &lam1, R * sizeof(double)));
  SPD_TRY(mem.get((void **) &tmpv, R * sizeof(double)));
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, lam1, (uint64_t) R, 1.0);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int StreamState::slice(uint32_t const * const * ind_in, double const * vals_b, uint64_t bn,
    double xnorm, double * local_err, uint64_t * outer_out, double * secs_out)
{
  SPD_HIP(hipStreamSynchronize(st));
  double const t0 = sst_wall_s();
  double * const d_rho = sc + 0;
  double * const d_inner = sc + 4;
  // change sums of every mode of an outer iteration, read back together
  double * const d_diffm = sc + 16 + 4 + R;
  double * const d_sums = sc + 8;
  // the streamed mode of a slice is all zeros
  if(bn > zero_cap) {
    uint64_t nc = zero_cap ? zero_cap : 1024;
    while(nc < bn) nc *= 2;
    SPD_TRY(mem.get((void **) &d_zero_ind, nc * sizeof(uint32_t)));
    SPD_TRY(spd_zero_async(d_zero_ind, nc * sizeof(uint32_t), st));
    zero_cap = nc;
  }
  uint32_t const * ind_b[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) ind_b[m] = (m == t) ? d_zero_ind : ind_in[m];
  // slice dims (ParserSimple.cc:131-140): running maxima of the labels
  uint64_t bd[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) {
    if(m == t) { bd[m] = 1; continue; }
    bd[m] = (batch_max[m] + 1 > prev_dim[m]) ? batch_max[m] + 1 : prev_dim[m];
    if(bd[m] > cap[m]) {
      spd_set_error(SPD_EINPUT, "stream: mode %d needs %llu rows, capacity %llu", m,
          (unsigned long long) bd[m], (unsigned long long) cap[m]);
      return SPD_EINPUT;
    }
    prev_dim[m] = bd[m];
  }
  double const * aTa_c[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) aTa_c[m] = aTa[m];

  // grow_mats (StreamCPD.cc:192-220): new rows from rand(), Grams
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    if(bd[m] > rows[m]) {
      uint64_t const nnew = (bd[m] - rows[m]) * R;
      hrand.resize(nnew);
      cfg.fill_rand(cfg.rand_ctx, hrand.data(), nnew);
      SPD_TRY(st_h2d(Unew[m] + rows[m] * R, hrand.data(), nnew * sizeof(double), st));
      rows[m] = bd[m];
    }
    SPD_TRY(dn_gram(Unew[m], rows[m], R, nullptr, nullptr, aTa[m], st));
  }
  if(nslices == 0) {
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      SPD_TRY(dn_colnorm(Unew[m], rows[m], R, tmpv, st));
      SPD_TRY(dn_div_cols(Unew[m], rows[m], R, tmpv, nullptr, st));
    }
  }
  double const * mats[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) mats[m] = (m == t) ? time_row : Unew[m];

  // copies of the slice sorted by each mode after the first one following
  // the streamed mode, so that their MTTKRP outputs come in runs of rows
  // and add up before one atomic per run
  int const lead = (t + 1) % N;
  if(bn > sort_cap) {
    uint64_t nc = sort_cap ? sort_cap : 4096;
    while(nc < bn) nc *= 2;
    SpdMem fresh;
    fresh.st = sort_mem.st;
    if(!fresh.st) SPD_HIP(hipStreamSynchronize(st));
    for(int k=0; k < 2; ++k) {
      SPD_TRY(fresh.get((void **) &sk[k], nc * sizeof(uint64_t)));
      SPD_TRY(fresh.get((void **) &sv[k], nc * sizeof(uint32_t)));
    }
    for(int m=0; m < N; ++m) {
      by_vals[m] = nullptr;
      for(int k=0; k < N; ++k) by_ind[m][k] = nullptr;
      if(m == t || m == lead) continue;
      SPD_TRY(fresh.get((void **) &by_vals[m], nc * sizeof(double)));
      for(int k=0; k < N; ++k) {
        if(k != t) SPD_TRY(fresh.get((void **) &by_ind[m][k], nc * sizeof(uint32_t)));
      }
    }
    sort_mem.ptrs.swap(fresh.ptrs);
    sort_cap = nc;
  }
  uint32_t const * ind_m[SPD_MAX_MODES][SPD_MAX_MODES];
  double const * vals_m[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) {
    vals_m[m] = vals_b;
    for(int k=0; k < N; ++k) ind_m[m][k] = ind_b[k];
    if(m == t || m == lead || bn < 2) continue;
    unsigned const grid = (unsigned) (spd_cdiv(bn, 256) < 4096 ? spd_cdiv(bn, 256) : 4096);
    hipLaunchKernelGGL(k_key_pos, dim3(grid), dim3(256), 0, st, ind_b[m], bn, sk[0], sv[0]);
    SPD_LAUNCH_CHECK();
    int bits = 1;
    while(bits < 32 && (1ull << bits) < bd[m]) ++bits;
    int which = 0;
    SPD_TRY(spd_radix_sort(sk[0], sv[0], sk[1], sv[1], bn, bits, &which, st));
    SliceCopy cp = {};
    for(int k=0; k < N; ++k) {
      if(k == t) continue;
      cp.in[k] = ind_b[k];
      cp.out[k] = by_ind[m][k];
      ind_m[m][k] = by_ind[m][k];
    }
    hipLaunchKernelGGL(k_gather_slice, dim3(grid), dim3(256), 0, st, cp, N, vals_b,
        by_vals[m], (uint32_t const *) sv[which], bn);
    SPD_LAUNCH_CHECK();
    vals_m[m] = by_vals[m];
  }

  double prev_delta = 0.;
  uint64_t outer = 0;
  for(outer=0; outer < cfg.max_its; ++outer) {
    double delta = 0.;
    // time row: MTTKRP to one row and the closed-form solve
    SPD_TRY(coo_mttkrp(ind_b, vals_b, bn, N, t, mats, R, 1, Mbuf, st));
    SPD_TRY(dn_chol_inv(aTa_c, N, t, R, cfg.con[t].diag_add, 0, ginv, d_rho, info, st));
    SPD_TRY(dn_apply(Mbuf, SPD_F64, ginv, 1, R, time_row, nullptr, st));
    hipLaunchKernelGGL(k_time_gram, dim3(1), dim3(256), 0, st, aTa[t],
        (double const *) time_row, old_gram, R);
    SPD_LAUNCH_CHECK();
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      uint64_t const I = bd[m];
      SPD_TRY(coo_mttkrp(ind_m[m], vals_m[m], bn, N, m, mats, R, I, Mbuf, st));
      // add_historical (StreamCPD.cc:223-294)
      HistSrc hs = {};
      int nc = 0;
      for(int mm=0; mm < N; ++mm) {
        if(mm == m || mm == t) continue;
        double * cm = cross + (size_t) nc * R * R;
        SPD_TRY(dn_xgram(Uold[mm], Unew[mm], rows[mm], R, cm, st));
        hs.m[nc++] = cm;
      }
      hipLaunchKernelGGL(k_hist, dim3(1), dim3(256), 0, st, (double const *) old_gram,
          hs, nc, R, RP, H);
      SPD_LAUNCH_CHECK();
      SPD_TRY(dn_apply_acc(Uold[m], H, I, R, Mbuf, st));
      spd_con const & con = cfg.con[m];
      if(con.kind == SPD_CON_ADMM) {
        // the duals restart from zero for every mode's solve (the solve
        // takes that as given instead of reading a cleared matrix)
        SPD_TRY(dn_chol_inv(aTa_c, N, m, R, con.diag_add, 1, ginv, d_rho, info, st));
        SPD_TRY(dn_admm_solve(Mbuf, SPD_F64, ginv, d_rho, I, R, Unew[m], Dual[m], aux,
            init_buf, con.prox, con.lam, cfg.max_inner, cfg.inner_tol, d_sums, ctl, st, 1));
      } else {
        SPD_TRY(dn_chol_inv(aTa_c, N, m, R, con.diag_add, 0, ginv, d_rho, info, st));
        SPD_TRY(dn_apply(Mbuf, SPD_F64, ginv, I, R, Unew[m], nullptr, st));
      }
      SPD_TRY(dn_gram(Unew[m], I, R, nullptr, nullptr, aTa[m], st));
      SPD_TRY(diff2(Uold[m], Unew[m], I * R, d_diffm + 2 * m, st));
    }
    // one readback per outer iteration; the modes' terms added in mode order
    double dd[2 * SPD_MAX_MODES];
    SPD_TRY(st_d2h(dd, d_diffm, 2 * (size_t) N * sizeof(double), st));
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      delta += sqrt(dd[2 * m]) / sqrt(dd[2 * m + 1]);
    }
    if(cfg.verbosity > 1) {
      printf("  delta: %e prev_delta: %e (%e diff)\n", delta, prev_delta,
          fabs(delta - prev_delta));
    }
    if(outer > 0 && fabs(delta - prev_delta) < cfg.tol) {
      if(cfg.verbosity > 1) {
        printf("  converged in: %llu\n", (unsigned long long) (outer + 1));
      }
      ++outer;
      break;
    }
    prev_delta = delta;
  }
  // forget, old <- new (StreamCPD.cc:488-506)
  hipLaunchKernelGGL(k_scale, dim3((unsigned) spd_cdiv((uint64_t) R * R, 256)), dim3(256),
      0, st, aTa[t], (uint64_t) R * R, cfg.forget);
  SPD_LAUNCH_CHECK();
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    SPD_TRY(st_d2d(Uold[m], Unew[m], rows[m] * R * sizeof(double), st));
  }
  SPD_HIP(hipStreamSynchronize(st));
  if(secs_out) *secs_out = sst_wall_s() - t0;
  if(outer_out) *outer_out = outer;
  ++nslices;

  // local error of this slice (StreamCPD.cc:77-89): the model of the last
  // time row against the slice (time index 0)
  if(local_err) {
    int small = 0;
    for(int m=1; m < N; ++m) if(bd[m] < bd[small]) small = m;
    uint64_t const nr = bd[small];
    SPD_TRY(coo_mttkrp(ind_b, vals_b, bn, N, small, mats, R, nr, Mbuf, st));
    double const * fsm = (small == t) ? time_row : Unew[small];
    SPD_TRY(dn_innerprod(fsm, Mbuf, SPD_F64, nr, R, lam1, d_inner, st));
    SPD_TRY(dn_gram(time_row, 1, R, nullptr, nullptr, tgram, st));
    double const * grams[SPD_MAX_MODES];
    for(int m=0; m < N; ++m) grams[m] = (m == t) ? tgram : aTa[m];
    SPD_TRY(dn_cpd_norm(grams, N, R, lam1, d_inner + 1, st));
    double h[2];
    SPD_TRY(st_d2h(h, d_inner, sizeof(h), st));
    *local_err = (xnorm + h[1] - 2. * h[0]) / xnorm;
  }
  return SPD_OK;
}

int StreamState::cholesky_info(int * out)
{
  return st_d2h(out, info, sizeof(int), st);
}

extern "C" int spd_stream_cpd(int N, uint64_t nnz, uint64_t const * dims,
    uint32_t const * const * inds, double const * vals, spd_stream_cfg const * cfg_in,
    spd_stream_result * res)
{
  SPD_TRY(spd_init());
  hipStream_t const st = spd_hs(nullptr);
  spd_stream_cfg const cfg = *cfg_in;
  if(cfg.fill_rand == nullptr) {
    spd_set_error(SPD_EINPUT, "stream: no random fill for new factor rows");
    return SPD_EINPUT;
  }
  int const t = cfg.stream_mode;
  uint64_t const R64 = cfg.rank;
  if(N < 2 || N > SPD_MAX_MODES || t < 0 || t >= N || R64 == 0 || R64 > 256 ||
      nnz == 0 || nnz >= (1ull << 32)) {
    spd_set_error(SPD_EINPUT, "stream: bad shape (nmodes %d, stream mode %d, rank %llu,"
        " nnz %llu)", N, t, (unsigned long long) R64, (unsigned long long) nnz);
    return SPD_EINPUT;
  }
  SPD_TRY(sst_check_cfg(N, cfg));
  uint32_t const R = (uint32_t) R64;

  // ---- batches (ParserSimple::next_batch, ParserSimple.cc:89-149) ----
  std::vector<uint64_t> bstart, bend;
  std::vector<std::vector<uint64_t>> bmax;
  {
    uint64_t ptr = 0;
    uint64_t it = 0;
    while(ptr < nnz) {
      uint64_t s = ptr;
      while(s < nnz && inds[t][s] < it) ++s;
      uint64_t e = s;
      while(e < nnz && inds[t][e] < it + 1) ++e;
      if(e == s) {
        fprintf(stderr, "SPLATT: time slice %llu is empty; the stream ends there.\n",
            (unsigned long long) it);
        break;
      }
      std::vector<uint64_t> mx(N, 0);
      for(int m=0; m < N; ++m) {
        if(m == t) continue;
        for(uint64_t x=s; x < e; ++x) if(inds[m][x] > mx[m]) mx[m] = inds[m][x];
      }
      bstart.push_back(s);
      bend.push_back(e);
      bmax.push_back(mx);
      ptr = e;
      ++it;
    }
  }
  uint64_t const T = bstart.size();
  uint64_t caps[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) {
    caps[m] = 1;
    for(uint64_t b=0; b < T; ++b) if(bmax[b][m] + 1 > caps[m]) caps[m] = bmax[b][m] + 1;
  }

  SpdMem mem;
  uint32_t * d_ind[SPD_MAX_MODES] = {};
  double * d_vals = nullptr;
  for(int m=0; m < N; ++m) {
    SPD_TRY(mem.get((void **) &d_ind[m], nnz * sizeof(uint32_t)));
    SPD_TRY(spd_h2d(d_ind[m], inds[m], nnz * sizeof(uint32_t)));
  }
  SPD_TRY(mem.get((void **) &d_vals, nnz * sizeof(double)));
  SPD_TRY(spd_h2d(d_vals, vals, nnz * sizeof(double)));
  double * gtime = nullptr;
  SPD_TRY(mem.get((void **) &gtime, (T ? T : 1) * R * sizeof(double)));
  SPD_TRY(spd_zero_async(gtime, (T ? T : 1) * R * sizeof(double), st));

  StreamState s;
  SPD_TRY(s.init(N, cfg, caps, st));

  // cpd_error (src/cpd/cpd.c:580-656) of slices [b0, b1) against the factors
  // (time rows gtime[b0 .. b1), the other modes' current rows); wd are the
  // dims of that tensor, whose smallest mode takes the MTTKRP. Squared.
  double * M2 = nullptr;
  {
    uint64_t mr = T;
    for(int m=0; m < N; ++m) if(dims[m] > mr) mr = dims[m];
    SPD_TRY(mem.get((void **) &M2, (mr ? mr : 1) * R * sizeof(double)));
  }
  double * const d_inner = s.sc + 4;
  auto window_errsq = [&](uint64_t b0, uint64_t b1, uint64_t const * wd,
      double * out) -> int {
    int small = 0;
    for(int m=1; m < N; ++m) if(wd[m] < wd[small]) small = m;
    uint64_t const x0 = bstart[b0];
    uint64_t const x1 = bend[b1 - 1];
    uint64_t const nt = b1 - b0;
    double const * fm[SPD_MAX_MODES];
    for(int m=0; m < N; ++m) fm[m] = (m == t) ? gtime : s.Unew[m];
    uint64_t frows = 0;
    if(small == t) {
      // the stream is sorted by time: one single-row reduction per slice
      frows = nt;
      SPD_TRY(spd_zero_async(M2, nt * R * sizeof(double), st));
      for(uint64_t b=b0; b < b1; ++b) {
        uint32_t const * ib[SPD_MAX_MODES];
        for(int m=0; m < N; ++m) ib[m] = d_ind[m] + bstart[b];
        SPD_TRY(coo_mttkrp(ib, d_vals + bstart[b], bend[b] - bstart[b], N, t, fm, R, 1,
            M2 + (b - b0) * R, st));
      }
    } else {
      frows = s.rows[small];
      uint32_t const * ib[SPD_MAX_MODES];
      for(int m=0; m < N; ++m) ib[m] = d_ind[m] + x0;
      SPD_TRY(coo_mttkrp(ib, d_vals + x0, x1 - x0, N, small, fm, R, wd[small], M2, st));
    }
    double const * fsm = (small == t) ? gtime + b0 * R : s.Unew[small];
    SPD_TRY(dn_innerprod(fsm, M2, SPD_F64, frows, R, s.lam1, d_inner, st));
    SPD_TRY(dn_gram(gtime + b0 * R, nt, R, nullptr, nullptr, s.tgram, st));
    double const * grams[SPD_MAX_MODES];
    for(int m=0; m < N; ++m) grams[m] = (m == t) ? s.tgram : s.aTa[m];
    SPD_TRY(dn_cpd_norm(grams, N, R, s.lam1, d_inner + 1, st));
    double h[2];
    SPD_TRY(spd_d2h(h, d_inner, sizeof(h)));
    double xn = 0.;
    for(uint64_t x=x0; x < x1; ++x) xn += vals[x] * vals[x];
    double const e = sqrt(xn + h[1] - 2. * h[0]) / sqrt(xn);
    *out = e * e;
    return SPD_OK;
  };

  for(uint64_t b=0; b < T; ++b) {
    uint32_t const * ib[SPD_MAX_MODES];
    for(int m=0; m < N; ++m) {
      ib[m] = d_ind[m] + bstart[b];
      s.batch_max[m] = bmax[b][m];
    }
    double xn = 0.;
    for(uint64_t x=bstart[b]; x < bend[b]; ++x) xn += vals[x] * vals[x];
    double le = 0., secs = 0.;
    uint64_t outer = 0;
    SPD_TRY(s.slice(ib, d_vals + bstart[b], bend[b] - bstart[b], xn,
        res->local_err ? &le : nullptr, &outer, &secs));
    SPD_TRY(spd_d2d(gtime + b * R, s.time_row, R * sizeof(double)));
    if(res->local_err) res->local_err[b] = le;
    if(res->batch_seconds) res->batch_seconds[b] = secs;
    if(res->outer_its) res->outer_its[b] = outer;

    // checkpoints (StreamCPD.cc:517-528)
    uint64_t const done = b + 1;
    double ge = -1., l10 = -1., ce = -1.;
    if(cfg.check_interval > 0 && (done % cfg.check_interval == 0 || bend[b] == nnz)) {
      uint64_t wd[SPD_MAX_MODES];
      for(int m=0; m < N; ++m) wd[m] = (m == t) ? done : s.rows[m];
      SPD_TRY(window_errsq(0, done, wd, &ge));
      uint64_t const b0 = (10 < done) ? done - 10 : 0;
      wd[t] = done - b0;
      SPD_TRY(window_errsq(b0, done, wd, &l10));
      if(cfg.refit) {
        wd[t] = done;
        ce = cfg.refit(cfg.refit_ctx, N, bend[b], wd, inds, vals, R);
        if(isnan(ce)) ce = -1.;
      }
    }
    if(res->global_err) res->global_err[b] = ge;
    if(res->local10_err) res->local10_err[b] = l10;
    if(res->cpd_err) res->cpd_err[b] = ce;
    if(cfg.verbosity > 0) {
      printf("batch %5llu: %7llu nnz (%0.5fs) (%0.3e NNZ/s) cpd: %+0.5f global: "
          "%+0.5f local-1: %+0.5f local-10: %+0.5f\n", (unsigned long long) done,
          (unsigned long long) (bend[b] - bstart[b]), secs,
          (double) (bend[b] - bstart[b]) / secs, ce, ge, le, l10);
      fflush(stdout);
    }
  }
  int hinfo = 0;
  SPD_TRY(s.cholesky_info(&hinfo));
  if(hinfo != 0) {
    fprintf(stderr, "SPLATT: Cholesky factorisation failed in the stream (pivot %d).\n",
        hinfo);
  }

  // final error of the whole stream (StreamCPD.cc:549-556)
  res->nbatches = T;
  for(int m=0; m < N; ++m) res->dims[m] = (m == t) ? T : s.rows[m];
  if(res->final_err && T > 0) {
    SPD_TRY(window_errsq(0, T, dims, res->final_err));
  }
  for(int m=0; m < N; ++m) {
    if(res->factors[m] == nullptr) continue;
    if(m == t) {
      SPD_TRY(spd_d2h(res->factors[m], gtime, T * R * sizeof(double)));
    } else {
      SPD_TRY(spd_d2h(res->factors[m], s.Unew[m], s.rows[m] * R * sizeof(double)));
    }
  }
  return SPD_OK;
}

int sst_check_cfg(int N, spd_stream_cfg const & cfg)
{
  int const t = cfg.stream_mode;
  for(int m=0; m < N; ++m) {
    if(m != t && cfg.con[m].kind == SPD_CON_NONE) {
      spd_set_error(SPD_EINPUT, "stream: mode %d needs a constraint", m);
      return SPD_EINPUT;
    }
  }
  if(cfg.con[t].kind == SPD_CON_ADMM) {
    spd_set_error(SPD_EINPUT, "stream: the streamed mode takes closed-form "
        "constraints only");
    return SPD_EINPUT;
  }
  return SPD_OK;
}

// ---- batched launches (spd_batch.hpp): every slice of a batch at once ----

#define SB_ALL 0
#define SB_LIVE 1
#define SB_SMALL 2        // slices whose local error uses this mode

__device__ __forceinline__ bool sb_gate(BSlot const & s, int gate, int mode)
{
  if(gate == SB_SMALL) return s.small == mode;
  return gate == SB_ALL || s.c->done == 0;
}

__host__ __device__ __forceinline__ uint32_t sb_coo_nb(uint64_t nnz, uint32_t L, int reduce)
{
  uint64_t const per_block = (uint64_t) (CO_THREADS / L) * CO_CHUNK;
  uint64_t const nb = ((nnz > 0 ? nnz : 1) + per_block - 1) / per_block;
  uint64_t const cap = reduce ? 2048u : 8192u;
  return (uint32_t) (nb < cap ? nb : cap);
}

// COO MTTKRP of output mode `mode` over copy `copy` of each slice's
// nonzeros: the streamed mode reduces to one row (block partials in part),
// other modes add into Mbuf (cleared before)
template<int L, int NC>
__global__ void __launch_bounds__(CO_THREADS)
kb_coo(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int copy,
    int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  int const red = (mode == hd->t);
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(sb_gate(s, gate, mode)) nb = sb_coo_nb(s.nnz, L, red);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    CooArgs a = {};
    for(int m=0; m < hd->N; ++m) {
      a.ind[m] = s.ind[copy][m];
      a.mats[m] = (m == hd->t) ? s.time_row : s.Unew[m];
    }
    a.vals = s.vals[copy];
    a.nnz = s.nnz;
    a.R = s.R;
    a.N = hd->N;
    a.mode = mode;
    a.reduce = red;
    a.out = s.Mbuf;
    a.part = s.part;
    b_coo_mttkrp<L, NC>(a, w - s_it.pref[q], s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

// the streamed mode's row (the only mode reduced to one row): Mbuf[0, :] =
// sum of the block partials
__global__ void __launch_bounds__(1024)
kb_rowsum(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int L, int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(sb_gate(s, gate, hd->t)) nb = (s.R + 31) / 32;
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    b_rowsum(s.part, sb_coo_nb(s.nnz, (uint32_t) L, 1), s.R, s.Mbuf, w - s_it.pref[q], 0);
    __syncthreads();
  }
}

// what 0: Mbuf rows of `mode`; 1: Dual rows of `mode`
__global__ void kb_zero(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode,
    int what, int gate)
{
  uint32_t const q = blockIdx.y;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(!sb_gate(s, gate, mode)) return;
  double * const x = (what == 0) ? s.Mbuf : s.Dual[mode];
  uint64_t const n = s.rows[mode] * s.R;
  for(uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; e < n;
      e += (uint64_t) gridDim.x * blockDim.x) {
    x[e] = 0.;
  }
}

__global__ void kb_time_gram(BSlot const * __restrict__ S, BHead const * __restrict__ hd)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(s.c->done) return;
  b_time_gram(s.aTa[hd->t], s.time_row, s.old_gram, s.R);
}

// H of `mode` from the cross Grams of the other non-streamed modes
__global__ void kb_hist(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int ncross)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(s.c->done) return;
  HistSrc hs = {};
  for(int k=0; k < ncross; ++k) hs.m[k] = s.cross + (uint64_t) k * s.R * s.R;
  b_hist(s.old_gram, hs, ncross, s.R, dn_rpad(s.R), s.H);
}

__host__ __device__ __forceinline__ uint32_t sb_diff_nb(uint64_t n)
{
  uint64_t const nb = ((n > 0 ? n : 1) + 256 * 8 - 1) / (256 * 8);
  return (uint32_t) (nb < 1024 ? nb : 1024);
}

__global__ void __launch_bounds__(256)
kb_diff2(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(s.c->done == 0) nb = sb_diff_nb(s.rows[mode] * s.R);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    b_diff2(s.Uold[mode], s.Unew[mode], s.rows[mode] * s.R, s.part, w - s_it.pref[q],
        s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

// the mode's change measure added to the outer iteration's delta, in the
// per-slice order: delta += sqrt(sum (old-new)^2) / sqrt(sum new^2)
__global__ void __launch_bounds__(256)
kb_sum2_delta(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(s.c->done) return;
  b_sum2(s.part, sb_diff_nb(s.rows[mode] * s.R), s.c->diff);
  __syncthreads();
  if(threadIdx.x == 0) s.c->delta += sqrt(s.c->diff[0]) / sqrt(s.c->diff[1]);
}

__global__ void kb_ctl_reset(BSlot const * __restrict__ S, BHead const * __restrict__ hd)
{
  uint32_t const q = threadIdx.x;
  if(q >= hd->K) return;
  BCtl * const c = S[q].c;
  c->admm[0] = 0;
  c->admm[1] = 0;
}

// start of a slice's step: the outer loop runs, nothing measured yet
__global__ void kb_begin(BSlot const * __restrict__ S, BHead const * __restrict__ hd)
{
  uint32_t const q = threadIdx.x;
  if(q >= hd->K) return;
  BCtl * const c = S[q].c;
  c->done = 0;
  c->outer = 0;
  c->delta = 0.;
  c->prev_delta = 0.;
  c->admm[0] = 0;
  c->admm[1] = 0;
  c->info = 0;
}

// end of outer iteration o (StreamCPD.cc:476-486): stop when o > 0 and the
// change measure moved less than the tolerance
__global__ void kb_outer_check(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int o)
{
  uint32_t const q = threadIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  BCtl * const c = s.c;
  if(c->done) return;
  double const delta = c->delta;
  c->outer = o + 1;
  if(o > 0 && fabs(delta - c->prev_delta) < s.tol) {
    c->done = 1;
  } else {
    c->prev_delta = delta;
  }
  if((uint32_t) (o + 1) >= s.max_its) c->done = 1;
  c->delta = 0.;
}

// after the outer loop: the time Gram forgets, old <- new
__global__ void kb_finish(BSlot const * __restrict__ S, BHead const * __restrict__ hd)
{
  uint32_t const q = blockIdx.y;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  uint32_t const R = s.R;
  if(blockIdx.x == 0) {
    double * const g = s.aTa[hd->t];
    for(uint32_t e=threadIdx.x; e < R * R; e += blockDim.x) g[e] *= s.forget;
  }
  for(int m=0; m < hd->N; ++m) {
    if(m == hd->t) continue;
    uint64_t const n = s.rows[m] * R;
    for(uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; e < n;
        e += (uint64_t) gridDim.x * blockDim.x) {
      s.Uold[m][e] = s.Unew[m][e];
    }
  }
}

// new factor rows (values drawn on the host) into Unew
__global__ void kb_grow(BSlot const * __restrict__ S, BHead const * __restrict__ hd)
{
  uint32_t const q = blockIdx.y;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  for(int m=0; m < hd->N; ++m) {
    if(m == hd->t || s.grow[m] == nullptr) continue;
    uint64_t const n = (s.rows[m] - s.grow0[m]) * s.R;
    double * const dst = s.Unew[m] + s.grow0[m] * s.R;
    for(uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; e < n;
        e += (uint64_t) gridDim.x * blockDim.x) {
      dst[e] = s.grow[m][e];
    }
  }
}

// per slice results for the host: time row (R), outer its, <X,Z>, ||Z||^2,
// Cholesky status
__global__ void kb_collect(BSlot const * __restrict__ S, BHead const * __restrict__ hd,
    double * __restrict__ out, uint32_t ld)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  double * const o = out + (uint64_t) q * ld;
  for(uint32_t r=threadIdx.x; r < s.R; r += blockDim.x) o[r] = s.time_row[r];
  if(threadIdx.x == 0) {
    o[s.R] = (double) s.c->outer;
    o[s.R + 1] = s.c->inner[0];
    o[s.R + 2] = s.c->inner[1];
    o[s.R + 3] = (double) s.c->info;
  }
}

namespace {
uint32_t sb_grid(uint64_t tot)
{
  if(tot < 1) tot = 1;
  return (uint32_t) (tot < BT_GRID ? tot : BT_GRID);
}
}

int sstb_coo(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, int copy, int gate, int t, hipStream_t st)
{
  int const L = (TT == 1) ? 16 : ((TT == 2) ? 32 : 64);
  int const red = (mode == t);
  uint64_t tot = 0;
  uint64_t maxn = 1;
  for(uint32_t q=0; q < K; ++q) {
    tot += sb_coo_nb(hS[q].nnz, (uint32_t) L, red);
    uint64_t const n = hS[q].rows[mode] * hS[q].R;
    if(n > maxn) maxn = n;
  }
  if(!red) {
    uint64_t gx = (maxn + 255) / 256;
    if(gx > 512) gx = 512;
    hipLaunchKernelGGL(kb_zero, dim3((unsigned) gx, K), dim3(256), 0, st, dS, dH, mode, 0, gate);
    SPD_LAUNCH_CHECK();
  }
  if(L == 16) hipLaunchKernelGGL((kb_coo<16, 1>), dim3(sb_grid(tot)), dim3(CO_THREADS), 0, st, dS, dH, mode, copy, gate);
  else if(L == 32) hipLaunchKernelGGL((kb_coo<32, 1>), dim3(sb_grid(tot)), dim3(CO_THREADS), 0, st, dS, dH, mode, copy, gate);
  else hipLaunchKernelGGL((kb_coo<64, 1>), dim3(sb_grid(tot)), dim3(CO_THREADS), 0, st, dS, dH, mode, copy, gate);
  SPD_LAUNCH_CHECK();
  if(red) {
    hipLaunchKernelGGL(kb_rowsum, dim3(sb_grid((uint64_t) K * 2)), dim3(1024), 0, st, dS, dH, L,
        gate);
    SPD_LAUNCH_CHECK();
  }
  return SPD_OK;
}

int sstb_zero_dual(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st)
{
  uint64_t maxn = 1;
  for(uint32_t q=0; q < K; ++q) {
    uint64_t const n = hS[q].rows[mode] * hS[q].R;
    if(n > maxn) maxn = n;
  }
  uint64_t gx = (maxn + 255) / 256;
  if(gx > 512) gx = 512;
  hipLaunchKernelGGL(kb_zero, dim3((unsigned) gx, K), dim3(256), 0, st, dS, dH, mode, 1, SB_LIVE);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_time_gram(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st)
{
  hipLaunchKernelGGL(kb_time_gram, dim3(K), dim3(256), 0, st, dS, dH);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_hist(uint32_t K, BSlot const * dS, BHead const * dH, int ncross, hipStream_t st)
{
  hipLaunchKernelGGL(kb_hist, dim3(K), dim3(256), 0, st, dS, dH, ncross);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_delta(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st)
{
  uint64_t tot = 0;
  for(uint32_t q=0; q < K; ++q) tot += sb_diff_nb(hS[q].rows[mode] * hS[q].R);
  hipLaunchKernelGGL(kb_diff2, dim3(sb_grid(tot)), dim3(256), 0, st, dS, dH, mode);
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL(kb_sum2_delta, dim3(K), dim3(256), 0, st, dS, dH, mode);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_ctl_reset(BSlot const * dS, BHead const * dH, hipStream_t st)
{
  hipLaunchKernelGGL(kb_ctl_reset, dim3(1), dim3(64), 0, st, dS, dH);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_begin(BSlot const * dS, BHead const * dH, hipStream_t st)
{
  hipLaunchKernelGGL(kb_begin, dim3(1), dim3(64), 0, st, dS, dH);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_outer_check(BSlot const * dS, BHead const * dH, int o, hipStream_t st)
{
  hipLaunchKernelGGL(kb_outer_check, dim3(1), dim3(64), 0, st, dS, dH, o);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_finish(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st)
{
  hipLaunchKernelGGL(kb_finish, dim3(256, K), dim3(256), 0, st, dS, dH);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_grow(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st)
{
  hipLaunchKernelGGL(kb_grow, dim3(64, K), dim3(256), 0, st, dS, dH);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int sstb_collect(uint32_t K, BSlot const * dS, BHead const * dH, double * out, uint32_t ld,
    hipStream_t st)
{
  hipLaunchKernelGGL(kb_collect, dim3(K), dim3(64), 0, st, dS, dH, out, ld);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}
