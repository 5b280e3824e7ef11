// spd_dense.hip -- dense FP64 factor algebra for CPD on gfx950.
//
// The reference does these steps with BLAS/LAPACK (src/matrix.c: syrk for
// A^T A, potrf + potrs for the normal equations) and OpenMP loops (column
// norms, ADMM updates, fit). Here:
//   - Grams A^T A run on MFMA (v_mfma_f64_16x16x4f64): a wave holds the
//     16x16 output tiles of the upper triangle in registers and streams
//     4-row slices of A through them; blocks write partial Grams that a
//     second kernel sums in a fixed order (results are deterministic).
//   - The R x R normal-equation matrix is formed, Cholesky-factored and
//     inverted in one single-wave kernel (R^3/3 flops, latency-bound).
//   - The solve of every row against it (potrs in the reference) becomes a
//     product with the inverse on MFMA, with the rows of the right-hand
//     side staged through LDS; the column sums of squares needed for
//     normalisation or the max-column-norm prox are fused into it.
//   - Column reductions use fixed-order block partials, so repeated runs
//     give bit-identical results.
// Ranks up to 64 take the MFMA kernels; larger ranks take plain kernels.
#include "spd_dense.hpp"
#include "spd_batch.hpp"

typedef double d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16). Lane l holds A[l%16][l/16] and B[l/16][l%16];
// D comes back as D[l/16 + 4*r][l%16] in component r.
__device__ __forceinline__ d4 mfma16(double a, double b, d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

#define DN_THREADS 256
#define DN_MAX_BLOCKS 1024
#define GR_U 4                 // 4-row MFMA steps per wave batch

static uint32_t dn_blocks(uint64_t work_items, uint64_t per_block)
{
  uint64_t nb = spd_cdiv(work_items, per_block);
  if(nb < 1) nb = 1;
  if(nb > DN_MAX_BLOCKS) nb = DN_MAX_BLOCKS;
  return (uint32_t) nb;
}

// Sums of block partials, part[b * ld + src], b < nb, for up to CW entries
// at once by one block of NT threads: thread t takes entry t % CW and every
// G-th partial starting at t / CW (G = NT / CW), four loads in flight; the
// G partial sums are then added in a fixed order. Returns the total in
// threads t < CW (src of the entry is given by those threads).
// Deterministic for fixed nb, CW and NT.
template<int CW, int NT = DN_THREADS>
__device__ __forceinline__ double block_partsum(double const * __restrict__ part,
    uint32_t nb, uint32_t ld, bool valid, uint32_t src, double (*s)[CW])
{
  constexpr int G = NT / CW;
  int const t = threadIdx.x;
  int const e = t % CW;
  int const q = t / CW;
  double v = 0.;
  if(q < G && valid) {
    uint32_t b = q;
    for(; b + 3 * G < nb; b += 4 * G) {
      double const x0 = part[(uint64_t) b * ld + src];
      double const x1 = part[(uint64_t) (b + G) * ld + src];
      double const x2 = part[(uint64_t) (b + 2 * G) * ld + src];
      double const x3 = part[(uint64_t) (b + 3 * G) * ld + src];
      v += ((x0 + x1) + (x2 + x3));
    }
    for(; b < nb; b += G) v += part[(uint64_t) b * ld + src];
  }
  if(q < G) s[q][e] = v;
  __syncthreads();
  double tot = 0.;
  if(t < CW) {
    for(int k=0; k < G; ++k) tot += s[k][t];
  }
  __syncthreads();
  return tot;
}

// ---- Gram on MFMA (R <= 64) -------------------------------------------------

template<int TT, bool DIV, bool CROSS>
__device__ __forceinline__ void
b_gram_mfma(double * __restrict__ X, uint64_t I, uint32_t R,
    double const * __restrict__ div, float * __restrict__ X32,
    double * __restrict__ part, double const * __restrict__ Y,
    uint32_t blk, uint32_t nblk)
{
  constexpr int RP = TT * 16;
  __shared__ double s_acc[RP * RP];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const kk = lane >> 4;
  int const cc = lane & 15;
  for(int e=threadIdx.x; e < RP * RP; e += DN_THREADS) s_acc[e] = 0.;

  d4 acc[TT][TT];
  #pragma unroll
  for(int i=0; i < TT; ++i) {
    #pragma unroll
    for(int j=0; j < TT; ++j) acc[i][j] = (d4) 0.;
  }
  bool cok[TT];
  double dv[TT];
  #pragma unroll
  for(int t=0; t < TT; ++t) {
    uint32_t const col = t * 16 + cc;
    cok[t] = col < R;
    dv[t] = (DIV && cok[t]) ? div[col] : 1.;
  }
  uint64_t const nsteps = (I + 3) / 4;
  uint64_t const stride = (uint64_t) nblk * 4 * GR_U;
  for(uint64_t s0 = ((uint64_t) blk * 4 + wave) * GR_U; s0 < nsteps; s0 += stride) {
    double a[GR_U][TT];
    double b[CROSS ? GR_U : 1][TT];
    #pragma unroll
    for(int u=0; u < GR_U; ++u) {
      uint64_t const row = (s0 + u) * 4 + kk;
      #pragma unroll
      for(int t=0; t < TT; ++t) {
        bool const ok = (row < I) && cok[t];
        a[u][t] = ok ? X[row * R + t * 16 + cc] : 0.;
        if(CROSS) b[CROSS ? u : 0][t] = ok ? Y[row * R + t * 16 + cc] : 0.;
      }
    }
    if(DIV || X32 != nullptr) {
      #pragma unroll
      for(int u=0; u < GR_U; ++u) {
        uint64_t const row = (s0 + u) * 4 + kk;
        #pragma unroll
        for(int t=0; t < TT; ++t) {
          if((row < I) && cok[t]) {
            uint64_t const idx = row * R + t * 16 + cc;
            double v = a[u][t];
            if(DIV) {
              v = v / dv[t];
              X[idx] = v;
              a[u][t] = v;
            }
            if(X32 != nullptr) X32[idx] = (float) v;
          }
        }
      }
    }
    #pragma unroll
    for(int u=0; u < GR_U; ++u) {
      #pragma unroll
      for(int i=0; i < TT; ++i) {
        #pragma unroll
        for(int j=(CROSS ? 0 : i); j < TT; ++j) {
          acc[i][j] = mfma16(a[u][i], CROSS ? b[CROSS ? u : 0][j] : a[u][j], acc[i][j]);
        }
      }
    }
  }
  // waves add their tiles in a fixed order
  __syncthreads();
  for(int w=0; w < 4; ++w) {
    if(wave == w) {
      #pragma unroll
      for(int i=0; i < TT; ++i) {
        #pragma unroll
        for(int j=(CROSS ? 0 : i); j < TT; ++j) {
          #pragma unroll
          for(int r=0; r < 4; ++r) {
            s_acc[(i * 16 + kk + 4 * r) * RP + j * 16 + cc] += acc[i][j][r];
          }
        }
      }
    }
    __syncthreads();
  }
  double * const out = part + (uint64_t) blk * RP * RP;
  for(int e=threadIdx.x; e < RP * RP; e += DN_THREADS) out[e] = s_acc[e];
}

template<int TT, bool DIV, bool CROSS>
__global__ void __launch_bounds__(DN_THREADS)
k_gram_mfma(double * __restrict__ X, uint64_t I, uint32_t R,
    double const * __restrict__ div, float * __restrict__ X32,
    double * __restrict__ part, double const * __restrict__ Y)
{
  b_gram_mfma<TT, DIV, CROSS>(X, I, R, div, X32, part, Y, blockIdx.x, gridDim.x);
}

// Sum of the block partials into the full R x R: the upper triangle
// mirrored (sym) or every entry as computed (cross products). 32 entries
// per block.
#define GR_RED_THREADS 1024
__device__ __forceinline__ void
b_gram_reduce(double const * __restrict__ part, uint32_t nb,
    uint32_t R, uint32_t ld, double * __restrict__ out, int sym,
    uint32_t blk, uint32_t nblk)
{
  __shared__ double s[GR_RED_THREADS / 32][32];
  uint32_t const e = blk * 32 + threadIdx.x % 32;
  bool const valid = e < R * R;
  uint32_t const i = valid ? e / R : 0;
  uint32_t const j = valid ? e - i * R : 0;
  uint32_t const a = (sym && j < i) ? j : i;
  uint32_t const b = (sym && j < i) ? i : j;
  double const tot = block_partsum<32, GR_RED_THREADS>(part, nb, ld * ld, valid, a * ld + b, s);
  if(threadIdx.x < 32 && valid) out[e] = tot;
}

__global__ void __launch_bounds__(GR_RED_THREADS)
k_gram_reduce(double const * __restrict__ part, uint32_t nb,
    uint32_t R, uint32_t ld, double * __restrict__ out, int sym)
{
  b_gram_reduce(part, nb, R, ld, out, sym, blockIdx.x, gridDim.x);
}

// ---- plain kernels for any rank -----------------------------------------------

// Gram block partials for ranks above 64: entry (i, j) of X^T Y over the
// block's rows (Y == X: upper triangle only).
__global__ void __launch_bounds__(DN_THREADS)
k_gram_plain(double const * __restrict__ X, double const * __restrict__ Y,
    uint64_t I, uint32_t R, uint64_t rows_per_block, double * __restrict__ part)
{
  uint64_t const r0 = (uint64_t) blockIdx.x * rows_per_block;
  uint64_t const r1 = (r0 + rows_per_block < I) ? r0 + rows_per_block : I;
  double * const out = part + (uint64_t) blockIdx.x * R * R;
  bool const sym = (X == Y);
  for(uint32_t e=threadIdx.x; e < R * R; e += DN_THREADS) {
    uint32_t const i = e / R;
    uint32_t const j = e - i * R;
    if(sym && i > j) continue;
    double s = 0.;
    for(uint64_t r=r0; r < r1; ++r) s += X[r * R + i] * Y[r * R + j];
    out[e] = s;
  }
}

// X[:,r] /= div[r], FP32 copy.
__device__ __forceinline__ void
b_div_cols(double * __restrict__ X, uint64_t n, uint32_t R,
    double const * __restrict__ div, float * __restrict__ X32,
    uint32_t blk, uint32_t nblk)
{
  uint64_t e = (uint64_t) blk * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) nblk * blockDim.x;
  uint32_t c = (uint32_t) (e % R);
  uint32_t const sc = (uint32_t) (stride % R);
  for(; e < n; e += stride) {
    double v = X[e];
    if(div != nullptr) {
      v = v / div[c];
      X[e] = v;
    }
    if(X32 != nullptr) X32[e] = (float) v;
    c += sc;
    if(c >= R) c -= R;
  }
}

__global__ void
k_div_cols(double * __restrict__ X, uint64_t n, uint32_t R,
    double const * __restrict__ div, float * __restrict__ X32)
{
  b_div_cols(X, n, R, div, X32, blockIdx.x, gridDim.x);
}

// Per block and column: sum over the block's rows of f(X[i,c]) * Y[i,c]
// (Y == nullptr: f(X)^2, f = max(x,0) when CLAMP). Block partials in
// part[b * R + c].
template<typename TY, bool CLAMP>
__device__ __forceinline__ void
b_coldot(double const * __restrict__ X, TY const * __restrict__ Y, uint64_t I,
    uint32_t R, uint64_t rows_per_block, double * __restrict__ part,
    uint32_t blk, uint32_t nblk)
{
  __shared__ double s[DN_THREADS];
  uint64_t const r0 = (uint64_t) blk * rows_per_block;
  uint64_t const r1 = (r0 + rows_per_block < I) ? r0 + rows_per_block : I;
  int const t = threadIdx.x;
  if(R <= DN_THREADS) {
    uint32_t const nq = DN_THREADS / R;
    uint32_t const c = t % R;
    uint32_t const q = t / R;
    double acc = 0.;
    if(q < nq) {
      for(uint64_t i=r0 + q; i < r1; i += nq) {
        double x = X[i * R + c];
        if(CLAMP) x = (x < 0.) ? 0. : x;
        double const y = (Y == nullptr) ? x : (double) Y[i * R + c];
        acc += x * y;
      }
    }
    s[t] = acc;
    __syncthreads();
    if(t < (int) R) {
      double tot = 0.;
      for(uint32_t k=0; k < nq; ++k) tot += s[t + k * R];
      part[(uint64_t) blk * R + t] = tot;
    }
  } else {
    for(uint32_t c=t; c < R; c += DN_THREADS) {
      double acc = 0.;
      for(uint64_t i=r0; i < r1; ++i) {
        double x = X[i * R + c];
        if(CLAMP) x = (x < 0.) ? 0. : x;
        double const y = (Y == nullptr) ? x : (double) Y[i * R + c];
        acc += x * y;
      }
      part[(uint64_t) blk * R + c] = acc;
    }
  }
}

template<typename TY, bool CLAMP>
__global__ void __launch_bounds__(DN_THREADS)
k_coldot(double const * __restrict__ X, TY const * __restrict__ Y, uint64_t I,
    uint32_t R, uint64_t rows_per_block, double * __restrict__ part)
{
  b_coldot<TY, CLAMP>(X, Y, I, R, rows_per_block, part, blockIdx.x, gridDim.x);
}

// Column totals of block partials (row stride ld): mode 0 writes
// out[c] = sqrt(total_c); mode 1 writes out[0] = sum_c total_c * lambda[c]
// (columns taken in order). One block.
__device__ __forceinline__ void
b_colreduce(double const * __restrict__ part, uint32_t nb, uint32_t R, uint32_t ld,
    int mode, double const * __restrict__ lambda, double * __restrict__ out)
{
  __shared__ double s[DN_THREADS / 32][32];
  double mine = 0.;
  for(uint32_t c0=0; c0 < R; c0 += 32) {
    uint32_t const c = c0 + threadIdx.x % 32;
    double const v = block_partsum<32>(part, nb, ld, c < R, c, s);
    if(threadIdx.x < 32 && c < R) {
      if(mode == 0) out[c] = sqrt(v);
      else mine += v * lambda[c];
    }
  }
  if(mode == 1) {
    s[0][threadIdx.x % 32] = 0.;
    __syncthreads();
    if(threadIdx.x < 32) s[0][threadIdx.x] = mine;
    __syncthreads();
    if(threadIdx.x == 0) {
      double tot = 0.;
      for(int k=0; k < 32; ++k) tot += s[0][k];
      out[0] = tot;
    }
  }
}

__global__ void __launch_bounds__(DN_THREADS)
k_colreduce(double const * __restrict__ part, uint32_t nb, uint32_t R, uint32_t ld,
    int mode, double const * __restrict__ lambda, double * __restrict__ out)
{
  b_colreduce(part, nb, R, ld, mode, lambda, out);
}

// ---- normal equations: form, Cholesky, inverse -----------------------------------

struct GramSrc { double const * aTa[SPD_MAX_MODES]; };

// One wave. W is 2 R^2 doubles (LDS when it fits, else global scratch).
__global__ void __launch_bounds__(64)
k_chol_inv(GramSrc g, int nmodes, int mode, uint32_t R, uint32_t RP,
    double diag_add, int admm, double * gW, double * __restrict__ ginv,
    double * __restrict__ rho_out, int * __restrict__ info)
{
  extern __shared__ double dyn[];
  double * const A = gW ? gW : dyn;       // R x R working matrix, then L
  double * const Li = A + (uint64_t) R * R;   // R x R, L^{-1}
  int const t = threadIdx.x;

  for(uint32_t e=t; e < R * R; e += 64) {
    uint32_t const i = e / R;
    uint32_t const j = e - i * R;
    uint32_t const a = i < j ? i : j;
    uint32_t const b = i < j ? j : i;
    double h = 1.;
    for(int m=0; m < nmodes; ++m) {
      if(m != mode) h *= g.aTa[m][a * R + b];
    }
    A[e] = h;
    Li[e] = 0.;
  }
  __syncthreads();
  if(t == 0) {
    for(uint32_t i=0; i < R; ++i) A[i * R + i] += diag_add;
    if(admm) {
      double tr = 0.;
      for(uint32_t i=0; i < R; ++i) tr += A[i * R + i];
      double const rho = tr / (double) R;
      for(uint32_t i=0; i < R; ++i) A[i * R + i] += rho;
      *rho_out = rho;
    }
  }
  __syncthreads();

  // right-looking Cholesky on the lower triangle; thread t owns rows t+64k.
  // The first bad pivot is kept in a register and stored once at the end.
  int bad = 0;
  for(uint32_t j=0; j < R; ++j) {
    double const d = A[j * R + j];
    if(t == 0 && !(d > 0.) && bad == 0) bad = (int) j + 1;
    double const ljj = sqrt(d);
    __syncthreads();
    for(uint32_t i=j + 1 + t; i < R; i += 64) A[i * R + j] /= ljj;
    if(t == 0) A[j * R + j] = ljj;
    __syncthreads();
    for(uint32_t i=j + 1 + t; i < R; i += 64) {
      double const lij = A[i * R + j];
      for(uint32_t k=j + 1; k <= i; ++k) A[i * R + k] -= lij * A[k * R + j];
    }
    __syncthreads();
  }
  if(t == 0 && bad != 0 && *info == 0) *info = bad;
  // L^{-1}, column c by thread c (forward substitution against e_c)
  for(uint32_t c=t; c < R; c += 64) {
    Li[c * R + c] = 1. / A[c * R + c];
    for(uint32_t i=c + 1; i < R; ++i) {
      double s = 0.;
      for(uint32_t k=c; k < i; ++k) s += A[i * R + k] * Li[k * R + c];
      Li[i * R + c] = -s / A[i * R + i];
    }
  }
  __syncthreads();
  // G^{-1} = L^{-T} L^{-1}; padded entries are zero
  for(uint32_t e=t; e < RP * RP; e += 64) {
    uint32_t const i = e / RP;
    uint32_t const j = e - i * RP;
    if(i >= R || j >= R) ginv[e] = 0.;
  }
  for(uint32_t j=t; j < R; j += 64) {
    for(uint32_t i=0; i <= j; ++i) {
      double s = 0.;
      for(uint32_t k=j; k < R; ++k) s += Li[k * R + i] * Li[k * R + j];
      ginv[i * RP + j] = s;
      ginv[j * RP + i] = s;
    }
  }
}

// The same for R <= RP <= 64 by Gauss-Jordan elimination in place (no
// pivoting: G is symmetric positive definite, and its pivots are the
// squares of the Cholesky diagonal, so a non-positive one is reported the
// same way). Every step updates all RP x RP entries spread over the NT
// threads, so the work per thread is RP^3/NT independent LDS updates rather
// than a chain of dependent ones. The result is written from its upper
// triangle, so it is exactly symmetric.
template<int RP, int NT>
__device__ __forceinline__ void
b_gj_inv(GramSrc g, int nmodes, int mode, uint32_t R, double diag_add, int admm,
    double * __restrict__ ginv, double * __restrict__ rho_out, int * __restrict__ info)
{
  __shared__ double A[RP][RP + 1];
  __shared__ double colj[RP];
  __shared__ double s_rho;
  constexpr int NE = RP * RP / NT;
  static_assert(NE * NT == RP * RP && RP <= NT, "entries must split evenly");
  int const t = threadIdx.x;
  int const iR = (int) R;
  // the Hadamard product into A, one Gram at a time in chunks of 16 entries
  // per thread whose loads are all in flight together (the Grams were just
  // written by other kernels and come from far away)
  #pragma unroll
  for(int q=0; q < NE; ++q) {
    int const e = t + NT * q;
    int const r = e / RP;
    int const k = e - r * RP;
    A[r][k] = (r < iR && k < iR) ? 1. : ((k == r) ? 1. : 0.);
  }
  constexpr int CH = (NE < 16) ? NE : 16;
  for(int m=0; m < nmodes; ++m) {
    if(m == mode) continue;
    double const * const gm = g.aTa[m];
    #pragma unroll
    for(int q0=0; q0 < NE; q0 += CH) {
      // chunks past NE (when CH does not divide NE) are skipped at compile time
      double gv[CH];
      #pragma unroll
      for(int q=0; q < CH; ++q) {
        if(q0 + q < NE) {
          int const e = t + NT * (q0 + q);
          int const r = e / RP;
          int const k = e - r * RP;
          bool const ok = (r < iR && k < iR);
          int const lo = r < k ? r : k;
          int const hi = r < k ? k : r;
          gv[q] = gm[ok ? lo * R + hi : 0];
        }
      }
      #pragma unroll
      for(int q=0; q < CH; ++q) {
        if(q0 + q < NE) {
          int const e = t + NT * (q0 + q);
          int const r = e / RP;
          int const k = e - r * RP;
          A[r][k] *= (r < iR && k < iR) ? gv[q] : 1.;   // branch-free: x * 1 == x
        }
      }
    }
  }
  // the diagonal entry (t, t) belongs to another thread of the staging
  // (in another wave when NT > 64): wait for the Hadamard product first
  __syncthreads();
  if(t < iR) A[t][t] += diag_add;
  __syncthreads();
  if(admm) {
    if(t == 0) {
      double tr = 0.;
      for(int q=0; q < iR; ++q) tr += A[q][q];
      double const rho = tr / (double) R;
      *rho_out = rho;
      s_rho = rho;
    }
    __syncthreads();
    if(t < iR) A[t][t] += s_rho;
    __syncthreads();
  }
  // the first bad pivot is kept in a register and stored once at the end
  int bad = 0;
  // Step j: colj = column j (pivot entry zeroed); row j scaled by p = 1/d
  // with p in the pivot; column j cleared elsewhere. Then every entry takes
  // the same update A[i][k] -= colj[i] * A[j][k]: row j passes unchanged
  // (colj[j] = 0), column j becomes -colj[i] * p, the rest is eliminated.
  // The update runs in chunks whose LDS reads are all issued before use.
  constexpr int UC = (NE < 16) ? NE : 16;
  for(int j=0; j < RP; ++j) {
    double const d = A[j][j];
    if(t == 0 && j < iR && !(d > 0.) && bad == 0) bad = j + 1;
    double const p = 1. / d;
    double cj = 0.;
    if(t < RP) cj = A[t][j];
    __syncthreads();
    if(t < RP) {
      colj[t] = (t == j) ? 0. : cj;
      if(t == j) {
        A[j][j] = p;
      } else {
        A[j][t] = A[j][t] * p;
        A[t][j] = 0.;
      }
    }
    __syncthreads();
    // when RP divides NT a thread's entries share one column k = t % RP, so
    // the row-j factor is read once per step (row j itself passes unchanged)
    constexpr bool KF = (NT % RP) == 0;
    double const ajk0 = KF ? A[j][t % RP] : 0.;
    #pragma unroll
    for(int q0=0; q0 < NE; q0 += UC) {
      double ci[UC], ajk[UC], aik[UC];
      #pragma unroll
      for(int q=0; q < UC; ++q) {
        if(q0 + q < NE) {
          int const e = t + NT * (q0 + q);
          int const i = e / RP;
          int const k = e - i * RP;
          ci[q] = colj[i];
          ajk[q] = KF ? ajk0 : A[j][k];
          aik[q] = A[i][k];
        }
      }
      #pragma unroll
      for(int q=0; q < UC; ++q) {
        if(q0 + q < NE) {
          int const e = t + NT * (q0 + q);
          int const i = e / RP;
          int const k = e - i * RP;
          A[i][k] = aik[q] - ci[q] * ajk[q];
        }
      }
    }
    __syncthreads();
  }
  if(t == 0 && bad != 0 && *info == 0) *info = bad;
  for(int e=t; e < RP * RP; e += NT) {
    int const q = e / RP;
    int const i = e - q * RP;
    bool const in = (i < iR) && (q < iR);
    ginv[e] = in ? ((q <= i) ? A[q][i] : A[i][q]) : 0.;
  }
}

template<int RP, int NT>
__global__ void __launch_bounds__(NT)
k_gj_inv(GramSrc g, int nmodes, int mode, uint32_t R, double diag_add, int admm,
    double * __restrict__ ginv, double * __restrict__ rho_out, int * __restrict__ info)
{
  b_gj_inv<RP, NT>(g, nmodes, mode, R, diag_add, admm, ginv, rho_out, info);
}

// ---- apply the inverse on MFMA (R <= 64) ---------------------------------------

template<typename TM>
struct ApplyArgs {
  TM const * M;
  double const * ginv;      // RP x RP
  uint64_t I;
  uint32_t R;
  double * X;               // KIND 0: output; KIND 1: primal (in: old, out: aux - D)
  double * aux;             // KIND 1
  double const * dual;      // KIND 1
  double * init;            // KIND 1
  double const * rho;       // KIND 1
  int clamp;                // KIND 1: column sums of max(P, 0)^2
  double * colsq;           // per-block partial sums of squares (RP per block), or null
};

template<int TT, typename TM, int KIND>
__device__ __forceinline__ void
b_apply_mfma(ApplyArgs<TM> a,
    uint32_t blk, uint32_t nblk)
{
  constexpr int RP = TT * 16;
  constexpr int LD = RP + 1;
  constexpr int RPI = (64 / RP) > 0 ? 64 / RP : 1;   // rows staged per pass
  __shared__ double s_g[RP * RP];
  __shared__ double s_m[4][16 * LD];
  __shared__ double s_cs[RP];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const kk = lane >> 4;
  int const cc = lane & 15;
  uint64_t const I = a.I;
  uint32_t const R = a.R;
  for(int e=threadIdx.x; e < RP * RP; e += DN_THREADS) s_g[e] = a.ginv[e];
  for(int e=threadIdx.x; e < RP; e += DN_THREADS) s_cs[e] = 0.;
  __syncthreads();
  double const rho = (KIND == 1) ? *a.rho : 0.;
  double cs[TT];
  #pragma unroll
  for(int t=0; t < TT; ++t) cs[t] = 0.;

  int const lr = lane / RP;
  int const lc = lane - lr * RP;
  uint64_t const ngroups = (I + 15) / 16;
  for(uint64_t gi = (uint64_t) blk * 4 + wave; gi < ngroups; gi += (uint64_t) nblk * 4) {
    uint64_t const row0 = gi * 16;
    // stage 16 rows of the right-hand side (coalesced: the rows are contiguous)
    #pragma unroll
    for(int rr=0; rr < 16; rr += RPI) {
      int const r = rr + lr;
      if(lr < RPI && r < 16) {
        uint64_t const row = row0 + r;
        double v = 0.;
        if(row < I && (uint32_t) lc < R) {
          uint64_t const idx = row * R + lc;
          if(KIND != 1) {
            v = (double) a.M[idx];
          } else {
            double const p = a.X[idx];
            double const d = a.dual[idx];
            a.init[idx] = p;
            v = (double) a.M[idx] + rho * (p + d);
          }
        }
        s_m[wave][r * LD + lc] = v;
      }
    }
    __builtin_amdgcn_wave_barrier();
    d4 acc[TT];
    #pragma unroll
    for(int t=0; t < TT; ++t) acc[t] = (d4) 0.;
    #pragma unroll
    for(int k0=0; k0 < RP; k0 += 4) {
      double const av = s_m[wave][cc * LD + k0 + kk];
      #pragma unroll
      for(int t=0; t < TT; ++t) {
        acc[t] = mfma16(av, s_g[(k0 + kk) * RP + t * 16 + cc], acc[t]);
      }
    }
    __builtin_amdgcn_wave_barrier();
    #pragma unroll
    for(int t=0; t < TT; ++t) {
      uint32_t const col = t * 16 + cc;
      #pragma unroll
      for(int r=0; r < 4; ++r) {
        uint64_t const row = row0 + kk + 4 * r;
        if(row < I && col < R) {
          uint64_t const idx = row * R + col;
          double const x = acc[t][r];
          if(KIND == 0) {
            a.X[idx] = x;
            cs[t] += x * x;
          } else if(KIND == 2) {
            a.X[idx] += x;
          } else {
            a.aux[idx] = x;
            double const p = x - a.dual[idx];
            a.X[idx] = p;
            double const q = (a.clamp && p < 0.) ? 0. : p;
            cs[t] += q * q;
          }
        }
      }
    }
  }
  if(a.colsq != nullptr) {
    #pragma unroll
    for(int t=0; t < TT; ++t) {
      cs[t] += __shfl_xor(cs[t], 16, 64);
      cs[t] += __shfl_xor(cs[t], 32, 64);
    }
    for(int w=0; w < 4; ++w) {
      if(wave == w && kk == 0) {
        #pragma unroll
        for(int t=0; t < TT; ++t) s_cs[t * 16 + cc] += cs[t];
      }
      __syncthreads();
    }
    for(int e=threadIdx.x; e < RP; e += DN_THREADS) {
      a.colsq[(uint64_t) blk * RP + e] = s_cs[e];
    }
  }
}

template<int TT, typename TM, int KIND>
__global__ void __launch_bounds__(DN_THREADS)
k_apply_mfma(ApplyArgs<TM> a)
{
  b_apply_mfma<TT, TM, KIND>(a, blockIdx.x, gridDim.x);
}

// ---- ADMM with the iterate held as the pre-prox matrix (R <= 64) -----------
//
// The reference's inner iteration k (src/cpd/admm.c:240-308) computes
// aux_k = (M + rho (P_{k-1} + D_{k-1})) G^{-1}, P'_k = aux_k - D_{k-1},
// P_k = prox(P'_k) and D_k = D_{k-1} + P_k - aux_k = P_k - P'_k. So the
// state after iteration k is P'_k alone (plus the prox's column norms):
// P_k = prox(P'_k), D_k = P_k - P'_k. Pass j reads M and P'_{j-1} (the
// state it updates) and P'_{j-2} (for the convergence terms of iteration
// j-1, which need P_{j-2} and D_{j-2}), and writes P'_j: four matrix passes
// per iteration instead of the eight of a solve/prox/dual sweep pair. The
// first two passes take P_0 and D_0 as given. Iteration j-1's test is known
// after pass j; when it converges, pass j's output is discarded and P, D are
// rebuilt from P'_{j-1}.
template<typename TM>
struct Admm2Args {
  TM const * M;
  double const * ginv;      // RP x RP
  uint64_t I;
  uint32_t R;
  double const * rho;
  int prox;
  double lam;               // lasso multiplier
  double const * P0;        // explicit state (P_0, D_0)
  double const * D0;
  int d0zero;               // D_0 is zero: not read
  double const * cur;       // P'_{j-1} (states >= 1)
  double const * cur_n;     // its prox column norms
  double const * prev;      // P'_{j-2} (test against a state >= 1)
  double const * prev_n;
  double * out;             // P'_j
  double * part;            // RP column sums, then 4 test sums, each over the blocks
  int const * ctl;          // ctl[0] != 0: converged, nothing to do
};

// nc: the column's norm (max-col-norm variants) or the entry-wise scalar of
// dn_prox_scalar (src/cpd/constraints/lasso.c:30-52, nonneg_frob.c:32-55,
// nonneg_lasso.c:25-45)
__device__ __forceinline__ double dn_prox_apply(double p, int prox, double nc)
{
  if(prox == DN_PROX_NONNEG || dn_prox_scalar_kind(prox)) {
    p = dn_prox_entry(p, prox, nc);
  } else if(prox == DN_PROX_MAXCOLNORM || prox == DN_PROX_MAXCOLNORM_NONNEG) {
    if(prox == DN_PROX_MAXCOLNORM_NONNEG && p < 0.) p = 0.;
    if(!(nc <= 1.)) p /= nc;
  }
  return p;
}

// MODE 0: pass 1 (state P_0, D_0; no test). MODE 1: pass 2 (state P'_1,
// test of iteration 1 against P_0, D_0). MODE 2: pass j >= 3 (state
// P'_{j-1}, test against P'_{j-2}). Block partials: column sums of squares
// of P'_j and the four test terms.
// O32: every matrix of the pass is below 2 GiB, so element offsets are
// formed as 32-bit byte offsets from the (uniform) base pointers: one
// offset register per load instead of a 64-bit address per matrix.
template<int TT, typename TM, int MODE, bool O32 = false>
__device__ __forceinline__ void
b_admm_pass(Admm2Args<TM> a,
    uint32_t blk, uint32_t nblk)
{
  if(a.ctl[0] != 0) return;
  constexpr int RP = TT * 16;
  constexpr int LD = RP + 1;
  constexpr int RPI = (64 / RP) > 0 ? 64 / RP : 1;
  constexpr int NQ = (16 + RPI - 1) / RPI;     // staged entries per lane
  constexpr int E = RP + 4;                    // partials per block
  // s_w holds each wave's 16-row tile during the sweep, then the block's
  // partial sums, then the last block's wave totals (one after the other)
  constexpr int WS = 4 * 16 * LD;
  static_assert(WS >= 4 * DN_THREADS && WS >= 4 * E, "LDS work area too small");
  __shared__ double s_g[RP * RP];
  __shared__ double s_w[WS];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const kk = lane >> 4;
  int const cc = lane & 15;
  uint64_t const I = a.I;
  uint32_t const R = a.R;
  int const prox = a.prox;
  for(int e=threadIdx.x; e < RP * RP; e += DN_THREADS) s_g[e] = a.ginv[e];
  __syncthreads();
  double const rho = *a.rho;
  double cs[TT];
  #pragma unroll
  for(int t=0; t < TT; ++t) cs[t] = 0.;
  double pn = 0., pr = 0., dr = 0., dn = 0.;
  int const lr = lane / RP;
  int const lc = lane - lr * RP;
  bool const lok = (lr < RPI) && ((uint32_t) lc < R);
  double const thr = dn_prox_scalar(prox, a.lam, rho);
  bool const sk = dn_prox_scalar_kind(prox);
  double const ncur = sk ? thr : ((MODE >= 1 && lok) ? a.cur_n[lc] : 1.);
  double const nprev = sk ? thr : ((MODE == 2 && lok) ? a.prev_n[lc] : 1.);
  uint64_t const ngroups = (I + 15) / 16;
  for(uint64_t gi = (uint64_t) blk * 4 + wave; gi < ngroups; gi += (uint64_t) nblk * 4) {
    uint64_t const row0 = gi * 16;
    // stage M + rho (P + D) of the state j-1; D kept for the output; the
    // test terms of iteration j-1. Every load of the group is issued first.
    double mv[NQ], cv[NQ], xv[NQ], yv[NQ];
    #pragma unroll
    for(int q=0; q < NQ; ++q) {
      int const r = q * RPI + lr;
      uint64_t const row = row0 + r;
      bool const ok = lok && r < 16 && row < I;
      uint64_t const idx = ok ? row * R + lc : 0;
      if constexpr (O32) {
        uint32_t const i32 = (uint32_t) idx;
        uint32_t const bm = i32 * (uint32_t) sizeof(TM);
        uint32_t const bd = i32 * (uint32_t) sizeof(double);
        auto ld = [&](double const * base) -> double {
          return *(double const *) ((char const *) base + bd);
        };
        mv[q] = ok ? (double) *(TM const *) ((char const *) a.M + bm) : 0.;
        if(MODE == 0) {
          cv[q] = ok ? ld(a.P0) : 0.;
          xv[q] = (ok && !a.d0zero) ? ld(a.D0) : 0.;
        } else if(MODE == 1) {
          cv[q] = ok ? ld(a.cur) : 0.;
          xv[q] = ok ? ld(a.P0) : 0.;
          yv[q] = (ok && !a.d0zero) ? ld(a.D0) : 0.;
        } else {
          cv[q] = ok ? ld(a.cur) : 0.;
          xv[q] = ok ? ld(a.prev) : 0.;
        }
      } else {
        mv[q] = ok ? (double) a.M[idx] : 0.;
        if(MODE == 0) {
          cv[q] = ok ? a.P0[idx] : 0.;
          xv[q] = (ok && !a.d0zero) ? a.D0[idx] : 0.;
        } else if(MODE == 1) {
          cv[q] = ok ? a.cur[idx] : 0.;
          xv[q] = ok ? a.P0[idx] : 0.;
          yv[q] = (ok && !a.d0zero) ? a.D0[idx] : 0.;
        } else {
          cv[q] = ok ? a.cur[idx] : 0.;
          xv[q] = ok ? a.prev[idx] : 0.;
        }
      }
    }
    double dv[NQ];
    #pragma unroll
    for(int q=0; q < NQ; ++q) {
      int const r = q * RPI + lr;
      uint64_t const row = row0 + r;
      bool const ok = lok && r < 16 && row < I;
      double P, D;
      if(MODE == 0) {
        P = cv[q];
        D = xv[q];
      } else {
        P = dn_prox_apply(cv[q], prox, ncur);
        D = P - cv[q];
      }
      double const v = ok ? mv[q] + rho * (P + D) : 0.;
      if(MODE >= 1 && ok) {
        double Pp, Dp;
        if(MODE == 2) {
          Pp = dn_prox_apply(xv[q], prox, nprev);
          Dp = Pp - xv[q];
        } else {
          Pp = xv[q];
          Dp = yv[q];
        }
        double const x = D - Dp;
        double const w = P - Pp;
        pn += P * P;
        pr += x * x;
        dr += w * w;
        dn += D * D;
      }
      dv[q] = ok ? D : 0.;
      if(lr < RPI && r < 16) s_w[wave * 16 * LD + r * LD + lc] = v;
    }
    __builtin_amdgcn_wave_barrier();
    d4 acc[TT];
    #pragma unroll
    for(int t=0; t < TT; ++t) acc[t] = (d4) 0.;
    #pragma unroll
    for(int k0=0; k0 < RP; k0 += 4) {
      double const av = s_w[wave * 16 * LD + cc * LD + k0 + kk];
      #pragma unroll
      for(int t=0; t < TT; ++t) {
        acc[t] = mfma16(av, s_g[(k0 + kk) * RP + t * 16 + cc], acc[t]);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // D of the state, in the layout of the product
    #pragma unroll
    for(int q=0; q < NQ; ++q) {
      int const r = q * RPI + lr;
      if(lr < RPI && r < 16) s_w[wave * 16 * LD + r * LD + lc] = dv[q];
    }
    __builtin_amdgcn_wave_barrier();
    // P'_j = aux_j - D_{j-1}; column sums of what the prox will measure
    #pragma unroll
    for(int t=0; t < TT; ++t) {
      uint32_t const col = t * 16 + cc;
      #pragma unroll
      for(int r=0; r < 4; ++r) {
        int const rl = kk + 4 * r;
        uint64_t const row = row0 + rl;
        if(row < I && col < R) {
          double const p = acc[t][r] - s_w[wave * 16 * LD + rl * LD + col];
          if constexpr (O32) {
            *(double *) ((char *) a.out + (uint32_t) (row * R + col) * (uint32_t) sizeof(double)) = p;
          } else {
            a.out[row * R + col] = p;
          }
          double const q = (prox == DN_PROX_MAXCOLNORM_NONNEG && p < 0.) ? 0. : p;
          cs[t] += q * q;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // block partials in a fixed order: RP column sums, then the four terms
  __syncthreads();
  // entry-major: part[e * gridDim.x + block], so the reduction reads rows
  double * const part = a.part + blk;
  uint64_t const pstride = nblk;
  #pragma unroll
  for(int t=0; t < TT; ++t) {
    cs[t] += __shfl_xor(cs[t], 16, 64);
    cs[t] += __shfl_xor(cs[t], 32, 64);
  }
  if(kk == 0) {
    #pragma unroll
    for(int t=0; t < TT; ++t) s_w[wave * DN_THREADS + t * 16 + cc] = cs[t];
  }
  __syncthreads();
  for(int e=threadIdx.x; e < RP; e += DN_THREADS) {
    part[(uint64_t) e * pstride] =
        ((s_w[e] + s_w[DN_THREADS + e]) + s_w[2 * DN_THREADS + e]) + s_w[3 * DN_THREADS + e];
  }
  __syncthreads();
  s_w[threadIdx.x] = pn;
  s_w[DN_THREADS + threadIdx.x] = pr;
  s_w[2 * DN_THREADS + threadIdx.x] = dr;
  s_w[3 * DN_THREADS + threadIdx.x] = dn;
  __syncthreads();
  for(int h=DN_THREADS / 2; h > 0; h >>= 1) {
    if((int) threadIdx.x < h) {
      #pragma unroll
      for(int k=0; k < 4; ++k) s_w[k * DN_THREADS + threadIdx.x] += s_w[k * DN_THREADS + threadIdx.x + h];
    }
    __syncthreads();
  }
  if(threadIdx.x < 4) part[(uint64_t) (RP + threadIdx.x) * pstride] = s_w[threadIdx.x * DN_THREADS];
}

template<int TT, typename TM, int MODE, bool O32>
__global__ void __launch_bounds__(DN_THREADS)
k_admm_pass(Admm2Args<TM> a)
{
  b_admm_pass<TT, TM, MODE, O32>(a, blockIdx.x, gridDim.x);
}

// After pass j (one block of RD_THREADS): the block partials (entry-major)
// are added by wave w for entries w, w + 16, ...: lane l sums blocks l,
// l + 64, ... in order, then a butterfly. Norms of P'_j for the next prox,
// the test of iteration j-1, and ctl (ctl[1] = index of the state that is
// final so far: j, or j-1 on convergence, which sets ctl[0]). The code is
// kept short: a single block runs it and its instruction fetches are cold.
#define RD_THREADS 1024
#define ADMM_MAX_BLOCKS 1024       // blocks of an ADMM pass, at most (a multiple of 64)
template<int E>
__device__ __forceinline__ void
b_admm_pass_reduce(double const * __restrict__ part, uint32_t nb, uint32_t R,
    double * __restrict__ norms, double * __restrict__ sums, int * __restrict__ ctl,
    double tol, int check, int j)
{
  if(ctl[0] != 0) return;
  constexpr int RP = E - 4;
  constexpr int NW = RD_THREADS / 64;
  __shared__ double s_t[E];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  // a lane's partials of an entry (blocks lane, lane + 64, ... < nb <=
  // ADMM_MAX_BLOCKS) are loaded at once, the next entry's while this one is
  // summed in block order
  constexpr int PB = ADMM_MAX_BLOCKS / 64;
  double xa[PB], xb[PB];
  auto fetch = [&](int e, double (&x)[PB]) {
    double const * const pe = part + (uint64_t) e * nb;
    #pragma unroll
    for(int k=0; k < PB; ++k) {
      uint32_t const b = lane + 64u * k;
      x[k] = (b < nb) ? pe[b] : 0.;
    }
  };
  auto total = [&](double const (&x)[PB]) {
    double v = 0.;
    #pragma unroll
    for(int k=0; k < PB; ++k) {
      if(lane + 64u * k < nb) v += x[k];
    }
    for(int off=32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  };
  if(wave < E) fetch(wave, xa);
  for(int e=wave; e < E; e += 2 * NW) {
    if(e + NW < E) fetch(e + NW, xb);
    double const va = total(xa);
    if(lane == 0) s_t[e] = va;
    if(e + NW < E) {
      if(e + 2 * NW < E) fetch(e + 2 * NW, xa);
      double const vb = total(xb);
      if(lane == 0) s_t[e + NW] = vb;
    }
  }
  __syncthreads();
  if(threadIdx.x < RP) {
    if(threadIdx.x < (int) R) norms[threadIdx.x] = sqrt(s_t[threadIdx.x]);
  } else if(threadIdx.x == RP) {
    bool conv = false;
    if(check) {
      for(int k=0; k < 4; ++k) sums[k] = s_t[RP + k];
      conv = (s_t[RP + 1] <= tol * s_t[RP] && s_t[RP + 2] <= tol * s_t[RP + 3]);
    }
    if(conv) {
      ctl[0] = 1;
      ctl[1] = j - 1;
    } else {
      ctl[1] = j;
    }
  }
}

template<int E>
__global__ void __launch_bounds__(RD_THREADS)
k_admm_pass_reduce(double const * __restrict__ part, uint32_t nb, uint32_t R,
    double * __restrict__ norms, double * __restrict__ sums, int * __restrict__ ctl,
    double tol, int check, int j)
{
  b_admm_pass_reduce<E>(part, nb, R, norms, sums, ctl, tol, check, j);
}

// The same reduction with one 64-lane block per entry, all entries at once:
// block e's lane l sums blocks l, l + 64, ... of entry e in order, then the
// butterfly (the sums of b_admm_pass_reduce, bit for bit) and writes
// tot[e]; the last block to finish (a ticket counter, reset by it) runs the
// norms and the test of iteration j-1.
template<int E>
__global__ void __launch_bounds__(64)
k_admm_pass_reduce_e(double const * __restrict__ part, uint32_t nb, uint32_t R,
    double * __restrict__ norms, double * __restrict__ sums, int * __restrict__ ctl,
    double tol, int check, int j, double * __restrict__ tot, unsigned * __restrict__ ticket)
{
  if(ctl[0] != 0) return;                  // every block: converged earlier
  constexpr int RP = E - 4;
  constexpr int PB = ADMM_MAX_BLOCKS / 64;
  int const lane = threadIdx.x;
  int const e = blockIdx.x;
  double const * const pe = part + (uint64_t) e * nb;
  double x[PB];
  #pragma unroll
  for(int k=0; k < PB; ++k) {
    uint32_t const b = lane + 64u * k;
    x[k] = (b < nb) ? pe[b] : 0.;
  }
  double v = 0.;
  #pragma unroll
  for(int k=0; k < PB; ++k) {
    if(lane + 64u * k < nb) v += x[k];
  }
  for(int off=32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ int s_last;
  if(lane == 0) {
    tot[e] = v;
    __threadfence();
    s_last = (atomicAdd(ticket, 1u) == (unsigned) (E - 1));
  }
  __syncthreads();
  if(!s_last) return;
  __threadfence();
  for(int c=lane; c < RP; c += 64) {
    if(c < (int) R) norms[c] = sqrt(tot[c]);
  }
  if(lane == 0) {
    bool conv = false;
    if(check) {
      for(int k=0; k < 4; ++k) sums[k] = tot[RP + k];
      conv = (tot[RP + 1] <= tol * tot[RP] && tot[RP + 2] <= tol * tot[RP + 3]);
    }
    if(conv) {
      ctl[0] = 1;
      ctl[1] = j - 1;
    } else {
      ctl[1] = j;
    }
    *ticket = 0u;
  }
}

struct Bufs3 { double * b[3]; double * n[3]; };

// P = prox(P'_k), D = P - P'_k for the final state k = ctl[1] (k >= 1).
__device__ __forceinline__ void
b_admm_final(Bufs3 bf, double * __restrict__ P, double * __restrict__ D,
    uint64_t n, uint32_t R, int prox, double lam, double const * __restrict__ rho,
    int const * __restrict__ ctl,
    uint32_t blk, uint32_t nblk)
{
  int const k = ctl[1];
  double const thr = dn_prox_scalar(prox, lam, *rho);
  double const * const src = bf.b[k % 3];
  double const * const nrm = bf.n[k % 3];
  uint64_t e = (uint64_t) blk * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) nblk * blockDim.x;
  uint32_t c = (uint32_t) (e % R);
  uint32_t const sc = (uint32_t) (stride % R);
  for(; e < n; e += stride) {
    double const p1 = src[e];
    double const p = dn_prox_apply(p1, prox, dn_prox_scalar_kind(prox) ? thr : nrm[c]);
    P[e] = p;
    D[e] = p - p1;
    c += sc;
    if(c >= R) c -= R;
  }
}

__global__ void
k_admm_final(Bufs3 bf, double * __restrict__ P, double * __restrict__ D,
    uint64_t n, uint32_t R, int prox, double lam, double const * __restrict__ rho,
    int const * __restrict__ ctl)
{
  b_admm_final(bf, P, D, n, R, prox, lam, rho, ctl, blockIdx.x, gridDim.x);
}

__global__ void k_ctl_reset(int * ctl, unsigned * ticket)
{
  if(threadIdx.x < 2) ctl[threadIdx.x] = 0;
  if(threadIdx.x == 2 && ticket != nullptr) *ticket = 0u;
}

// Chunked ADMM for a row-separable proximal operator (src/cpd/admm.c:355-400
// around p_admm_iterate_chunk, :240-308): block b owns rows [b * chunk,
// min((b + 1) * chunk, I)) and iterates them to their own convergence. One
// pass per iteration and row tile: aux = M + rho (P + D) into LDS, then per
// entry the product with the inverse, P = prox(aux - D), D += P - aux and
// the residual sums. Sums over the chunk are added in a fixed order
// (thread, wave butterfly, waves in order). GL: the inverse (R <= 64) sits
// in LDS. SX: the prox projects every row onto the probability simplex
// (src/cpd/constraints/simplex.cc:32-66): the tile's pre-prox rows are kept
// in LDS, each value's place in its row's descending order is counted, the
// row's threshold comes from the running sum in that order, and each entry
// becomes max(p - theta, 0). its_rows += iterations x rows of the chunk.
#define CH_TILE 3072
template<typename TM, bool GL, bool SX>
__global__ void __launch_bounds__(DN_THREADS)
k_admm_chunk(TM const * __restrict__ M, double const * __restrict__ ginv, uint32_t RP,
    double const * __restrict__ rho_p, uint64_t I, uint32_t R, uint64_t chunk,
    double * __restrict__ P, double * __restrict__ D, int prox, double lam, uint32_t max_inner,
    double tol, unsigned long long * __restrict__ its_rows)
{
  __shared__ double s_aux[CH_TILE];      // aux; with SX, then the rows in descending order
  __shared__ double s_x[SX ? CH_TILE : 1];
  __shared__ double s_p[SX ? CH_TILE : 1];
  __shared__ double s_th[SX ? CH_TILE : 1];
  __shared__ double s_g[GL ? 64 * 64 : 1];
  __shared__ double s_w[DN_THREADS / 64][4];
  __shared__ int s_done;
  int const t = threadIdx.x;
  uint64_t const r0 = (uint64_t) blockIdx.x * chunk;
  uint64_t const r1 = (r0 + chunk < I) ? r0 + chunk : I;
  if(GL) {
    for(uint32_t e=t; e < R * R; e += DN_THREADS) {
      uint32_t const i = e / R;
      s_g[e] = ginv[(uint64_t) i * RP + (e - i * R)];
    }
  }
  uint32_t const TR = CH_TILE / R;
  double const rho = *rho_p;
  double const thr = dn_prox_scalar(prox, lam, rho);
  uint32_t it = 0;
  while(it < max_inner) {
    double pn = 0., pr = 0., dr = 0., dn = 0.;
    for(uint64_t t0=r0; t0 < r1; t0 += TR) {
      uint64_t const nr = (t0 + TR < r1) ? TR : r1 - t0;
      uint32_t const ne = (uint32_t) nr * R;
      uint64_t const base = t0 * R;
      __syncthreads();
      for(uint32_t e=t; e < ne; e += DN_THREADS) {
        s_aux[e] = (double) M[base + e] + rho * (P[base + e] + D[base + e]);
      }
      __syncthreads();
      if constexpr (SX) {
        for(uint32_t e=t; e < ne; e += DN_THREADS) {
          uint32_t const i = e / R;
          uint32_t const j = e - i * R;
          double const * const arow = s_aux + (uint64_t) i * R;
          double x = 0.;
          for(uint32_t k=0; k < R; ++k) {
            double const g = GL ? s_g[k * R + j] : ginv[(uint64_t) k * RP + j];
            x += arow[k] * g;
          }
          s_x[e] = x;
          s_p[e] = x - D[base + e];
        }
        __syncthreads();
        // place of each value in its row's non-increasing order (ties by
        // column), written into the row's slot of s_aux
        for(uint32_t e=t; e < ne; e += DN_THREADS) {
          uint32_t const i = e / R;
          uint32_t const j = e - i * R;
          double const v = s_p[e];
          uint32_t pos = 0;
          for(uint32_t k=0; k < R; ++k) {
            double const u = s_p[i * R + k];
            pos += (u > v || (u == v && k < j)) ? 1u : 0u;
          }
          s_aux[i * R + pos] = v;
        }
        __syncthreads();
        for(uint32_t i=t; i < (uint32_t) nr; i += DN_THREADS) {
          double const * const b = s_aux + (uint64_t) i * R;
          double run = -1.;
          uint32_t piv;
          for(piv=0; piv < R; ++piv) {
            run += b[piv];
            double const v = b[piv] - (run / (double) (piv + 1));
            if(v <= 0.) {
              run -= b[piv];
              break;
            }
          }
          s_th[i] = run / (double) piv;
        }
        __syncthreads();
      }
      for(uint32_t e=t; e < ne; e += DN_THREADS) {
        uint32_t const i = e / R;
        uint32_t const j = e - i * R;
        double x;
        if constexpr (SX) {
          x = s_x[e];
        } else {
          double const * const arow = s_aux + (uint64_t) i * R;
          x = 0.;
          for(uint32_t k=0; k < R; ++k) {
            double const g = GL ? s_g[k * R + j] : ginv[(uint64_t) k * RP + j];
            x += arow[k] * g;
          }
        }
        double const p0 = P[base + e];
        double const d0 = D[base + e];
        double p = x - d0;
        if constexpr (SX) {
          p = s_p[e] - s_th[i];
          p = (p > 0.) ? p : 0.;
        } else {
          p = dn_prox_entry(p, prox, thr);
        }
        double const d = d0 + (p - x);
        P[base + e] = p;
        D[base + e] = d;
        double const pd = p - x;
        double const dd = p - p0;
        pn += p * p;
        pr += pd * pd;
        dr += dd * dd;
        dn += d * d;
      }
    }
    for(int off=32; off > 0; off >>= 1) {
      pn += __shfl_xor(pn, off, 64);
      pr += __shfl_xor(pr, off, 64);
      dr += __shfl_xor(dr, off, 64);
      dn += __shfl_xor(dn, off, 64);
    }
    if((t & 63) == 0) {
      s_w[t >> 6][0] = pn;
      s_w[t >> 6][1] = pr;
      s_w[t >> 6][2] = dr;
      s_w[t >> 6][3] = dn;
    }
    __syncthreads();
    if(t == 0) {
      double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
      for(int w=0; w < DN_THREADS / 64; ++w) {
        a0 += s_w[w][0];
        a1 += s_w[w][1];
        a2 += s_w[w][2];
        a3 += s_w[w][3];
      }
      s_done = (a1 <= tol * a0 && a2 <= tol * a3) ? 1 : 0;
    }
    __syncthreads();
    ++it;
    if(s_done) break;
  }
  if(t == 0) atomicAdd(its_rows, (unsigned long long) it * (unsigned long long) (r1 - r0));
}

// Plain product for ranks above 64: out[i,j] = sum_k In[i,k] ginv[k,j].
template<typename TM>
__global__ void k_apply_plain(TM const * __restrict__ In, double const * __restrict__ ginv,
    uint64_t I, uint32_t R, uint32_t ld, double * __restrict__ out)
{
  uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const n = I * R;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; e < n; e += stride) {
    uint64_t const i = e / R;
    uint32_t const j = (uint32_t) (e - i * R);
    double s = 0.;
    for(uint32_t k=0; k < R; ++k) s += (double) In[i * R + k] * ginv[(uint64_t) k * ld + j];
    out[e] = s;
  }
}

// Plain accumulate for ranks above 64: out[i,j] += sum_k In[i,k] G[k,j].
__global__ void k_apply_plain_acc(double const * __restrict__ In,
    double const * __restrict__ G, uint64_t I, uint32_t R, uint32_t ld,
    double * __restrict__ out)
{
  uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const n = I * R;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; e < n; e += stride) {
    uint64_t const i = e / R;
    uint32_t const j = (uint32_t) (e - i * R);
    double s = 0.;
    for(uint32_t k=0; k < R; ++k) s += In[i * R + k] * G[(uint64_t) k * ld + j];
    out[e] += s;
  }
}

// ADMM, plain path: T = M + rho (P + D), init = P.
template<typename TM>
__global__ void k_admm_form(TM const * __restrict__ M, double const * __restrict__ rho,
    uint64_t n, double const * __restrict__ P, double const * __restrict__ D,
    double * __restrict__ T, double * __restrict__ init)
{
  uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  double const r = *rho;
  for(; e < n; e += stride) {
    double const p = P[e];
    init[e] = p;
    T[e] = (double) M[e] + r * (p + D[e]);
  }
}

// ADMM, plain path: P = aux - D.
__global__ void k_admm_sub(double const * __restrict__ aux, double const * __restrict__ D,
    uint64_t n, double * __restrict__ P)
{
  uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; e < n; e += stride) P[e] = aux[e] - D[e];
}

// ADMM: prox, dual update and the four residual sums (block partials).
__global__ void __launch_bounds__(DN_THREADS)
k_admm_post(double * __restrict__ P, double const * __restrict__ aux,
    double * __restrict__ D, double const * __restrict__ init,
    double const * __restrict__ norms, int prox, double lam, double const * __restrict__ rho,
    uint64_t n, uint32_t R, double * __restrict__ part)
{
  double const thr = dn_prox_scalar(prox, lam, *rho);
  __shared__ double s[4][DN_THREADS];
  uint64_t e = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  uint32_t c = (uint32_t) (e % R);
  uint32_t const sc = (uint32_t) (stride % R);
  double pn = 0., pr = 0., dr = 0., dn = 0.;
  for(; e < n; e += stride) {
    double p = P[e];
    if(prox == DN_PROX_NONNEG || dn_prox_scalar_kind(prox)) {
      p = dn_prox_entry(p, prox, thr);
    } else if(prox == DN_PROX_MAXCOLNORM || prox == DN_PROX_MAXCOLNORM_NONNEG) {
      if(prox == DN_PROX_MAXCOLNORM_NONNEG && p < 0.) p = 0.;
      double const nc = norms[c];
      if(!(nc <= 1.)) p /= nc;
    }
    P[e] = p;
    double const x = aux[e];
    double const d = D[e] + (p - x);
    D[e] = d;
    double const q = p - x;
    double const w = p - init[e];
    pn += p * p;
    pr += q * q;
    dr += w * w;
    dn += d * d;
    c += sc;
    if(c >= R) c -= R;
  }
  int const t = threadIdx.x;
  s[0][t] = pn; s[1][t] = pr; s[2][t] = dr; s[3][t] = dn;
  __syncthreads();
  for(int h=DN_THREADS / 2; h > 0; h >>= 1) {
    if(t < h) {
      #pragma unroll
      for(int k=0; k < 4; ++k) s[k][t] += s[k][t + h];
    }
    __syncthreads();
  }
  if(t < 4) part[(uint64_t) blockIdx.x * 4 + t] = s[t][0];
}

__global__ void __launch_bounds__(DN_THREADS)
k_sum4(double const * __restrict__ part, uint32_t nb, double * __restrict__ out)
{
  __shared__ double s[DN_THREADS / 4][4];
  double const v = block_partsum<4>(part, nb, 4, true, threadIdx.x % 4, s);
  if(threadIdx.x < 4) out[threadIdx.x] = v;
}

__device__ __forceinline__ void
b_cpd_norm(GramSrc g, int nmodes, uint32_t R,
    double const * __restrict__ lambda, double * __restrict__ out)
{
  __shared__ double s[64];
  int const t = threadIdx.x;
  double mine = 0.;
  for(uint32_t i=t; i < R; i += 64) {
    double row = 0.;
    for(uint32_t j=i; j < R; ++j) {
      double h = 1.;
      for(int m=0; m < nmodes; ++m) h *= g.aTa[m][i * R + j];
      row += (i == j) ? h * lambda[i] * lambda[i] : h * lambda[i] * lambda[j] * 2.;
    }
    mine += row;
  }
  s[t] = mine;
  __syncthreads();
  if(t == 0) {
    double tot = 0.;
    for(int k=0; k < 64; ++k) tot += s[k];
    out[0] = fabs(tot);
  }
}

__global__ void
k_cpd_norm(GramSrc g, int nmodes, uint32_t R,
    double const * __restrict__ lambda, double * __restrict__ out)
{
  b_cpd_norm(g, nmodes, R, lambda, out);
}

// ---- host wrappers -----------------------------------------------------------------

#define SLOT_GRAM 4
#define SLOT_COL 5
#define SLOT_CHOL 6
#define SLOT_TMP 7

template<int TT>
static int launch_gram_mfma(double * X, uint64_t I, uint32_t R, double const * div,
    float * X32, double * part, uint32_t nb, hipStream_t st, double const * Y)
{
  if(Y) {
    hipLaunchKernelGGL((k_gram_mfma<TT, false, true>), dim3(nb), dim3(DN_THREADS), 0, st,
        X, I, R, div, X32, part, Y);
  } else if(div) {
    hipLaunchKernelGGL((k_gram_mfma<TT, true, false>), dim3(nb), dim3(DN_THREADS), 0, st,
        X, I, R, div, X32, part, Y);
  } else {
    hipLaunchKernelGGL((k_gram_mfma<TT, false, false>), dim3(nb), dim3(DN_THREADS), 0, st,
        X, I, R, div, X32, part, Y);
  }
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dn_div_cols(double * X, uint64_t I, uint32_t R, double const * div,
    float * X32, hipStream_t st)
{
  uint64_t const n = I * R;
  if(n == 0) return SPD_OK;
  hipLaunchKernelGGL(k_div_cols, dim3(dn_blocks(n, DN_THREADS * 8)), dim3(DN_THREADS),
      0, st, X, n, R, div, X32);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

template<typename TY, bool CLAMP>
static int coldot(double const * X, TY const * Y, uint64_t I, uint32_t R, int mode,
    double const * lambda, double * out, hipStream_t st)
{
  uint32_t const nb = dn_blocks(I, 64);
  uint64_t const rpb = spd_cdiv(I > 0 ? I : 1, nb);
  void * part = nullptr;
  SPD_TRY(spd_scratch(SLOT_COL, (size_t) nb * R * sizeof(double), &part));
  hipLaunchKernelGGL((k_coldot<TY, CLAMP>), dim3(nb), dim3(DN_THREADS), 0, st,
      X, Y, I, R, rpb, (double *) part);
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_colreduce, dim3(1), dim3(DN_THREADS), 0, st,
      (double const *) part, nb, R, R, mode, lambda, out);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dn_colnorm(double const * X, uint64_t I, uint32_t R, double * lam, hipStream_t st)
{
  return coldot<double, false>(X, (double const *) nullptr, I, R, 0, nullptr, lam, st);
}

int dn_innerprod(double const * X, void const * M, int mdtype, uint64_t I,
    uint32_t R, double const * lambda, double * out, hipStream_t st)
{
  if(mdtype == SPD_F32) {
    return coldot<float, false>(X, (float const *) M, I, R, 1, lambda, out, st);
  }
  return coldot<double, false>(X, (double const *) M, I, R, 1, lambda, out, st);
}

static int gram_impl(double * X, double const * Y, uint64_t I, uint32_t R,
    double const * div, float * X32, double * out, hipStream_t st)
{
  int const sym = (Y == nullptr);
  if(R > 64) {
    // plain path: divide / copy first, then one entry per thread
    if(div || X32) SPD_TRY(dn_div_cols(X, I, R, div, X32, st));
    uint32_t const nb = dn_blocks(I, 64);
    uint64_t const rpb = spd_cdiv(I > 0 ? I : 1, nb);
    void * part = nullptr;
    SPD_TRY(spd_scratch(SLOT_GRAM, (size_t) nb * R * R * sizeof(double), &part));
    hipLaunchKernelGGL(k_gram_plain, dim3(nb), dim3(DN_THREADS), 0, st,
        (double const *) X, sym ? (double const *) X : Y, I, R, rpb, (double *) part);
    SPD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned) spd_cdiv((uint64_t) R * R, 32)),
        dim3(GR_RED_THREADS), 0, st, (double const *) part, nb, R, R, out, sym);
    SPD_LAUNCH_CHECK();
    return SPD_OK;
  }
  uint32_t const TT = dn_rpad(R) / 16;
  uint32_t const RP = TT * 16;
  uint64_t const nsteps = (I + 3) / 4;
  uint32_t const nb = dn_blocks(nsteps, 4 * GR_U * 4);
  void * part = nullptr;
  SPD_TRY(spd_scratch(SLOT_GRAM, (size_t) nb * RP * RP * sizeof(double), &part));
  switch(TT) {
  case 1: SPD_TRY(launch_gram_mfma<1>(X, I, R, div, X32, (double *) part, nb, st, Y)); break;
  case 2: SPD_TRY(launch_gram_mfma<2>(X, I, R, div, X32, (double *) part, nb, st, Y)); break;
  case 3: SPD_TRY(launch_gram_mfma<3>(X, I, R, div, X32, (double *) part, nb, st, Y)); break;
  default: SPD_TRY(launch_gram_mfma<4>(X, I, R, div, X32, (double *) part, nb, st, Y)); break;
  }
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned) spd_cdiv((uint64_t) R * R, 32)),
      dim3(GR_RED_THREADS), 0, st, (double const *) part, nb, R, RP, out, sym);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dn_gram(double * X, uint64_t I, uint32_t R, double const * div, float * X32,
    double * out, hipStream_t st)
{
  return gram_impl(X, nullptr, I, R, div, X32, out, st);
}

int dn_xgram(double const * X, double const * Y, uint64_t I, uint32_t R, double * out,
    hipStream_t st)
{
  return gram_impl((double *) X, Y, I, R, nullptr, nullptr, out, st);
}

int dn_chol_inv(double const * const * aTa, int nmodes, int mode, uint32_t R,
    double diag_add, int admm, double * ginv, double * rho, int * info,
    hipStream_t st)
{
  if(nmodes > SPD_MAX_MODES || R == 0) {
    spd_set_error(SPD_EINPUT, "bad Gram shape");
    return SPD_EINPUT;
  }
  GramSrc g = {};
  for(int m=0; m < nmodes; ++m) g.aTa[m] = aTa[m];
  uint32_t const RP = dn_rpad(R);
  if(RP <= 64) {
    switch(RP) {
    case 16: hipLaunchKernelGGL((k_gj_inv<16, 64>), dim3(1), dim3(64), 0, st, g, nmodes, mode,
                 R, diag_add, admm, ginv, rho, info); break;
    case 32: hipLaunchKernelGGL((k_gj_inv<32, 128>), dim3(1), dim3(128), 0, st, g, nmodes, mode,
                 R, diag_add, admm, ginv, rho, info); break;
    case 48: hipLaunchKernelGGL((k_gj_inv<48, 128>), dim3(1), dim3(128), 0, st, g, nmodes, mode,
                 R, diag_add, admm, ginv, rho, info); break;
    default: hipLaunchKernelGGL((k_gj_inv<64, 256>), dim3(1), dim3(256), 0, st, g, nmodes, mode,
                 R, diag_add, admm, ginv, rho, info); break;
    }
    SPD_LAUNCH_CHECK();
    return SPD_OK;
  }
  size_t const wbytes = 2ull * R * R * sizeof(double);
  double * gW = nullptr;
  size_t shmem = 0;
  if(wbytes <= 64 * 1024) {
    shmem = wbytes;
  } else {
    void * p = nullptr;
    SPD_TRY(spd_scratch(SLOT_CHOL, wbytes, &p));
    gW = (double *) p;
  }
  hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(64), shmem, st, g, nmodes, mode, R,
      dn_rpad(R), diag_add, admm, gW, ginv, rho, info);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

template<int TT, typename TM, int KIND>
static int launch_apply(ApplyArgs<TM> const & a, hipStream_t st, uint32_t nb)
{
  hipLaunchKernelGGL((k_apply_mfma<TT, TM, KIND>), dim3(nb), dim3(DN_THREADS), 0, st, a);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

template<typename TM, int KIND>
static int dispatch_apply(ApplyArgs<TM> const & a, uint32_t TT, uint32_t nb, hipStream_t st)
{
  switch(TT) {
  case 1: return launch_apply<1, TM, KIND>(a, st, nb);
  case 2: return launch_apply<2, TM, KIND>(a, st, nb);
  case 3: return launch_apply<3, TM, KIND>(a, st, nb);
  default: return launch_apply<4, TM, KIND>(a, st, nb);
  }
}

static uint32_t apply_blocks(uint64_t I)
{
  return dn_blocks((I + 15) / 16, 4 * 4);
}

template<typename TM>
static int apply_t(TM const * M, double const * ginv, uint64_t I, uint32_t R,
    double * X, double * colnorm, hipStream_t st)
{
  if(I == 0) return SPD_OK;
  uint32_t const RP = dn_rpad(R);
  if(R > 64) {
    hipLaunchKernelGGL((k_apply_plain<TM>), dim3(dn_blocks(I * R, DN_THREADS * 4)),
        dim3(DN_THREADS), 0, st, M, ginv, I, R, RP, X);
    SPD_LAUNCH_CHECK();
    if(colnorm) SPD_TRY(dn_colnorm(X, I, R, colnorm, st));
    return SPD_OK;
  }
  uint32_t const nb = apply_blocks(I);
  void * part = nullptr;
  if(colnorm) SPD_TRY(spd_scratch(SLOT_COL, (size_t) nb * RP * sizeof(double), &part));
  ApplyArgs<TM> a = {};
  a.M = M;
  a.ginv = ginv;
  a.I = I;
  a.R = R;
  a.X = X;
  a.colsq = (double *) part;
  SPD_TRY((dispatch_apply<TM, 0>(a, RP / 16, nb, st)));
  if(colnorm) {
    hipLaunchKernelGGL(k_colreduce, dim3(1), dim3(DN_THREADS), 0, st,
        (double const *) part, nb, R, RP, 0, (double const *) nullptr, colnorm);
    SPD_LAUNCH_CHECK();
  }
  return SPD_OK;
}

int dn_apply_acc(double const * In, double const * G, uint64_t I, uint32_t R,
    double * out, hipStream_t st)
{
  if(I == 0) return SPD_OK;
  uint32_t const RP = dn_rpad(R);
  if(R > 64) {
    hipLaunchKernelGGL(k_apply_plain_acc, dim3(dn_blocks(I * R, DN_THREADS * 4)),
        dim3(DN_THREADS), 0, st, In, G, I, R, RP, out);
    SPD_LAUNCH_CHECK();
    return SPD_OK;
  }
  ApplyArgs<double> a = {};
  a.M = In;
  a.ginv = G;
  a.I = I;
  a.R = R;
  a.X = out;
  a.colsq = nullptr;
  return dispatch_apply<double, 2>(a, RP / 16, apply_blocks(I), st);
}

int dn_apply(void const * M, int mdtype, double const * ginv, uint64_t I,
    uint32_t R, double * X, double * colnorm, hipStream_t st)
{
  if(mdtype == SPD_F32) return apply_t<float>((float const *) M, ginv, I, R, X, colnorm, st);
  return apply_t<double>((double const *) M, ginv, I, R, X, colnorm, st);
}

// ---- smoothness prox (src/cpd/constraints/smooth.c:30-184) ---------------
// Each column x of P becomes rho A^{-1} x with A = lam K + rho I, K the
// reference's banded second-difference matrix (p_form_banded: stencil
// [1 -4 6 -4 1], first and last diagonal entries 5). A is symmetric
// positive definite, so it is factored A = L D L^T without pivoting (L unit
// lower with two subdiagonals) by one thread; then one lane per column runs
// the forward and backward substitutions down the rows (rows are read
// coalesced across the lanes). Both passes are sequential in I: this is a
// latency-bound path for a constraint off the hot path.
#define SLOT_SMOOTH 8

// f[3i] = D_i, f[3i+1] = L(i, i-1), f[3i+2] = L(i, i-2)
__global__ void k_smooth_factor(uint64_t I, double lam, double const * __restrict__ rho_p,
    double * __restrict__ f)
{
  if(threadIdx.x != 0 || blockIdx.x != 0) return;
  double const rho = *rho_p;
  double dm2 = 0., dm1 = 0., l1m1 = 0.;
  for(uint64_t i=0; i < I; ++i) {
    double const a = ((i == 0 || i == I - 1) ? 5. * lam : 6. * lam) + rho;
    double const l2 = (i >= 2) ? lam / dm2 : 0.;
    double const l1 = (i >= 1) ? (-4. * lam - l2 * dm2 * l1m1) / dm1 : 0.;
    double const d = a - l2 * l2 * dm2 - l1 * l1 * dm1;
    f[3 * i] = d;
    f[3 * i + 1] = l1;
    f[3 * i + 2] = l2;
    dm2 = dm1;
    dm1 = d;
    l1m1 = l1;
  }
}

__global__ void k_smooth_solve(double * __restrict__ P, uint64_t I, uint32_t R,
    double const * __restrict__ f, double const * __restrict__ rho_p)
{
  uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
  if(c >= R) return;
  double const rho = *rho_p;
  double y1 = 0., y2 = 0.;
  for(uint64_t i=0; i < I; ++i) {
    double const y = P[i * R + c] - f[3 * i + 1] * y1 - f[3 * i + 2] * y2;
    P[i * R + c] = y / f[3 * i];
    y2 = y1;
    y1 = y;
  }
  double x1 = 0., x2 = 0.;
  for(uint64_t i=I; i-- > 0;) {
    double const l1n = (i + 1 < I) ? f[3 * (i + 1) + 1] : 0.;
    double const l2n = (i + 2 < I) ? f[3 * (i + 2) + 2] : 0.;
    double const x = P[i * R + c] - l1n * x1 - l2n * x2;
    P[i * R + c] = x * rho;
    x2 = x1;
    x1 = x;
  }
}

// ---- column simplex prox (src/cpd/constraints/simplex.cc:120-158) ---------
// Each column x (I entries) becomes max(x - theta, 0) with theta = (sum of
// the support - 1) / |support|, the support being the entries above theta.
// One block per column finds it by Michelot's iteration: start from every
// entry, then keep the entries above the current theta until the support
// stops shrinking (it shrinks at each other step, so at most I rounds). The
// support is the one the reference's sort and running sum find; theta's sum
// is taken in block order instead of sorted order.
__global__ void __launch_bounds__(DN_THREADS)
k_colsimp(double * __restrict__ P, uint64_t I, uint32_t R)
{
  __shared__ double s_sum[DN_THREADS / 64];
  __shared__ unsigned long long s_cnt[DN_THREADS / 64];
  __shared__ double s_theta;
  __shared__ int s_stop;
  uint32_t const c = blockIdx.x;
  int const t = threadIdx.x;
  int const lane = t & 63;
  int const w = t >> 6;
  double theta = -HUGE_VAL;
  unsigned long long prev = ~0ull;
  while(true) {
    double sum = 0.;
    unsigned long long cnt = 0;
    for(uint64_t i=t; i < I; i += DN_THREADS) {
      double const v = P[i * R + c];
      if(v > theta) {
        sum += v;
        ++cnt;
      }
    }
    for(int off=32; off > 0; off >>= 1) {
      sum += __shfl_xor(sum, off, 64);
      cnt += __shfl_xor(cnt, off, 64);
    }
    if(lane == 0) {
      s_sum[w] = sum;
      s_cnt[w] = cnt;
    }
    __syncthreads();
    if(t == 0) {
      double a = 0.;
      unsigned long long n = 0;
      for(int k=0; k < DN_THREADS / 64; ++k) {
        a += s_sum[k];
        n += s_cnt[k];
      }
      s_stop = (n == prev || n == 0) ? 1 : 0;
      if(n > 0) s_theta = (a - 1.) / (double) n;
      prev = n;
    }
    __syncthreads();
    theta = s_theta;
    if(s_stop) break;
    __syncthreads();
  }
  for(uint64_t i=t; i < I; i += DN_THREADS) {
    double const v = P[i * R + c] - theta;
    P[i * R + c] = (v > 0.) ? v : 0.;
  }
}

// Factor the smoothness matrix for rows I and the device rho into scratch.
static int smooth_factor(uint64_t I, double lam, double const * rho, double ** f,
    hipStream_t st)
{
  void * p = nullptr;
  SPD_TRY(spd_scratch(SLOT_SMOOTH, 3 * (I ? I : 1) * sizeof(double), &p));
  hipLaunchKernelGGL(k_smooth_factor, dim3(1), dim3(64), 0, st, I, lam, rho, (double *) p);
  SPD_LAUNCH_CHECK();
  *f = (double *) p;
  return SPD_OK;
}

// sf: the factors of the smoothness matrix when prox is DN_PROX_SMOOTH
template<typename TM>
static int admm_t(TM const * M, double const * ginv, double const * rho, uint64_t I,
    uint32_t R, double * P, double * aux, double * D, double * init, int prox, double lam,
    double const * sf, double * sums, hipStream_t st)
{
  if(I == 0) return SPD_OK;
  uint32_t const RP = dn_rpad(R);
  uint64_t const n = I * R;
  bool const need_norms = (prox == DN_PROX_MAXCOLNORM || prox == DN_PROX_MAXCOLNORM_NONNEG);
  int const clamp = (prox == DN_PROX_MAXCOLNORM_NONNEG);
  // column norms go into the tail of the sums buffer: sums[4 .. 4+R)
  double * const norms = sums + 4;
  if(R > 64) {
    void * tmp = nullptr;
    SPD_TRY(spd_scratch(SLOT_TMP, n * sizeof(double), &tmp));
    uint32_t const nbe = dn_blocks(n, DN_THREADS * 4);
    hipLaunchKernelGGL((k_admm_form<TM>), dim3(nbe), dim3(DN_THREADS), 0, st,
        M, rho, n, (double const *) P, (double const *) D, (double *) tmp, init);
    SPD_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_apply_plain<double>), dim3(nbe), dim3(DN_THREADS), 0, st,
        (double const *) tmp, ginv, I, R, RP, aux);
    SPD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_admm_sub, dim3(nbe), dim3(DN_THREADS), 0, st,
        (double const *) aux, (double const *) D, n, P);
    SPD_LAUNCH_CHECK();
    if(need_norms) {
      if(clamp) {
        SPD_TRY((coldot<double, true>(P, (double const *) nullptr, I, R, 0, nullptr, norms, st)));
      } else {
        SPD_TRY((coldot<double, false>(P, (double const *) nullptr, I, R, 0, nullptr, norms, st)));
      }
    }
  } else {
    uint32_t const nb = apply_blocks(I);
    void * part = nullptr;
    if(need_norms) SPD_TRY(spd_scratch(SLOT_COL, (size_t) nb * RP * sizeof(double), &part));
    ApplyArgs<TM> a = {};
    a.M = M;
    a.ginv = ginv;
    a.I = I;
    a.R = R;
    a.X = P;
    a.aux = aux;
    a.dual = D;
    a.init = init;
    a.rho = rho;
    a.clamp = clamp;
    a.colsq = (double *) part;
    SPD_TRY((dispatch_apply<TM, 1>(a, RP / 16, nb, st)));
    if(need_norms) {
      hipLaunchKernelGGL(k_colreduce, dim3(1), dim3(DN_THREADS), 0, st,
          (double const *) part, nb, R, RP, 0, (double const *) nullptr, norms);
      SPD_LAUNCH_CHECK();
    }
  }
  if(prox == DN_PROX_SMOOTH) {
    hipLaunchKernelGGL(k_smooth_solve, dim3((unsigned) spd_cdiv(R, 64)), dim3(64), 0, st,
        P, I, R, sf, rho);
    SPD_LAUNCH_CHECK();
  }
  if(prox == DN_PROX_COLSIMP) {
    hipLaunchKernelGGL(k_colsimp, dim3(R), dim3(DN_THREADS), 0, st, P, I, R);
    SPD_LAUNCH_CHECK();
  }
  uint32_t const nb2 = dn_blocks(n, DN_THREADS * 8);
  void * part2 = nullptr;
  SPD_TRY(spd_scratch(SLOT_GRAM, (size_t) nb2 * 4 * sizeof(double), &part2));
  hipLaunchKernelGGL(k_admm_post, dim3(nb2), dim3(DN_THREADS), 0, st, P,
      (double const *) aux, D, (double const *) init, (double const *) norms, prox, lam, rho,
      n, R, (double *) part2);
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum4, dim3(1), dim3(DN_THREADS), 0, st, (double const *) part2, nb2, sums);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

template<int TT, typename TM>
static int launch_admm2(Admm2Args<TM> const & a, int j, uint32_t nb, double * norms,
    double * sums, int * ctl, double tol, double * tot, unsigned * ticket, hipStream_t st)
{
  constexpr int E = TT * 16 + 4;
  // every matrix of the pass (I x R, 8 bytes at most) below 2 GiB
  bool const o32 = (uint64_t) a.I * a.R * sizeof(double) < (1ull << 31);
  if(o32) {
    if(j == 1) {
      hipLaunchKernelGGL((k_admm_pass<TT, TM, 0, true>), dim3(nb), dim3(DN_THREADS), 0, st, a);
    } else if(j == 2) {
      hipLaunchKernelGGL((k_admm_pass<TT, TM, 1, true>), dim3(nb), dim3(DN_THREADS), 0, st, a);
    } else {
      hipLaunchKernelGGL((k_admm_pass<TT, TM, 2, true>), dim3(nb), dim3(DN_THREADS), 0, st, a);
    }
  } else {
    if(j == 1) {
      hipLaunchKernelGGL((k_admm_pass<TT, TM, 0, false>), dim3(nb), dim3(DN_THREADS), 0, st, a);
    } else if(j == 2) {
      hipLaunchKernelGGL((k_admm_pass<TT, TM, 1, false>), dim3(nb), dim3(DN_THREADS), 0, st, a);
    } else {
      hipLaunchKernelGGL((k_admm_pass<TT, TM, 2, false>), dim3(nb), dim3(DN_THREADS), 0, st, a);
    }
  }
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_admm_pass_reduce_e<E>), dim3(E), dim3(64), 0, st,
      (double const *) a.part, nb, a.R, norms, sums, ctl, tol, j >= 2 ? 1 : 0, j, tot, ticket);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// The inner loop as passes over the pre-prox iterate (k_admm_pass): P'_j
// goes to bufs[j % 3] = {P, aux, init}; P and D hold P_0 and D_0 until pass
// 3 first overwrites P, and receive the final state at the end.
template<typename TM>
static int admm_solve2_t(TM const * M, double const * ginv, double const * rho, uint64_t I,
    uint32_t R, double * P, double * D, double * aux, double * init, int prox, double lam,
    uint64_t max_inner, double tol, double * sums, int * ctl, hipStream_t st, int d_zero)
{
  uint32_t const RP = dn_rpad(R);
  // one block per 64 rows, at most one resident round (four blocks per CU;
  // a larger grid only lengthens the one-block reduction)
  uint32_t nb = (uint32_t) (spd_cdiv(I, 64) < ADMM_MAX_BLOCKS ? spd_cdiv(I, 64)
                                                             : ADMM_MAX_BLOCKS);
  if(nb < 1) nb = 1;
  void * pc = nullptr;
  size_t const part_n = (size_t) nb * (RP + 4);
  // block partials, the three states' column norms, the entry totals of the
  // reduction and its ticket counter
  SPD_TRY(spd_scratch(SLOT_COL, (part_n + 3 * (size_t) RP + (RP + 4) + 1) * sizeof(double), &pc));
  double * const part = (double *) pc;
  double * const tot = part + part_n + 3 * (size_t) RP;
  unsigned * const ticket = (unsigned *) (tot + RP + 4);
  Bufs3 bf;
  bf.b[0] = P;
  bf.b[1] = aux;
  bf.b[2] = init;
  for(int k=0; k < 3; ++k) bf.n[k] = part + part_n + (size_t) k * RP;
  hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, st, ctl, ticket);
  SPD_LAUNCH_CHECK();
  Admm2Args<TM> a = {};
  a.M = M;
  a.ginv = ginv;
  a.I = I;
  a.R = R;
  a.rho = rho;
  a.prox = prox;
  a.lam = lam;
  a.P0 = P;
  a.D0 = D;
  a.d0zero = d_zero;
  a.part = part;
  a.ctl = ctl;
  for(uint64_t j=1; j <= max_inner; ++j) {
    a.cur = (j >= 2) ? bf.b[(j - 1) % 3] : nullptr;
    a.cur_n = (j >= 2) ? bf.n[(j - 1) % 3] : nullptr;
    a.prev = (j >= 3) ? bf.b[(j - 2) % 3] : nullptr;
    a.prev_n = (j >= 3) ? bf.n[(j - 2) % 3] : nullptr;
    a.out = bf.b[j % 3];
    double * const nrm = bf.n[j % 3];
    switch(RP / 16) {
    case 1: SPD_TRY((launch_admm2<1, TM>(a, (int) j, nb, nrm, sums, ctl, tol, tot, ticket, st))); break;
    case 2: SPD_TRY((launch_admm2<2, TM>(a, (int) j, nb, nrm, sums, ctl, tol, tot, ticket, st))); break;
    case 3: SPD_TRY((launch_admm2<3, TM>(a, (int) j, nb, nrm, sums, ctl, tol, tot, ticket, st))); break;
    default: SPD_TRY((launch_admm2<4, TM>(a, (int) j, nb, nrm, sums, ctl, tol, tot, ticket, st))); break;
    }
  }
  uint64_t const n = I * R;
  hipLaunchKernelGGL(k_admm_final, dim3(dn_blocks(n, DN_THREADS * 8)), dim3(DN_THREADS), 0, st,
      bf, P, D, n, R, prox, lam, rho, (int const *) ctl);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dn_admm_solve(void const * M, int mdtype, double const * ginv, double const * rho,
    uint64_t I, uint32_t R, double * P, double * D, double * aux, double * init, int prox,
    double lam, uint64_t max_inner, double tol, double * sums, int * ctl, hipStream_t st,
    int d_zero)
{
  if(prox == DN_PROX_ROWSIMP) {
    spd_set_error(SPD_EINPUT, "the row-simplex prox runs in chunked ADMM only");
    return SPD_EINPUT;
  }
  if(I == 0) return SPD_OK;
  if(max_inner == 0) return d_zero ? spd_zero_async(D, I * R * sizeof(double), st) : SPD_OK;
  double * sf = nullptr;
  if(prox == DN_PROX_SMOOTH) SPD_TRY(smooth_factor(I, lam, rho, &sf, st));
  if(R <= 64 && prox != DN_PROX_SMOOTH && prox != DN_PROX_COLSIMP) {
    // the pass form reads D only as the starting dual, in its first two
    // passes, and writes all of it at the end
    if(mdtype == SPD_F32) {
      return admm_solve2_t<float>((float const *) M, ginv, rho, I, R, P, D, aux, init, prox,
          lam, max_inner, tol, sums, ctl, st, d_zero);
    }
    return admm_solve2_t<double>((double const *) M, ginv, rho, I, R, P, D, aux, init, prox,
        lam, max_inner, tol, sums, ctl, st, d_zero);
  }
  if(d_zero) SPD_TRY(spd_zero_async(D, I * R * sizeof(double), st));
  // ranks above 64 and the column-coupled smoothness and simplex proxes: one
  // step at a time, the test on the host
  int hctl[2] = {0, 0};
  for(uint64_t k=0; k < max_inner; ++k) {
    if(mdtype == SPD_F32) {
      SPD_TRY(admm_t<float>((float const *) M, ginv, rho, I, R, P, aux, D, init, prox, lam, sf,
          sums, st));
    } else {
      SPD_TRY(admm_t<double>((double const *) M, ginv, rho, I, R, P, aux, D, init, prox, lam,
          sf, sums, st));
    }
    double s4[4];
    SPD_HIP(hipMemcpyAsync(s4, sums, sizeof(s4), hipMemcpyDeviceToHost, st));
    SPD_HIP(hipStreamSynchronize(st));
    hctl[1] = (int) (k + 1);
    if(s4[1] <= tol * s4[0] && s4[2] <= tol * s4[3]) {
      hctl[0] = 1;
      break;
    }
  }
  SPD_HIP(hipMemcpyAsync(ctl, hctl, sizeof(hctl), hipMemcpyHostToDevice, st));
  SPD_HIP(hipStreamSynchronize(st));
  return SPD_OK;
}

int dn_admm_step(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint32_t R, double * P, double * aux,
    double * D, double * init, int prox, double lam, double * sums, hipStream_t st)
{
  if(prox == DN_PROX_ROWSIMP) {
    spd_set_error(SPD_EINPUT, "the row-simplex prox runs in chunked ADMM only");
    return SPD_EINPUT;
  }
  double * sf = nullptr;
  if(prox == DN_PROX_SMOOTH) SPD_TRY(smooth_factor(I, lam, rho, &sf, st));
  if(mdtype == SPD_F32) {
    return admm_t<float>((float const *) M, ginv, rho, I, R, P, aux, D, init, prox, lam, sf,
        sums, st);
  }
  return admm_t<double>((double const *) M, ginv, rho, I, R, P, aux, D, init, prox, lam, sf,
      sums, st);
}

template<typename TM>
static int admm_chunked_t(TM const * M, double const * ginv, double const * rho, uint64_t I,
    uint32_t R, uint64_t chunk, double * P, double * D, int prox, double lam, uint32_t max_inner,
    double tol, unsigned long long * its_rows, hipStream_t st)
{
  uint64_t const nchunks = spd_cdiv(I, chunk);
  uint32_t const RP = dn_rpad(R);
  bool const sx = (prox == DN_PROX_ROWSIMP);
  if(R <= 64 && sx) {
    hipLaunchKernelGGL((k_admm_chunk<TM, true, true>), dim3((unsigned) nchunks),
        dim3(DN_THREADS), 0, st, M, ginv, RP, rho, I, R, chunk, P, D, prox, lam, max_inner, tol,
        its_rows);
  } else if(R <= 64) {
    hipLaunchKernelGGL((k_admm_chunk<TM, true, false>), dim3((unsigned) nchunks),
        dim3(DN_THREADS), 0, st, M, ginv, RP, rho, I, R, chunk, P, D, prox, lam, max_inner, tol,
        its_rows);
  } else if(sx) {
    hipLaunchKernelGGL((k_admm_chunk<TM, false, true>), dim3((unsigned) nchunks),
        dim3(DN_THREADS), 0, st, M, ginv, RP, rho, I, R, chunk, P, D, prox, lam, max_inner, tol,
        its_rows);
  } else {
    hipLaunchKernelGGL((k_admm_chunk<TM, false, false>), dim3((unsigned) nchunks),
        dim3(DN_THREADS), 0, st, M, ginv, RP, rho, I, R, chunk, P, D, prox, lam, max_inner, tol,
        its_rows);
  }
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dn_admm_chunked(void const * M, int mdtype, double const * ginv, double const * rho,
    uint64_t I, uint32_t R, uint64_t chunk, double * P, double * D, int prox, double lam,
    uint64_t max_inner, double tol, unsigned long long * its_rows, hipStream_t st)
{
  if(I == 0 || max_inner == 0) return SPD_OK;
  if(chunk == 0 || R == 0 || R > CH_TILE || spd_cdiv(I, chunk) > 0x7FFFFFFFull ||
      (prox != DN_PROX_NONNEG && !dn_prox_scalar_kind(prox) && prox != DN_PROX_ROWSIMP &&
       prox != DN_PROX_NONE)) {
    spd_set_error(SPD_EINPUT, "chunked ADMM: unsupported shape (rows %llu, rank %u, "
        "chunk %llu, prox %d)", (unsigned long long) I, R, (unsigned long long) chunk, prox);
    return SPD_EINPUT;
  }
  uint32_t const mi = (max_inner > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t) max_inner;
  if(mdtype == SPD_F32) {
    return admm_chunked_t<float>((float const *) M, ginv, rho, I, R, chunk, P, D, prox, lam,
        mi, tol, its_rows, st);
  }
  return admm_chunked_t<double>((double const *) M, ginv, rho, I, R, chunk, P, D, prox, lam,
      mi, tol, its_rows, st);
}

int dn_cpd_norm(double const * const * aTa, int nmodes, uint32_t R,
    double const * lambda, double * out, hipStream_t st)
{
  GramSrc g = {};
  for(int m=0; m < nmodes; ++m) g.aTa[m] = aTa[m];
  hipLaunchKernelGGL(k_cpd_norm, dim3(1), dim3(64), 0, st, g, nmodes, R, lambda, out);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// ---- exported for tests ------------------------------------------------------------

extern "C" int spd_dense_gram(double * X, uint64_t I, uint64_t R, double const * div,
    float * X32, double * out)
{
  SPD_TRY(spd_init());
  return dn_gram(X, I, (uint32_t) R, div, X32, out, spd_hs(nullptr));
}

extern "C" int spd_dense_chol_inv(double const * const * aTa, int nmodes, int mode,
    uint64_t R, double diag_add, int admm, double * ginv, double * rho, int * info)
{
  SPD_TRY(spd_init());
  SPD_TRY(spd_zero_async(info, sizeof(int), spd_hs(nullptr)));
  return dn_chol_inv(aTa, nmodes, mode, (uint32_t) R, diag_add, admm, ginv, rho, info,
      spd_hs(nullptr));
}

extern "C" int spd_dense_apply(void const * M, int mdtype, double const * ginv,
    uint64_t I, uint64_t R, double * X, double * colnorm)
{
  SPD_TRY(spd_init());
  return dn_apply(M, mdtype, ginv, I, (uint32_t) R, X, colnorm, spd_hs(nullptr));
}

extern "C" int spd_dense_admm_solve(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint64_t R, double * P, double * D, double * aux,
    double * init, int prox, double lam, uint64_t max_inner, double tol, double * sums,
    int * ctl)
{
  SPD_TRY(spd_init());
  return dn_admm_solve(M, mdtype, ginv, rho, I, (uint32_t) R, P, D, aux, init, prox, lam,
      max_inner, tol, sums, ctl, spd_hs(nullptr));
}

extern "C" int spd_dense_admm_chunked(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint64_t R, uint64_t chunk, double * P, double * D,
    int prox, double lam, uint64_t max_inner, double tol, uint64_t * its_rows)
{
  SPD_TRY(spd_init());
  hipStream_t const st = spd_hs(nullptr);
  SPD_TRY(spd_zero_async(its_rows, sizeof(uint64_t), st));
  return dn_admm_chunked(M, mdtype, ginv, rho, I, (uint32_t) R, chunk, P, D, prox, lam,
      max_inner, tol, (unsigned long long *) its_rows, st);
}

extern "C" int spd_dense_admm_step(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint64_t R, double * P, double * aux, double * D,
    double * init, int prox, double * sums)
{
  SPD_TRY(spd_init());
  return dn_admm_step(M, mdtype, ginv, rho, I, (uint32_t) R, P, aux, D, init, prox, 0.,
      sums, spd_hs(nullptr));
}

// ---- batched launches: one kernel covers every slice of a batch ----------
//
// Each kernel below runs the device body of its per-slice counterpart for
// work item (slice s, block b), with b < nb(s) computed from the slice's
// sizes by the same formula the per-slice launcher uses, so every slice
// gets exactly its per-slice arithmetic (spd_batch.hpp). Items of slices
// that a gate excludes (converged, other constraint kind, ...) are not
// counted.

#define BG_ALL 0          // every slice
#define BG_LIVE 1         // slices whose outer loop runs
#define BG_FIRST 2        // first slice of its stream
#define BG_ADMM 3         // live, mode solved by ADMM, ADMM not converged
#define BG_CLOSED 4       // live, mode solved in closed form
#define BG_SMALL 5        // slices whose local error uses this mode

__host__ __device__ __forceinline__ uint32_t bn_clamp(uint64_t nb, uint64_t hi)
{
  if(nb < 1) nb = 1;
  if(nb > hi) nb = hi;
  return (uint32_t) nb;
}
// the block counts of the per-slice launchers
__host__ __device__ __forceinline__ uint32_t bn_gram(uint64_t I)
{
  return bn_clamp(((I + 3) / 4 + 4 * GR_U * 4 - 1) / (4 * GR_U * 4), DN_MAX_BLOCKS);
}
__host__ __device__ __forceinline__ uint32_t bn_rows64(uint64_t I)
{
  return bn_clamp((I + 63) / 64, DN_MAX_BLOCKS);
}
__host__ __device__ __forceinline__ uint32_t bn_elem(uint64_t n)
{
  return bn_clamp((n + DN_THREADS * 8 - 1) / (DN_THREADS * 8), DN_MAX_BLOCKS);
}
__host__ __device__ __forceinline__ uint32_t bn_apply(uint64_t I)
{
  return bn_clamp(((I + 15) / 16 + 15) / 16, DN_MAX_BLOCKS);
}
__host__ __device__ __forceinline__ uint32_t bn_admm(uint64_t I)
{
  return bn_clamp((I + 63) / 64, ADMM_MAX_BLOCKS);
}

__device__ __forceinline__ bool bt_gate(BSlot const & s, int gate, int mode)
{
  switch(gate) {
  case BG_LIVE: return s.c->done == 0;
  case BG_FIRST: return s.first != 0;
  case BG_ADMM: return s.c->done == 0 && s.kind[mode] == SPD_CON_ADMM && s.c->admm[0] == 0;
  case BG_CLOSED: return s.c->done == 0 && s.kind[mode] != SPD_CON_ADMM;
  case BG_SMALL: return s.small == mode;
  default: return true;
  }
}

// Gram operands: what 0: Unew[mode] -> aTa[mode]; 1: Uold[mode]^T Unew[mode]
// -> cross + k R^2; 2: the time row (one row) -> tgram.
__device__ __forceinline__ double * bt_gram_x(BSlot const & s, int what, int mode)
{
  return (what == 0) ? s.Unew[mode] : ((what == 1) ? s.Uold[mode] : s.time_row);
}
__device__ __forceinline__ uint64_t bt_gram_rows(BSlot const & s, int what, int mode)
{
  return (what == 2) ? 1 : s.rows[mode];
}

template<int TT, bool CROSS>
__global__ void __launch_bounds__(DN_THREADS)
kb_gram(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int what,
    int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(bt_gate(s, gate, mode)) nb = bn_gram(bt_gram_rows(s, what, mode));
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    b_gram_mfma<TT, false, CROSS>(bt_gram_x(s, what, mode), bt_gram_rows(s, what, mode), s.R,
        nullptr, nullptr, s.part, CROSS ? s.Unew[mode] : nullptr, w - s_it.pref[q],
        s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

// the reduction: out = aTa[mode] (what 0), cross + k R^2 (1), tgram (2)
__global__ void __launch_bounds__(GR_RED_THREADS)
kb_gram_reduce(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode,
    int what, int k, int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(bt_gate(s, gate, mode)) nb = (s.R * s.R + 31) / 32;
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    uint32_t const RP = dn_rpad(s.R);
    double * const out = (what == 0) ? s.aTa[mode] :
        ((what == 1) ? s.cross + (uint64_t) k * s.R * s.R : s.tgram);
    b_gram_reduce(s.part, bn_gram(bt_gram_rows(s, what, mode)), s.R, RP, out,
        what == 1 ? 0 : 1, w - s_it.pref[q], 0);
    __syncthreads();
  }
}

// column sums of squares (colnorm) or of X o M (innerprod) of X = what 0:
// Unew[mode]; 1: the local-error factor of mode `mode` against Mbuf
__global__ void __launch_bounds__(DN_THREADS)
kb_coldot(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int what,
    int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  uint32_t const t = (uint32_t) hd->t;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(bt_gate(s, gate, mode)) nb = bn_rows64(s.rows[mode]);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    uint64_t const I = s.rows[mode];
    uint32_t const nb = s_it.pref[q + 1] - s_it.pref[q];
    uint64_t const rpb = spd_cdiv(I > 0 ? I : 1, nb);
    double const * X = ((uint32_t) mode == t) ? s.time_row : s.Unew[mode];
    b_coldot<double, false>(X, what == 0 ? (double const *) nullptr : (double const *) s.Mbuf,
        I, s.R, rpb, s.part, w - s_it.pref[q], nb);
    __syncthreads();
  }
}

// per slice: colnorm -> tmpv (what 0) or the inner product with lambda =
// ones -> c->inner[0] (what 1)
__global__ void __launch_bounds__(DN_THREADS)
kb_colreduce(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int what,
    int gate)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(!bt_gate(s, gate, mode)) return;
  b_colreduce(s.part, bn_rows64(s.rows[mode]), s.R, s.R, what,
      what == 1 ? s.lam1 : (double const *) nullptr, what == 1 ? &s.c->inner[0] : s.tmpv);
}

__global__ void __launch_bounds__(DN_THREADS)
kb_div_cols(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(bt_gate(s, gate, mode)) nb = bn_elem(s.rows[mode] * s.R);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    b_div_cols(s.Unew[mode], s.rows[mode] * s.R, s.R, s.tmpv, nullptr, w - s_it.pref[q],
        s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

// normal-equation inverse of `mode`, one block per slice: ADMM modes add
// rho (to c->rho), the streamed mode never
template<int RP, int NT>
__global__ void __launch_bounds__(NT)
kb_gj_inv(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int gate)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(!bt_gate(s, gate, mode)) return;
  GramSrc g = {};
  for(int m=0; m < hd->N; ++m) g.aTa[m] = s.aTa[m];
  int const admm = (mode != hd->t && s.kind[mode] == SPD_CON_ADMM) ? 1 : 0;
  b_gj_inv<RP, NT>(g, hd->N, mode, s.R, s.diag_add[mode], admm, s.ginv, &s.c->rho,
      &s.c->info);
}

// what 0: time row = Mbuf[0] ginv; 1: Unew[mode] = Mbuf ginv (closed form);
// 2: Mbuf += Uold[mode] H (history)
template<int TT, int KIND>
__global__ void __launch_bounds__(DN_THREADS)
kb_apply(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int what,
    int gate)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(bt_gate(s, gate, mode)) nb = bn_apply(what == 0 ? 1 : s.rows[mode]);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    ApplyArgs<double> a = {};
    a.R = s.R;
    if(what == 0) {
      a.M = s.Mbuf; a.ginv = s.ginv; a.I = 1; a.X = s.time_row;
    } else if(what == 1) {
      a.M = s.Mbuf; a.ginv = s.ginv; a.I = s.rows[mode]; a.X = s.Unew[mode];
    } else {
      a.M = s.Uold[mode]; a.ginv = s.H; a.I = s.rows[mode]; a.X = s.Mbuf;
    }
    a.colsq = nullptr;
    b_apply_mfma<TT, double, KIND>(a, w - s_it.pref[q], s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

__device__ __forceinline__ Bufs3 bt_bufs(BSlot const & s, int mode)
{
  Bufs3 bf;
  bf.b[0] = s.Unew[mode];
  bf.b[1] = s.aux;
  bf.b[2] = s.init;
  for(int k=0; k < 3; ++k) bf.n[k] = s.c->nrm[k];
  return bf;
}

// ADMM pass j of `mode` (the iterate in the three rotating buffers); the
// dual starts at zero (the stream clears it before each mode's ADMM)
template<int TT, int MODE>
__global__ void __launch_bounds__(DN_THREADS)
kb_admm_pass(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int j)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(bt_gate(s, BG_ADMM, mode) && (uint32_t) j <= s.max_inner) nb = bn_admm(s.rows[mode]);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    Bufs3 const bf = bt_bufs(s, mode);
    Admm2Args<double> a = {};
    a.M = s.Mbuf;
    a.ginv = s.ginv;
    a.I = s.rows[mode];
    a.R = s.R;
    a.rho = &s.c->rho;
    a.prox = s.prox[mode];
    a.lam = s.lam[mode];
    a.P0 = s.Unew[mode];
    a.D0 = s.Dual[mode];
    a.part = s.part;
    a.ctl = s.c->admm;
    a.cur = (j >= 2) ? bf.b[(j - 1) % 3] : nullptr;
    a.cur_n = (j >= 2) ? bf.n[(j - 1) % 3] : nullptr;
    a.prev = (j >= 3) ? bf.b[(j - 2) % 3] : nullptr;
    a.prev_n = (j >= 3) ? bf.n[(j - 2) % 3] : nullptr;
    a.out = bf.b[j % 3];
    b_admm_pass<TT, double, MODE>(a, w - s_it.pref[q], s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

template<int E>
__global__ void __launch_bounds__(RD_THREADS)
kb_admm_reduce(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode, int j)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  if(!bt_gate(s, BG_ADMM, mode) || (uint32_t) j > s.max_inner) return;
  b_admm_pass_reduce<E>(s.part, bn_admm(s.rows[mode]), s.R, s.c->nrm[j % 3], s.c->sums,
      s.c->admm, s.inner_tol, j >= 2 ? 1 : 0, j);
}

__global__ void __launch_bounds__(DN_THREADS)
kb_admm_final(BSlot const * __restrict__ S, BHead const * __restrict__ hd, int mode)
{
  __shared__ BItems s_it;
  uint32_t const K = hd->K;
  {
    uint32_t nb = 0;
    if(threadIdx.x < K) {
      BSlot const & s = S[threadIdx.x];
      if(s.c->done == 0 && s.kind[mode] == SPD_CON_ADMM) nb = bn_elem(s.rows[mode] * s.R);
    }
    bt_items(s_it, K, nb);
  }
  for(uint32_t w=blockIdx.x; w < s_it.total; w += gridDim.x) {
    uint32_t const q = bt_find(s_it, K, w);
    BSlot const & s = S[q];
    b_admm_final(bt_bufs(s, mode), s.Unew[mode], s.Dual[mode], s.rows[mode] * s.R, s.R,
        s.prox[mode], s.lam[mode], &s.c->rho, s.c->admm, w - s_it.pref[q],
        s_it.pref[q + 1] - s_it.pref[q]);
    __syncthreads();
  }
}

// ||Z||^2 of the slice's model (time Gram of its one row) -> c->inner[1]
__global__ void __launch_bounds__(64)
kb_cpd_norm(BSlot const * __restrict__ S, BHead const * __restrict__ hd)
{
  uint32_t const q = blockIdx.x;
  if(q >= hd->K) return;
  BSlot const & s = S[q];
  GramSrc g = {};
  for(int m=0; m < hd->N; ++m) g.aTa[m] = (m == hd->t) ? s.tgram : s.aTa[m];
  b_cpd_norm(g, hd->N, s.R, s.lam1, &s.c->inner[1]);
}

namespace {
uint32_t bt_grid(BSlot const * hS, uint32_t K, uint32_t (*nbf)(uint64_t), int mode, int what)
{
  uint64_t tot = 0;
  for(uint32_t q=0; q < K; ++q) {
    uint64_t const I = (what == 2) ? 1 : hS[q].rows[mode];
    tot += nbf(I);
  }
  if(tot < 1) tot = 1;
  return (uint32_t) (tot < BT_GRID ? tot : BT_GRID);
}
uint32_t bn_gram_h(uint64_t I) { return bn_gram(I); }
uint32_t bn_rows64_h(uint64_t I) { return bn_rows64(I); }
uint32_t bn_apply_h(uint64_t I) { return bn_apply(I); }
uint32_t bn_admm_h(uint64_t I) { return bn_admm(I); }
}

int dnb_gram(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, int what, int k, int gate, hipStream_t st)
{
  uint32_t const grid = bt_grid(hS, K, bn_gram_h, mode, what);
  bool const cross = (what == 1);
  switch(TT * 2 + (cross ? 1 : 0)) {
  case 2: hipLaunchKernelGGL((kb_gram<1, false>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 3: hipLaunchKernelGGL((kb_gram<1, true>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 4: hipLaunchKernelGGL((kb_gram<2, false>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 5: hipLaunchKernelGGL((kb_gram<2, true>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 6: hipLaunchKernelGGL((kb_gram<3, false>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 7: hipLaunchKernelGGL((kb_gram<3, true>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 8: hipLaunchKernelGGL((kb_gram<4, false>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  default: hipLaunchKernelGGL((kb_gram<4, true>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  }
  SPD_LAUNCH_CHECK();
  uint32_t const rg = K * ((TT * 16 * TT * 16 + 31) / 32);
  hipLaunchKernelGGL(kb_gram_reduce, dim3(rg < BT_GRID ? rg : BT_GRID), dim3(GR_RED_THREADS), 0,
      st, dS, dH, mode, what, k, gate);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dnb_colnorm_div(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st)
{
  hipLaunchKernelGGL(kb_coldot, dim3(bt_grid(hS, K, bn_rows64_h, mode, 0)), dim3(DN_THREADS), 0,
      st, dS, dH, mode, 0, BG_FIRST);
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL(kb_colreduce, dim3(K), dim3(DN_THREADS), 0, st, dS, dH, mode, 0, BG_FIRST);
  SPD_LAUNCH_CHECK();
  uint64_t tot = 0;
  for(uint32_t q=0; q < K; ++q) tot += bn_elem(hS[q].rows[mode] * hS[q].R);
  uint32_t const grid = (uint32_t) (tot < BT_GRID ? (tot ? tot : 1) : BT_GRID);
  hipLaunchKernelGGL(kb_div_cols, dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, BG_FIRST);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dnb_innerprod(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, int mode,
    hipStream_t st)
{
  hipLaunchKernelGGL(kb_coldot, dim3(bt_grid(hS, K, bn_rows64_h, mode, 0)), dim3(DN_THREADS), 0,
      st, dS, dH, mode, 1, BG_SMALL);
  SPD_LAUNCH_CHECK();
  hipLaunchKernelGGL(kb_colreduce, dim3(K), dim3(DN_THREADS), 0, st, dS, dH, mode, 1, BG_SMALL);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dnb_gj_inv(uint32_t K, BSlot const * dS, BHead const * dH, uint32_t TT, int mode, int gate,
    hipStream_t st)
{
  switch(TT) {
  case 1: hipLaunchKernelGGL((kb_gj_inv<16, 64>), dim3(K), dim3(64), 0, st, dS, dH, mode, gate); break;
  case 2: hipLaunchKernelGGL((kb_gj_inv<32, 128>), dim3(K), dim3(128), 0, st, dS, dH, mode, gate); break;
  case 3: hipLaunchKernelGGL((kb_gj_inv<48, 128>), dim3(K), dim3(128), 0, st, dS, dH, mode, gate); break;
  default: hipLaunchKernelGGL((kb_gj_inv<64, 256>), dim3(K), dim3(256), 0, st, dS, dH, mode, gate); break;
  }
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

template<int KIND>
static int dnb_apply_k(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K,
    uint32_t TT, int mode, int what, int gate, hipStream_t st)
{
  uint32_t const grid = bt_grid(hS, K, bn_apply_h, mode, what == 0 ? 2 : 0);
  switch(TT) {
  case 1: hipLaunchKernelGGL((kb_apply<1, KIND>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 2: hipLaunchKernelGGL((kb_apply<2, KIND>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  case 3: hipLaunchKernelGGL((kb_apply<3, KIND>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  default: hipLaunchKernelGGL((kb_apply<4, KIND>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, what, gate); break;
  }
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

int dnb_apply(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, int what, int gate, hipStream_t st)
{
  if(what == 2) return dnb_apply_k<2>(hS, dS, dH, K, TT, mode, what, gate, st);
  return dnb_apply_k<0>(hS, dS, dH, K, TT, mode, what, gate, st);
}

template<int TT>
static int dnb_admm_t(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K,
    int mode, uint32_t max_inner, hipStream_t st)
{
  constexpr int E = TT * 16 + 4;
  uint32_t const grid = bt_grid(hS, K, bn_admm_h, mode, 0);
  for(uint32_t j=1; j <= max_inner; ++j) {
    if(j == 1) {
      hipLaunchKernelGGL((kb_admm_pass<TT, 0>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, (int) j);
    } else if(j == 2) {
      hipLaunchKernelGGL((kb_admm_pass<TT, 1>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, (int) j);
    } else {
      hipLaunchKernelGGL((kb_admm_pass<TT, 2>), dim3(grid), dim3(DN_THREADS), 0, st, dS, dH, mode, (int) j);
    }
    SPD_LAUNCH_CHECK();
    hipLaunchKernelGGL((kb_admm_reduce<E>), dim3(K), dim3(RD_THREADS), 0, st, dS, dH, mode, (int) j);
    SPD_LAUNCH_CHECK();
  }
  uint64_t tot = 0;
  for(uint32_t q=0; q < K; ++q) tot += bn_elem(hS[q].rows[mode] * hS[q].R);
  uint32_t const g2 = (uint32_t) (tot < BT_GRID ? (tot ? tot : 1) : BT_GRID);
  hipLaunchKernelGGL(kb_admm_final, dim3(g2), dim3(DN_THREADS), 0, st, dS, dH, mode);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// the ADMM inner loop of `mode` for every live slice whose mode is solved
// by ADMM (its ctl was reset and its inverse with rho formed before)
int dnb_admm(BSlot const * hS, BSlot const * dS, BHead const * dH, uint32_t K, uint32_t TT,
    int mode, uint32_t max_inner, hipStream_t st)
{
  if(max_inner == 0) return SPD_OK;
  switch(TT) {
  case 1: return dnb_admm_t<1>(hS, dS, dH, K, mode, max_inner, st);
  case 2: return dnb_admm_t<2>(hS, dS, dH, K, mode, max_inner, st);
  case 3: return dnb_admm_t<3>(hS, dS, dH, K, mode, max_inner, st);
  default: return dnb_admm_t<4>(hS, dS, dH, K, mode, max_inner, st);
  }
}

int dnb_cpd_norm(uint32_t K, BSlot const * dS, BHead const * dH, hipStream_t st)
{
  hipLaunchKernelGGL(kb_cpd_norm, dim3(K), dim3(64), 0, st, dS, dH);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}
