// spd_dense.hpp -- dense FP64 factor algebra used by CPD-ALS and streaming
// CPD: Grams, the normal-equation solve, column norms, ADMM steps and the
// fit reductions. All matrices are row-major I x R in device memory; all
// results stay on the device (nothing here synchronises).
#pragma once
#include "spd_internal.hpp"

// Proximal operators recognised on the device (the reference's
// constraints/maxcolnorm.c, maxcolnorm-nonneg.c, nonneg.c).
#define DN_PROX_NONE 0
#define DN_PROX_MAXCOLNORM 1
#define DN_PROX_MAXCOLNORM_NONNEG 2
#define DN_PROX_NONNEG 3
#define DN_PROX_LASSO 4
#define DN_PROX_ROWSIMP 5      // chunked ADMM only
#define DN_PROX_SMOOTH 6       // one-block ADMM, one step at a time
#define DN_PROX_COLSIMP 7      // one-block ADMM, one step at a time
#define DN_PROX_NTF_FROB 8     // max(p * rho / (lam + rho), 0)
#define DN_PROX_NTF_LASSO 9    // max(p - lam / rho, 0)

// Entry-wise operators that take one scalar of lam and rho: the lasso
// thresholds (lam / rho) and the non-negative frob scale (rho / (lam + rho)).
__host__ __device__ __forceinline__ bool dn_prox_scalar_kind(int prox)
{
  return prox == DN_PROX_LASSO || prox == DN_PROX_NTF_LASSO || prox == DN_PROX_NTF_FROB;
}
__host__ __device__ __forceinline__ double dn_prox_scalar(int prox, double lam, double rho)
{
  return (prox == DN_PROX_NTF_FROB) ? rho / (lam + rho) : lam / rho;
}
// Entry-wise operators that need nothing but the entry and that scalar.
__host__ __device__ __forceinline__ double dn_prox_entry(double p, int prox, double thr)
{
  if(prox == DN_PROX_NONNEG) {
    p = (p > 0.) ? p : 0.;
  } else if(prox == DN_PROX_LASSO) {
    p = (p > thr) ? p - thr : ((p < -thr) ? p + thr : 0.);
  } else if(prox == DN_PROX_NTF_LASSO) {
    p = (p > thr) ? p - thr : 0.;
  } else if(prox == DN_PROX_NTF_FROB) {
    p = p * thr;
    p = (p > 0.) ? p : 0.;
  }
  return p;
}

// out (R x R, full, symmetric) = X^T X. When div != nullptr the columns of X
// are first divided in place by div[r] (X[:,r] /= div[r], the reference's
// normalisation); when X32 != nullptr an FP32 copy of the (divided) X is
// written as well.
int dn_gram(double * X, uint64_t I, uint32_t R, double const * div, float * X32,
    double * out, hipStream_t st);

// Normal-equation matrix of `mode` and its inverse: G = Hadamard product of
// aTa[m] over m != mode (upper triangle, mirrored), plus diag_add on the
// diagonal; with admm != 0 also rho = trace(G)/R is stored in *rho and added
// to the diagonal. G is Cholesky-factored (G = L L^T) and
// ginv (RP x RP, RP = R rounded up to 16, zero padded) = L^{-T} L^{-1}.
// *info gets the 1-based column of a non-positive pivot (0 if none); it must
// be zero on entry.
int dn_chol_inv(double const * const * aTa, int nmodes, int mode, uint32_t R,
    double diag_add, int admm, double * ginv, double * rho, int * info,
    hipStream_t st);

// Row-padded size used for ginv.
__host__ __device__ static inline uint32_t dn_rpad(uint32_t R) { return (R + 15) / 16 * 16; }

// X = M ginv (rows of M times the inverse, i.e. the Cholesky solve of the
// reference's potrs). M is FP64 (mdtype SPD_F64) or FP32. When colnorm is
// given, colnorm[r] = ||X[:,r]||_2.
int dn_apply(void const * M, int mdtype, double const * ginv, uint64_t I,
    uint32_t R, double * X, double * colnorm, hipStream_t st);

// One ADMM inner iteration (src/cpd/admm.c:240-308) on device:
//   init = P; aux = (M + rho (P + D)) ginv; P = prox(aux - D); D += P - aux;
// sums[0..3] = ||P||^2, ||P - aux||^2, ||P - init||^2, ||D||^2 (after the
// update); sums[4 .. 4+R) receive the column norms the prox used, so sums
// must hold 4 + R doubles. lam: the lasso multiplier (threshold lam / rho).
int dn_admm_step(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint32_t R, double * P, double * aux,
    double * D, double * init, int prox, double lam, double * sums, hipStream_t st);

// The whole ADMM inner loop (up to max_inner iterations, src/cpd/admm.c:
// 240-308): P and D are updated in place; sums (4 + R doubles) ends with the
// last iteration's residual terms and column norms; ctl (2 ints, device)
// gets ctl[0] = converged, ctl[1] = iterations run. For R <= 64 the test
// runs on the device and nothing synchronises; aux/init are scratch (I x R)
// used only above rank 64.
int dn_admm_solve(void const * M, int mdtype, double const * ginv, double const * rho,
    uint64_t I, uint32_t R, double * P, double * D, double * aux, double * init, int prox,
    double lam, uint64_t max_inner, double tol, double * sums, int * ctl, hipStream_t st,
    int d_zero = 0);

// ADMM in independent chunks of `chunk` rows for a row-separable prox
// (DN_PROX_NONNEG, DN_PROX_LASSO, DN_PROX_ROWSIMP or none; src/cpd/admm.c:
// 355-400): each chunk iterates to
// its own convergence, up to max_inner times, one block per chunk, no host
// synchronisation. *its_rows (device, zeroed by the caller) accumulates
// iterations x rows per chunk. Ranks up to 3072.
int dn_admm_chunked(void const * M, int mdtype, double const * ginv, double const * rho,
    uint64_t I, uint32_t R, uint64_t chunk, double * P, double * D, int prox, double lam,
    uint64_t max_inner, double tol, unsigned long long * its_rows, hipStream_t st);

// out (R x R, full) = X^T Y for two I x R matrices.
int dn_xgram(double const * X, double const * Y, uint64_t I, uint32_t R,
    double * out, hipStream_t st);

// out += In G for In (I x R) and a general G (RP x RP, zero padded, RP =
// R rounded up to 16).
int dn_apply_acc(double const * In, double const * G, uint64_t I, uint32_t R,
    double * out, hipStream_t st);

// lam[r] = ||X[:,r]||_2.
int dn_colnorm(double const * X, uint64_t I, uint32_t R, double * lam,
    hipStream_t st);

// X[:,r] /= div[r] (and the FP32 copy when X32 != nullptr).
int dn_div_cols(double * X, uint64_t I, uint32_t R, double const * div,
    float * X32, hipStream_t st);

// *out = |lambda^T (Hadamard of all aTa) lambda| over the upper triangle
// (src/cpd/cpd.c:453-487).
int dn_cpd_norm(double const * const * aTa, int nmodes, uint32_t R,
    double const * lambda, double * out, hipStream_t st);

// *out = sum_r lambda[r] sum_i X[i,r] M[i,r] (src/cpd/cpd.c:490-531).
int dn_innerprod(double const * X, void const * M, int mdtype, uint64_t I,
    uint32_t R, double const * lambda, double * out, hipStream_t st);
