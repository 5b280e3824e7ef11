/*
 * oracle_mttkrp.c -- TEST INFRASTRUCTURE (the oracle; see oracle.h).
 * MTTKRP restated on the CPU: brute force over the COO (long double) and the
 * CSF root / internal / leaf traversals with OpenMP threads, a 1024-entry
 * lock pool and privatised outputs for short modes.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- brute force (src/mttkrp.c:1699-1769, accumulation widened) ------- */

#define OR_DEF_COO(NAME, T)                                                   \
void NAME(or_coo const * tt, T ** mats, uint64_t rank, uint64_t mode,        \
    T * out)                                                                  \
{                                                                             \
  uint64_t const I = tt->dims[mode];                                          \
  long double * acc = calloc((I * rank) > 0 ? I * rank : 1, sizeof(long double)); \
  long double * row = malloc((rank ? rank : 1) * sizeof(long double));        \
  for(uint64_t n=0; n < tt->nnz; ++n) {                                       \
    for(uint64_t r=0; r < rank; ++r) {                                        \
      row[r] = (long double) tt->vals[n];                                     \
    }                                                                         \
    for(uint64_t m=0; m < tt->nmodes; ++m) {                                  \
      if(m == mode) continue;                                                 \
      T const * in = mats[m] + tt->ind[m][n] * rank;                          \
      for(uint64_t r=0; r < rank; ++r) {                                      \
        row[r] *= (long double) in[r];                                        \
      }                                                                       \
    }                                                                         \
    long double * o = acc + tt->ind[mode][n] * rank;                          \
    for(uint64_t r=0; r < rank; ++r) {                                        \
      o[r] += row[r];                                                         \
    }                                                                         \
  }                                                                           \
  for(uint64_t x=0; x < I * rank; ++x) {                                      \
    out[x] = (T) acc[x];                                                      \
  }                                                                           \
  free(row);                                                                  \
  free(acc);                                                                  \
}

OR_DEF_COO(or_mttkrp_coo, double)
OR_DEF_COO(or_mttkrp_coo_f, float)

/* ---- CSF traversals ---------------------------------------------------- */

#define OR_NLOCKS 1024
#define OR_LOCK_PAD 16

#ifdef _OPENMP
static omp_lock_t g_locks[OR_NLOCKS * OR_LOCK_PAD];
static int g_locks_init = 0;
static void p_locks_init(void)
{
  if(!g_locks_init) {
    for(int i=0; i < OR_NLOCKS; ++i) {
      omp_init_lock(g_locks + i * OR_LOCK_PAD);
    }
    g_locks_init = 1;
  }
}
#define OR_LOCK(row) omp_set_lock(g_locks + ((row) % OR_NLOCKS) * OR_LOCK_PAD)
#define OR_UNLOCK(row) omp_unset_lock(g_locks + ((row) % OR_NLOCKS) * OR_LOCK_PAD)
#else
static void p_locks_init(void) {}
#define OR_LOCK(row) ((void) 0)
#define OR_UNLOCK(row) ((void) 0)
#endif

/*
 * One generic traversal per precision. For the nodes under root slice s of
 * tile pt:
 *   sub(l, x) = sum over children c of x: U_{l+1}[fid(c)] o sub(l+1, c),
 *   sub(N-2, x) = sum over nonzeros j of x: v_j U_{N-1}[fid(j)],
 * and with pre(l) = U_0[root] o ... o U_l[node at l] along the path:
 *   root output:     out[root]    += sub(0, s)                (:390-732)
 *   internal d:      out[fid(x)]  += pre(d-1) o sub(d, x)     (:1032-1278)
 *   leaf output:     out[fid(j)]  += v_j pre(N-2)             (:802-1029)
 * Root ids are fids[0][s] when stored, else the slice position s (:428).
 */
#define OR_DEF_CSF(NAME, T)                                                   \
static void NAME##_sub(or_tile const * pt, uint64_t N, T ** um,              \
    uint64_t rank, uint64_t l, uint64_t x, T * acc, T * scratch)             \
{                                                                             \
  for(uint64_t r=0; r < rank; ++r) acc[r] = 0;                                \
  uint64_t const a = pt->fptr[l][x];                                          \
  uint64_t const b = pt->fptr[l][x+1];                                        \
  if(l + 2 == N) {                                                            \
    for(uint64_t j=a; j < b; ++j) {                                           \
      T const v = (T) pt->vals[j];                                            \
      T const * row = um[N-1] + pt->fids[N-1][j] * rank;                      \
      for(uint64_t r=0; r < rank; ++r) acc[r] += v * row[r];                  \
    }                                                                         \
    return;                                                                   \
  }                                                                           \
  T * child = scratch;                                                        \
  for(uint64_t c=a; c < b; ++c) {                                             \
    NAME##_sub(pt, N, um, rank, l+1, c, child, scratch + rank);               \
    T const * row = um[l+1] + pt->fids[l+1][c] * rank;                        \
    for(uint64_t r=0; r < rank; ++r) acc[r] += child[r] * row[r];             \
  }                                                                           \
}                                                                             \
                                                                              \
static void NAME##_down(or_tile const * pt, uint64_t N, T ** um,             \
    uint64_t rank, uint64_t depth, uint64_t l, uint64_t x, T const * pre,     \
    T * out, int locked, T * scratch)                                         \
{                                                                             \
  /* pre = product of rows for levels < l+? ; x is a node at level l */      \
  uint64_t const a = pt->fptr[l][x];                                          \
  uint64_t const b = pt->fptr[l][x+1];                                        \
  T * mine = scratch;                                                         \
  if(l + 1 == depth) {                                                        \
    if(depth == N - 1) {                                                      \
      for(uint64_t j=a; j < b; ++j) {                                         \
        T const v = (T) pt->vals[j];                                          \
        uint64_t const row = pt->fids[N-1][j];                                \
        T * o = out + row * rank;                                             \
        if(locked) OR_LOCK(row);                                              \
        for(uint64_t r=0; r < rank; ++r) o[r] += v * pre[r];                  \
        if(locked) OR_UNLOCK(row);                                            \
      }                                                                       \
    } else {                                                                  \
      for(uint64_t c=a; c < b; ++c) {                                         \
        NAME##_sub(pt, N, um, rank, depth, c, mine, scratch + rank);          \
        uint64_t const row = pt->fids[depth][c];                              \
        T * o = out + row * rank;                                             \
        if(locked) OR_LOCK(row);                                              \
        for(uint64_t r=0; r < rank; ++r) o[r] += pre[r] * mine[r];            \
        if(locked) OR_UNLOCK(row);                                            \
      }                                                                       \
    }                                                                         \
    return;                                                                   \
  }                                                                           \
  for(uint64_t c=a; c < b; ++c) {                                             \
    T const * row = um[l+1] + pt->fids[l+1][c] * rank;                        \
    for(uint64_t r=0; r < rank; ++r) mine[r] = pre[r] * row[r];               \
    NAME##_down(pt, N, um, rank, depth, l+1, c, mine, out, locked,            \
        scratch + rank);                                                      \
  }                                                                           \
}                                                                             \
                                                                              \
void NAME(or_csf const * ct, T ** mats, uint64_t rank, uint64_t mode,        \
    T * out, int nthreads)                                                    \
{                                                                             \
  uint64_t const N = ct->nmodes;                                              \
  uint64_t const depth = ct->dim_iperm[mode];                                 \
  uint64_t const I = ct->dims[mode];                                          \
  T * um[OR_MAX_NMODES];                                                      \
  for(uint64_t l=0; l < N; ++l) um[l] = mats[ct->dim_perm[l]];                \
  memset(out, 0, I * rank * sizeof(T));                                       \
  p_locks_init();                                                             \
  if(nthreads < 1) nthreads = 1;                                              \
  /* privatise short modes, src/mttkrp.c:221-236 with PRIVTHRESH 0.02 */     \
  int const priv = (nthreads > 1) && (depth > 0) &&                           \
      ((double) (I * (uint64_t) nthreads) <= 0.02 * (double) ct->nnz);        \
  T * pbuf = NULL;                                                            \
  if(priv) pbuf = calloc((size_t) nthreads * I * rank, sizeof(T));            \
  int const root_unique = (ct->ntiles == 1);                                  \
  uint64_t nwork = 0;                                                         \
  uint64_t * tstart = malloc((ct->ntiles + 1) * sizeof(uint64_t));            \
  for(uint64_t t=0; t < ct->ntiles; ++t) {                                    \
    tstart[t] = nwork;                                                        \
    nwork += (ct->pt[t].vals == NULL) ? 0 : ct->pt[t].nfibs[0];               \
  }                                                                           \
  tstart[ct->ntiles] = nwork;                                                 \
  _Pragma("omp parallel num_threads(nthreads)")                               \
  {                                                                           \
    int tid = 0;                                                              \
    OR_TID(tid);                                                              \
    T * scratch = malloc((size_t) (N + 2) * 2 * rank * sizeof(T) + 64);       \
    T * acc = malloc((size_t) rank * sizeof(T) + 64);                         \
    T * myout = priv ? pbuf + (size_t) tid * I * rank : out;                  \
    int const locked = !priv && !(depth == 0 && root_unique);                 \
    _Pragma("omp for schedule(dynamic, 16)")                                  \
    for(uint64_t w=0; w < nwork; ++w) {                                       \
      uint64_t t = 0;                                                         \
      while(tstart[t+1] <= w) ++t;                                            \
      or_tile const * pt = ct->pt + t;                                        \
      uint64_t const s = w - tstart[t];                                       \
      uint64_t const root = pt->fids[0] ? pt->fids[0][s] : s;                 \
      if(depth == 0) {                                                        \
        NAME##_sub(pt, N, um, rank, 0, s, acc, scratch);                      \
        T * o = myout + root * rank;                                          \
        if(locked) OR_LOCK(root);                                             \
        for(uint64_t r=0; r < rank; ++r) o[r] += acc[r];                      \
        if(locked) OR_UNLOCK(root);                                           \
      } else {                                                                \
        T const * rrow = um[0] + root * rank;                                 \
        for(uint64_t r=0; r < rank; ++r) acc[r] = rrow[r];                    \
        NAME##_down(pt, N, um, rank, depth, 0, s, acc, myout, locked,         \
            scratch);                                                         \
      }                                                                       \
    }                                                                         \
    free(scratch);                                                            \
    free(acc);                                                                \
  }                                                                           \
  if(priv) {                                                                  \
    _Pragma("omp parallel for num_threads(nthreads) schedule(static)")        \
    for(uint64_t x=0; x < I * rank; ++x) {                                    \
      T s = 0;                                                                \
      for(int t=0; t < nthreads; ++t) s += pbuf[(size_t) t * I * rank + x];   \
      out[x] = s;                                                             \
    }                                                                         \
    free(pbuf);                                                               \
  }                                                                           \
  free(tstart);                                                               \
}

#ifdef _OPENMP
#define OR_TID(v) ((v) = omp_get_thread_num())
#else
#define OR_TID(v) ((void) 0)
#endif

OR_DEF_CSF(or_mttkrp_csf, double)
OR_DEF_CSF(or_mttkrp_csf_f, float)
