/*
 * oracle_csf.c -- TEST INFRASTRUCTURE (the oracle; see oracle.h).
 * CSF construction restated on the CPU: mode ordering, canonical
 * lexicographic sort, dense tiling, fiber compression.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* ---- mode ordering, src/csf.c:93-218 and :608-651 ---------------------- */

/* Stable sort of modes by dim: ascending for small-first (src/csf.c:93-117),
 * descending for big-first (:185-218). Among equal dims the lower mode index
 * comes first in both cases, which is what the reference's greedy
 * "first unmatched mode with this dim" matching produces. */
static void p_order_by_dim(uint64_t const * dims, uint64_t nmodes, int desc,
    uint64_t * perm)
{
  for(uint64_t m=0; m < nmodes; ++m) {
    perm[m] = m;
  }
  for(uint64_t i=1; i < nmodes; ++i) {
    uint64_t v = perm[i];
    uint64_t j = i;
    while(j > 0) {
      uint64_t const a = dims[perm[j-1]];
      uint64_t const b = dims[v];
      int const before = desc ? (a < b) : (a > b);
      if(!before) {
        break;
      }
      perm[j] = perm[j-1];
      --j;
    }
    perm[j] = v;
  }
}

static void p_move_front(uint64_t * perm, uint64_t nmodes, uint64_t mode)
{
  for(uint64_t m=0; m < nmodes; ++m) {
    if(perm[m] == mode) {
      memmove(perm + 1, perm, m * sizeof(*perm));
      perm[0] = mode;
      return;
    }
  }
}

void or_mode_order(uint64_t const * dims, uint64_t nmodes, int type,
    uint64_t mode, uint64_t * perm)
{
  switch(type) {
  case 0:
    p_order_by_dim(dims, nmodes, 0, perm);
    break;
  case 1:
    p_order_by_dim(dims, nmodes, 1, perm);
    break;
  case 2:
    for(uint64_t m=0; m < nmodes; ++m) {
      perm[m] = m;
    }
    p_move_front(perm, nmodes, mode);
    break;
  case 3:
    p_order_by_dim(dims, nmodes, 0, perm);
    p_move_front(perm, nmodes, mode);
    break;
  default:
    break;
  }
}

/* ---- canonical order --------------------------------------------------- */

/* LSD radix sort of (key, payload) pairs, stable, 8-bit digits. */
static void p_radix_sort(uint64_t * key, uint64_t * pay, uint64_t n, int bits)
{
  uint64_t * k2 = malloc((n ? n : 1) * sizeof(uint64_t));
  uint64_t * p2 = malloc((n ? n : 1) * sizeof(uint64_t));
  for(int shift=0; shift < bits; shift += 8) {
    uint64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for(uint64_t i=0; i < n; ++i) {
      ++cnt[1 + ((key[i] >> shift) & 255)];
    }
    for(int d=0; d < 256; ++d) {
      cnt[d+1] += cnt[d];
    }
    for(uint64_t i=0; i < n; ++i) {
      uint64_t const d = (key[i] >> shift) & 255;
      uint64_t const o = cnt[d]++;
      k2[o] = key[i];
      p2[o] = pay[i];
    }
    memcpy(key, k2, n * sizeof(uint64_t));
    memcpy(pay, p2, n * sizeof(uint64_t));
  }
  free(k2);
  free(p2);
}

static or_coo const * g_cmp_tt;
static uint64_t const * g_cmp_perm;

static int p_cmp_lex(void const * a, void const * b)
{
  uint64_t const i = *(uint64_t const *) a;
  uint64_t const j = *(uint64_t const *) b;
  for(uint64_t l=0; l < g_cmp_tt->nmodes; ++l) {
    uint64_t const * ind = g_cmp_tt->ind[g_cmp_perm[l]];
    if(ind[i] < ind[j]) return -1;
    if(ind[i] > ind[j]) return 1;
  }
  return (i < j) ? -1 : (i > j);
}

/* Permutation putting the nonzeros in lexicographic order of their
 * coordinates taken in perm order (src/sort.c:761-905 yields this order for
 * tensors without duplicate coordinates). Ties keep input order. */
static uint64_t * p_lex_order(or_coo const * tt, uint64_t const * perm)
{
  uint64_t const n = tt->nnz;
  uint64_t * idx = malloc((n ? n : 1) * sizeof(uint64_t));
  for(uint64_t i=0; i < n; ++i) {
    idx[i] = i;
  }
  /* mixed-radix key if the product of the dims fits in 64 bits */
  unsigned __int128 prod = 1;
  for(uint64_t l=0; l < tt->nmodes; ++l) {
    prod *= (unsigned __int128) (tt->dims[perm[l]] ? tt->dims[perm[l]] : 1);
  }
  if(prod <= ((unsigned __int128) 1 << 64) - 1) {
    int bits = 0;
    while(bits < 64 && (((unsigned __int128) 1) << bits) < prod) {
      ++bits;
    }
    uint64_t * key = malloc((n ? n : 1) * sizeof(uint64_t));
    for(uint64_t i=0; i < n; ++i) {
      uint64_t k = 0;
      for(uint64_t l=0; l < tt->nmodes; ++l) {
        k = k * tt->dims[perm[l]] + tt->ind[perm[l]][i];
      }
      key[i] = k;
    }
    p_radix_sort(key, idx, n, bits);
    free(key);
  } else {
    g_cmp_tt = tt;
    g_cmp_perm = perm;
    qsort(idx, n, sizeof(uint64_t), p_cmp_lex);
  }
  return idx;
}

/* Sort the COO in place into the lexicographic order of perm (the order
 * tt_sort produces, src/sort.c:912-956). */
void or_coo_sort(or_coo * tt, uint64_t const * perm)
{
  uint64_t const n = tt->nnz;
  uint64_t * idx = p_lex_order(tt, perm);
  uint64_t * tmp = malloc((n ? n : 1) * sizeof(uint64_t));
  for(uint64_t m=0; m < tt->nmodes; ++m) {
    for(uint64_t i=0; i < n; ++i) tmp[i] = tt->ind[m][idx[i]];
    memcpy(tt->ind[m], tmp, n * sizeof(uint64_t));
  }
  double * tv = malloc((n ? n : 1) * sizeof(double));
  for(uint64_t i=0; i < n; ++i) tv[i] = tt->vals[idx[i]];
  memcpy(tt->vals, tv, n * sizeof(double));
  free(tv);
  free(tmp);
  free(idx);
}

/* ---- compression, src/csf.c:230-372 ------------------------------------ */

/* Build the sparsity of nonzeros [0, n) of the already-ordered arrays
 * lind[l] (index of level l) and vals. A node starts at level l where the
 * level-l index changes or a node starts at level l-1. */
static void p_compress(or_tile * pt, uint64_t nmodes, uint64_t * const * lind,
    double const * vals, uint64_t n, int store_root_ids)
{
  uint64_t const N = nmodes;
  /* level of first difference with the previous nonzero */
  unsigned char * lev = malloc(n ? n : 1);
  for(uint64_t i=0; i < n; ++i) {
    unsigned char d = (unsigned char) (N - 1);
    if(i == 0) {
      d = 0;
    } else {
      for(uint64_t l=0; l + 1 < N; ++l) {
        if(lind[l][i] != lind[l][i-1]) {
          d = (unsigned char) l;
          break;
        }
      }
    }
    lev[i] = d;
  }
  for(uint64_t l=0; l + 1 < N; ++l) {
    uint64_t cnt = 0;
    for(uint64_t i=0; i < n; ++i) {
      cnt += (lev[i] <= l);
    }
    pt->nfibs[l] = cnt;
    pt->fptr[l] = malloc((cnt + 1) * sizeof(uint64_t));
    pt->fids[l] = (l == 0 && !store_root_ids) ? NULL :
        malloc((cnt ? cnt : 1) * sizeof(uint64_t));
  }
  pt->nfibs[N-1] = n;
  pt->fids[N-1] = malloc((n ? n : 1) * sizeof(uint64_t));
  pt->fptr[N-1] = NULL;
  pt->vals = malloc((n ? n : 1) * sizeof(double));
  memcpy(pt->fids[N-1], lind[N-1], n * sizeof(uint64_t));
  memcpy(pt->vals, vals, n * sizeof(double));

  uint64_t node[OR_MAX_NMODES];
  for(uint64_t l=0; l < N; ++l) {
    node[l] = 0;
  }
  for(uint64_t i=0; i < n; ++i) {
    for(uint64_t l=0; l + 1 < N; ++l) {
      if(lev[i] <= l) {
        uint64_t const f = node[l]++;
        if(pt->fids[l] != NULL) {
          pt->fids[l][f] = lind[l][i];
        }
        /* the child started at this nonzero has id node[l+1] (before its
         * own increment), or i at the leaf level */
        pt->fptr[l][f] = (l + 2 == N) ? i : node[l+1];
      }
    }
  }
  for(uint64_t l=0; l + 1 < N; ++l) {
    pt->fptr[l][pt->nfibs[l]] = pt->nfibs[l+1];
  }
  free(lev);
}

/* get_tile_id, src/tile.c:326-342: last mode least significant. */
static uint64_t p_tile_id(uint64_t const * tile_dims, uint64_t nmodes,
    uint64_t const * coord)
{
  uint64_t id = 0;
  uint64_t mult = 1;
  for(uint64_t m=nmodes; m-- != 0; ) {
    id += coord[m] * mult;
    mult *= tile_dims[m];
  }
  return id;
}

or_csf * or_csf_build(or_coo * tt, uint64_t const * perm, int which_tile,
    uint64_t tilelevel, uint64_t nthreads, int reorder_tt)
{
  uint64_t const N = tt->nmodes;
  uint64_t const nnz = tt->nnz;
  or_csf * ct = calloc(1, sizeof(*ct));
  ct->nnz = nnz;
  ct->nmodes = N;
  for(uint64_t m=0; m < N; ++m) {
    ct->dims[m] = tt->dims[m];
    ct->dim_perm[m] = perm[m];
  }
  for(uint64_t l=0; l < N; ++l) {
    ct->dim_iperm[perm[l]] = l;
  }
  ct->which_tile = which_tile;

  uint64_t * order = p_lex_order(tt, perm);

  /* tiling (src/csf.c:427-451, src/tile.c:260-322) */
  uint64_t ntiles = 1;
  if(which_tile == 1) {
    ct->ntiled_modes = tilelevel < N ? tilelevel : N;
    uint64_t const tile_depth = N - ct->ntiled_modes;
    for(uint64_t m=0; m < N; ++m) {
      ct->tile_dims[m] = (ct->dim_iperm[m] >= tile_depth) ? nthreads : 1;
      ntiles *= ct->tile_dims[m];
    }
  } else {
    ct->ntiled_modes = 0;
    for(uint64_t m=0; m < N; ++m) {
      ct->tile_dims[m] = 1;
    }
  }
  ct->ntiles = ntiles;
  uint64_t * tptr = calloc(ntiles + 1, sizeof(uint64_t));
  if(ntiles > 1) {
    uint64_t tsizes[OR_MAX_NMODES];
    for(uint64_t m=0; m < N; ++m) {
      tsizes[m] = tt->dims[m] / ct->tile_dims[m];
      if(tsizes[m] < 1) tsizes[m] = 1;
    }
    uint64_t * tid = malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    for(uint64_t i=0; i < nnz; ++i) {
      uint64_t coord[OR_MAX_NMODES];
      uint64_t const x = order[i];
      for(uint64_t m=0; m < N; ++m) {
        uint64_t c = tt->ind[m][x] / tsizes[m];
        coord[m] = c < ct->tile_dims[m] - 1 ? c : ct->tile_dims[m] - 1;
      }
      tid[i] = p_tile_id(ct->tile_dims, N, coord);
      ++tptr[tid[i] + 1];
    }
    for(uint64_t t=0; t < ntiles; ++t) {
      tptr[t+1] += tptr[t];
    }
    /* stable bucket by tile id */
    uint64_t * fill = malloc(ntiles * sizeof(uint64_t));
    memcpy(fill, tptr, ntiles * sizeof(uint64_t));
    uint64_t * order2 = malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    for(uint64_t i=0; i < nnz; ++i) {
      order2[fill[tid[i]]++] = order[i];
    }
    free(fill);
    free(tid);
    free(order);
    order = order2;
  } else {
    tptr[1] = nnz;
  }

  /* level-ordered copies */
  uint64_t * lind[OR_MAX_NMODES];
  for(uint64_t l=0; l < N; ++l) {
    lind[l] = malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    uint64_t const * src = tt->ind[perm[l]];
    for(uint64_t i=0; i < nnz; ++i) {
      lind[l][i] = src[order[i]];
    }
  }
  double * svals = malloc((nnz ? nnz : 1) * sizeof(double));
  for(uint64_t i=0; i < nnz; ++i) {
    svals[i] = tt->vals[order[i]];
  }

  ct->pt = calloc(ntiles, sizeof(or_tile));
  for(uint64_t t=0; t < ntiles; ++t) {
    or_tile * pt = ct->pt + t;
    uint64_t const a = tptr[t];
    uint64_t const b = tptr[t+1];
    if(a == b) {
      /* empty tile, src/csf.c:467-479 */
      pt->fptr[0] = calloc(2, sizeof(uint64_t));
      continue;
    }
    uint64_t * sub[OR_MAX_NMODES];
    for(uint64_t l=0; l < N; ++l) {
      sub[l] = lind[l] + a;
    }
    p_compress(pt, N, sub, svals + a, b - a, ntiles > 1);
  }

  if(reorder_tt) {
    for(uint64_t m=0; m < N; ++m) {
      memcpy(tt->ind[m], lind[ct->dim_iperm[m]], nnz * sizeof(uint64_t));
    }
    memcpy(tt->vals, svals, nnz * sizeof(double));
  }

  for(uint64_t l=0; l < N; ++l) {
    free(lind[l]);
  }
  free(svals);
  free(tptr);
  free(order);
  return ct;
}

void or_csf_free(or_csf * ct)
{
  if(ct == NULL) {
    return;
  }
  for(uint64_t t=0; t < ct->ntiles; ++t) {
    for(uint64_t l=0; l < ct->nmodes; ++l) {
      free(ct->pt[t].fptr[l]);
      free(ct->pt[t].fids[l]);
    }
    free(ct->pt[t].vals);
  }
  free(ct->pt);
  free(ct);
}

double or_csf_frobsq(or_csf const * ct)
{
  double norm = 0;
  for(uint64_t t=0; t < ct->ntiles; ++t) {
    if(ct->pt[t].vals == NULL) {
      continue;
    }
    uint64_t const n = ct->pt[t].nfibs[ct->nmodes-1];
    for(uint64_t i=0; i < n; ++i) {
      norm += ct->pt[t].vals[i] * ct->pt[t].vals[i];
    }
  }
  return norm;
}
