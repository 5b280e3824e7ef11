/*
 * gen.c -- synthetic sparse tensors for tests and benchmarks.
 *
 * Coordinates: per mode either uniform or Zipf(s) truncated at the dim
 * (inverse CDF on a table), with Zipf ranks mapped to indices through a
 * seeded Fisher-Yates permutation so heavy indices are scattered. Draws are
 * deduplicated (sort + unique on a mixed-radix key) and topped up until the
 * target count of distinct coordinates is reached; optionally every mode is
 * then compacted (indices relabelled in order so no slice is empty).
 * Values: uniform(0,1), lognormal(0,1), or uniform(-1,1).
 * Random streams are SplitMix64, one per purpose, from a single seed.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GEN_MAX_MODES 8

static inline uint64_t sm64(uint64_t * s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline double u01(uint64_t * s)
{
  return (double) (sm64(s) >> 11) * (1.0 / 9007199254740992.0);
}

/* LSD radix sort of 64-bit keys, 16-bit digits, in place via a buffer. */
static void sort_u64(uint64_t * a, uint64_t n, int bits)
{
  uint64_t * b = malloc((n ? n : 1) * sizeof(uint64_t));
  uint64_t * cnt = malloc(65537 * sizeof(uint64_t));
  uint64_t * src = a;
  uint64_t * dst = b;
  for(int sh=0; sh < bits; sh += 16) {
    memset(cnt, 0, 65537 * sizeof(uint64_t));
    for(uint64_t i=0; i < n; ++i) cnt[1 + ((src[i] >> sh) & 0xFFFF)]++;
    for(int d=0; d < 65536; ++d) cnt[d+1] += cnt[d];
    for(uint64_t i=0; i < n; ++i) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
    uint64_t * t = src; src = dst; dst = t;
  }
  if(src != a) memcpy(a, src, n * sizeof(uint64_t));
  free(cnt);
  free(b);
}

typedef struct {
  double * cdf;      /* zipf only */
  uint64_t * perm;   /* rank -> index */
} mode_tab;

static uint64_t draw(mode_tab const * t, uint64_t dim, double zs, uint64_t * s)
{
  if(zs <= 0) {
    return sm64(s) % dim;
  }
  double const u = u01(s);
  uint64_t lo = 0, hi = dim - 1;
  while(lo < hi) {
    uint64_t const mid = (lo + hi) >> 1;
    if(t->cdf[mid] < u) lo = mid + 1; else hi = mid;
  }
  return t->perm[lo];
}

/*
 * Generate nnz distinct coordinates.
 *   zipf_s[m] <= 0: uniform in mode m, else Zipf(zipf_s[m]).
 *   val_kind: 0 uniform(0,1), 1 lognormal(0,1), 2 uniform(-1,1).
 *   compact: relabel each mode so that no index is empty; dims_out gets the
 *            resulting dims.
 * inds[m] (nnz uint64 each) and vals (nnz) are caller-allocated. The
 * coordinates come out in key order (mode 0 major); callers that need a
 * shuffled order can permute. Returns 0, or -1 if nnz exceeds the number of
 * cells.
 */
int splatt_gen_coord(int nmodes, uint64_t const * dims, double const * zipf_s,
    uint64_t nnz, int val_kind, int compact, uint64_t seed, uint64_t ** inds,
    double * vals, uint64_t * dims_out)
{
  unsigned __int128 cells = 1;
  for(int m=0; m < nmodes; ++m) cells *= dims[m];
  if((unsigned __int128) nnz > cells || cells >= ((unsigned __int128) 1 << 64)) {
    return -1;
  }
  int bits = 0;
  while(bits < 64 && ((unsigned __int128) 1 << bits) < cells) ++bits;

  uint64_t s_coord = seed ^ 0x1234567890ABCDEFull;
  uint64_t s_perm = seed ^ 0x0F1E2D3C4B5A6978ull;
  uint64_t s_val = seed ^ 0x5555AAAA3333CCCCull;

  mode_tab tabs[GEN_MAX_MODES];
  memset(tabs, 0, sizeof(tabs));
  for(int m=0; m < nmodes; ++m) {
    if(zipf_s[m] > 0) {
      uint64_t const d = dims[m];
      tabs[m].cdf = malloc(d * sizeof(double));
      double acc = 0;
      for(uint64_t k=0; k < d; ++k) {
        acc += pow((double) (k + 1), -zipf_s[m]);
        tabs[m].cdf[k] = acc;
      }
      for(uint64_t k=0; k < d; ++k) tabs[m].cdf[k] /= acc;
      tabs[m].cdf[d-1] = 1.0;
      tabs[m].perm = malloc(d * sizeof(uint64_t));
      for(uint64_t k=0; k < d; ++k) tabs[m].perm[k] = k;
      for(uint64_t k=d; k-- > 1; ) {
        uint64_t const j = sm64(&s_perm) % (k + 1);
        uint64_t const t = tabs[m].perm[k];
        tabs[m].perm[k] = tabs[m].perm[j];
        tabs[m].perm[j] = t;
      }
    }
  }

  uint64_t cap = nnz + nnz / 4 + 1024;
  uint64_t * keys = malloc(cap * sizeof(uint64_t));
  uint64_t have = 0;
  while(have < nnz) {
    uint64_t want = (nnz - have) + (nnz - have) / 8 + 64;
    if(have + want > cap) {
      cap = have + want;
      keys = realloc(keys, cap * sizeof(uint64_t));
    }
    for(uint64_t i=0; i < want; ++i) {
      uint64_t k = 0;
      for(int m=0; m < nmodes; ++m) {
        k = k * dims[m] + draw(tabs + m, dims[m], zipf_s[m], &s_coord);
      }
      keys[have + i] = k;
    }
    uint64_t const tot = have + want;
    sort_u64(keys, tot, bits);
    uint64_t u = 0;
    for(uint64_t i=0; i < tot; ++i) {
      if(u == 0 || keys[i] != keys[u-1]) keys[u++] = keys[i];
    }
    have = u;
  }
  /* keep a uniformly chosen subset of exactly nnz keys when overshooting */
  if(have > nnz) {
    uint64_t const extra = have - nnz;
    uint64_t s_drop = seed ^ 0x7777000011112222ull;
    unsigned char * drop = calloc(have, 1);
    uint64_t dropped = 0;
    while(dropped < extra) {
      uint64_t const j = sm64(&s_drop) % have;
      if(!drop[j]) { drop[j] = 1; ++dropped; }
    }
    uint64_t w = 0;
    for(uint64_t i=0; i < have; ++i) if(!drop[i]) keys[w++] = keys[i];
    free(drop);
  }
  for(uint64_t i=nnz; i-- > 0; ) {
    uint64_t k = keys[i];
    for(int m=nmodes; m-- > 0; ) {
      inds[m][i] = k % dims[m];
      k /= dims[m];
    }
  }
  free(keys);
  for(int m=0; m < nmodes; ++m) {
    free(tabs[m].cdf);
    free(tabs[m].perm);
    dims_out[m] = dims[m];
  }
  if(compact) {
    for(int m=0; m < nmodes; ++m) {
      uint64_t * map = calloc(dims[m], sizeof(uint64_t));
      for(uint64_t i=0; i < nnz; ++i) map[inds[m][i]] = 1;
      uint64_t k = 0;
      for(uint64_t j=0; j < dims[m]; ++j) {
        if(map[j]) map[j] = ++k;
      }
      for(uint64_t i=0; i < nnz; ++i) inds[m][i] = map[inds[m][i]] - 1;
      dims_out[m] = k;
      free(map);
    }
  }
  for(uint64_t i=0; i < nnz; ++i) {
    switch(val_kind) {
    case 1: {
      double u1 = u01(&s_val);
      double const u2 = u01(&s_val);
      if(u1 < 1e-300) u1 = 1e-300;
      vals[i] = exp(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
      break;
    }
    case 2:
      vals[i] = 2.0 * u01(&s_val) - 1.0;
      break;
    default:
      vals[i] = u01(&s_val);
      break;
    }
  }
  return 0;
}

/* Fisher-Yates shuffle of the nonzeros (all modes and values together). */
void splatt_gen_shuffle(int nmodes, uint64_t nnz, uint64_t ** inds,
    double * vals, uint64_t seed)
{
  uint64_t s = seed ^ 0x3C3C3C3C5A5A5A5Aull;
  for(uint64_t i=nnz; i-- > 1; ) {
    uint64_t const j = sm64(&s) % (i + 1);
    for(int m=0; m < nmodes; ++m) {
      uint64_t const t = inds[m][i]; inds[m][i] = inds[m][j]; inds[m][j] = t;
    }
    double const t = vals[i]; vals[i] = vals[j]; vals[j] = t;
  }
}

/* Row-major dense matrix of uniform(-1, 1) values (test/bench factors). */
void splatt_gen_matrix(uint64_t nrows, uint64_t ncols, uint64_t seed, double * out)
{
  uint64_t s = seed ^ 0x2468ACE013579BDFull;
  for(uint64_t i=0; i < nrows * ncols; ++i) out[i] = 2.0 * u01(&s) - 1.0;
}
