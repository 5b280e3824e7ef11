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
#include <omp.h>
#include <stdio.h>

#define GEN_MAX_MODES 8
#define GEN_CHUNK (1u << 20)

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

/* Linear cell keys: 128 bits, so tensors with more than 2^64 cells work. */
typedef unsigned __int128 gkey;

/* LSD radix sort of keys, 16-bit digits, OpenMP-parallel with per-thread
 * histograms (stable, so the result is thread-count independent). */
static void sort_keys(gkey * a, uint64_t n, int bits)
{
  if(n < 2) return;
  gkey * b = malloc(n * sizeof(gkey));
  int const T = omp_get_max_threads();
  uint64_t * cnt = malloc((size_t) T * 65536 * sizeof(uint64_t));
  gkey * src = a;
  gkey * dst = b;
  for(int sh=0; sh < bits; sh += 16) {
    #pragma omp parallel num_threads(T)
    {
      int const t = omp_get_thread_num();
      uint64_t const lo = n * (uint64_t) t / T;
      uint64_t const hi = n * (uint64_t) (t + 1) / T;
      uint64_t * c = cnt + (size_t) t * 65536;
      memset(c, 0, 65536 * sizeof(uint64_t));
      for(uint64_t i=lo; i < hi; ++i) c[(uint64_t) (src[i] >> sh) & 0xFFFF]++;
      #pragma omp barrier
      #pragma omp single
      {
        uint64_t run = 0;
        for(int d=0; d < 65536; ++d) {
          for(int u=0; u < T; ++u) {
            uint64_t const x = cnt[(size_t) u * 65536 + d];
            cnt[(size_t) u * 65536 + d] = run;
            run += x;
          }
        }
      }
      for(uint64_t i=lo; i < hi; ++i) dst[c[(uint64_t) (src[i] >> sh) & 0xFFFF]++] = src[i];
    }
    gkey * t = src; src = dst; dst = t;
  }
  if(src != a) memcpy(a, src, n * sizeof(gkey));
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
  for(int m=0; m < nmodes; ++m) {
    if(dims[m] >= ((uint64_t) 1 << 62)) return -1;
  }
  if((unsigned __int128) nnz > cells) {
    return -1;
  }
  int bits = 0;
  while(bits < 127 && ((unsigned __int128) 1 << bits) < cells) ++bits;

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

  /* Draw in rounds; each round draws fixed-size chunks whose streams depend
   * only on (seed, round, chunk), sorts the new keys, and merges them into
   * the sorted set of distinct keys. The size of the next round follows the
   * fraction of new distinct keys the last round produced. */
  gkey * keys = malloc((nnz + 1) * sizeof(gkey));
  uint64_t have = 0;
  double yield_est = 1.0;
  for(uint64_t round=0; have < nnz; ++round) {
    uint64_t const deficit = nnz - have;
    uint64_t want = (uint64_t) ((double) deficit / yield_est * 1.05) + 1024;
    gkey * fresh = malloc(want * sizeof(gkey));
    uint64_t const nchunks = (want + GEN_CHUNK - 1) / GEN_CHUNK;
    #pragma omp parallel for schedule(dynamic, 1)
    for(uint64_t ch=0; ch < nchunks; ++ch) {
      uint64_t st = s_coord ^ (0x9E3779B97F4A7C15ull * (round * 65536 + ch + 1));
      uint64_t const lo = ch * GEN_CHUNK;
      uint64_t const hi = (lo + GEN_CHUNK < want) ? lo + GEN_CHUNK : want;
      for(uint64_t i=lo; i < hi; ++i) {
        gkey k = 0;
        for(int m=0; m < nmodes; ++m) {
          k = k * dims[m] + draw(tabs + m, dims[m], zipf_s[m], &st);
        }
        fresh[i] = k;
      }
    }
    sort_keys(fresh, want, bits);
    uint64_t u = 0;
    for(uint64_t i=0; i < want; ++i) {
      if(u == 0 || fresh[i] != fresh[u-1]) fresh[u++] = fresh[i];
    }
    /* merge into the sorted set of distinct keys */
    gkey * merged = malloc((have + u) * sizeof(gkey));
    uint64_t i = 0, j = 0, w = 0;
    while(i < have || j < u) {
      gkey v;
      if(j >= u || (i < have && keys[i] <= fresh[j])) {
        v = keys[i++];
        if(j < u && fresh[j] == v) ++j;
      } else {
        v = fresh[j++];
      }
      merged[w++] = v;
    }
    uint64_t const added = w - have;
    yield_est = (double) added / (double) want;
    if(yield_est < 0.01) yield_est = 0.01;
    free(fresh);
    free(keys);
    keys = merged;
    have = w;
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
    gkey k = keys[i];
    for(int m=nmodes; m-- > 0; ) {
      inds[m][i] = (uint64_t) (k % dims[m]);
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
  uint64_t const vchunks = (nnz + GEN_CHUNK - 1) / GEN_CHUNK;
  #pragma omp parallel for schedule(dynamic, 1)
  for(uint64_t ch=0; ch < vchunks; ++ch) {
    uint64_t sv = s_val ^ (0xD1B54A32D192ED03ull * (ch + 1));
    uint64_t const hi = ((ch + 1) * GEN_CHUNK < nnz) ? (ch + 1) * GEN_CHUNK : nnz;
    for(uint64_t i=ch * GEN_CHUNK; i < hi; ++i) {
      switch(val_kind) {
      case 1: {
        double u1 = u01(&sv);
        double const u2 = u01(&sv);
        if(u1 < 1e-300) u1 = 1e-300;
        vals[i] = exp(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        break;
      }
      case 2:
        vals[i] = 2.0 * u01(&sv) - 1.0;
        break;
      default:
        vals[i] = u01(&sv);
        break;
      }
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

/* Write a tensor as .tns text (1-indexed, "i_1 ... i_N value\n", values
 * with 17 significant digits, %.17g, so they read back exactly). The lines
 * are formatted by all threads into per-thread chunks, then written in
 * order. 0 on success. */
int splatt_gen_write_tns(char const * path, int nmodes, uint64_t nnz,
    uint64_t const * const * inds, double const * vals)
{
  FILE * f = fopen(path, "wb");
  if(f == NULL) {
    return -1;
  }
  uint64_t const chunk = 1 << 16;
  uint64_t const nchunks = (nnz + chunk - 1) / chunk;
  int const T = omp_get_max_threads();
  /* rounds of T chunks: format in parallel, write in order */
  size_t const cap = (size_t) chunk * (size_t) (21 * nmodes + 32);
  char * bufs = malloc((size_t) T * cap);
  size_t * lens = malloc((size_t) T * sizeof(size_t));
  int rc = (bufs && lens) ? 0 : -1;
  for(uint64_t c0=0; c0 < nchunks && rc == 0; c0 += (uint64_t) T) {
    #pragma omp parallel for schedule(static, 1) num_threads(T)
    for(int t=0; t < T; ++t) {
      uint64_t const c = c0 + (uint64_t) t;
      lens[t] = 0;
      if(c >= nchunks) {
        continue;
      }
      char * p = bufs + (size_t) t * cap;
      uint64_t const lo = c * chunk;
      uint64_t const hi = (lo + chunk < nnz) ? lo + chunk : nnz;
      for(uint64_t n=lo; n < hi; ++n) {
        for(int m=0; m < nmodes; ++m) {
          p += sprintf(p, "%llu ", (unsigned long long) (inds[m][n] + 1));
        }
        p += sprintf(p, "%.17g\n", vals[n]);
      }
      lens[t] = (size_t) (p - (bufs + (size_t) t * cap));
    }
    for(int t=0; t < T && rc == 0; ++t) {
      if(lens[t] && fwrite(bufs + (size_t) t * cap, 1, lens[t], f) != lens[t]) {
        rc = -1;
      }
    }
  }
  free(bufs);
  free(lens);
  if(fclose(f) != 0) {
    rc = -1;
  }
  return rc;
}
