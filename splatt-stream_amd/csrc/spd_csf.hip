// spd_csf.hip -- CSF construction on the GPU.
//
// Pipeline for one tensor (level order perm[]):
//   1. key per nonzero from its level-ordered coordinates (bit-packed when
//      the level widths fit in 64 bits, mixed radix when the product of the
//      dims does, else several 64-bit keys sorted least significant first);
//   2. stable LSD radix sort of (key, nonzero id) -> canonical order, the
//      full lexicographic order the reference's tt_sort produces for tensors
//      without duplicate coordinates (src/sort.c:761-905);
//   3. optional dense tiling: a second stable sort by tile id (the
//      reference's tt_densetile bucket, src/tile.c:260-322);
//   4. fiber compression: per nonzero the first level whose index differs
//      from the previous nonzero (level 0 at tile starts); node ids per
//      level are block counts + scan + in-wave ballot ranks; fptr/fids are
//      written in one pass (the reference's p_mk_outerptr/p_mk_fptr,
//      src/csf.c:230-372, as a data-parallel rule).
// The untiled structure always stays on the device for the kernels; a tiled
// structure is kept alongside for the host mirror.
#include "spd_internal.hpp"
#include "spd_csf_dev.hpp"
#include <stdlib.h>
#include <string.h>
#include <vector>

// ---- COO upload -----------------------------------------------------------

__global__ void k_check_range(uint64_t const * __restrict__ ind, uint64_t n,
    uint64_t dim, uint32_t * __restrict__ bad)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    if(ind[i] >= dim) {
      atomicOr(bad, 1u);
    }
  }
}

extern "C" int spd_coo_upload(int nmodes, uint64_t nnz, uint64_t const * dims,
    uint64_t const * const * inds, double const * vals, spd_coo ** out)
{
  *out = nullptr;
  SPD_TRY(spd_init());
  if(nmodes < 2 || nmodes > SPD_MAX_MODES) {
    spd_set_error(SPD_EINPUT, "nmodes %d not in [2, %d]", nmodes, SPD_MAX_MODES);
    return SPD_EINPUT;
  }
  if(nnz >= (1ull << 32)) {
    spd_set_error(SPD_EINPUT, "nnz %llu exceeds the 32-bit device index range",
        (unsigned long long) nnz);
    return SPD_EINPUT;
  }
  for(int m=0; m < nmodes; ++m) {
    if(dims[m] >= (1ull << 32)) {
      spd_set_error(SPD_EINPUT, "dim %d (%llu) exceeds 32-bit device indices", m,
          (unsigned long long) dims[m]);
      return SPD_EINPUT;
    }
  }
  spd_coo * c = (spd_coo *) calloc(1, sizeof(spd_coo));
  if(c == nullptr) {
    spd_set_error(SPD_ENOMEM, "host allocation of the COO handle failed");
    return SPD_ENOMEM;
  }
  c->nmodes = nmodes;
  c->nnz = nnz;
  for(int m=0; m < nmodes; ++m) {
    c->dims[m] = dims[m];
  }
  hipStream_t st = spd_hs(nullptr);
  int rc = SPD_OK;
  uint32_t * bad = nullptr;
  rc = spd_malloc((void **) &bad, sizeof(uint32_t));
  if(rc == SPD_OK) rc = spd_zero_async(bad, sizeof(uint32_t), st);
  for(int m=0; m < nmodes && rc == SPD_OK; ++m) {
    rc = spd_malloc((void **) &c->ind[m], nnz * sizeof(uint64_t));
    if(rc == SPD_OK) rc = spd_h2d(c->ind[m], inds[m], nnz * sizeof(uint64_t));
    if(rc == SPD_OK && nnz > 0) {
      unsigned grid = (unsigned) (spd_cdiv(nnz, 256) < 4096 ? spd_cdiv(nnz, 256) : 4096);
      hipLaunchKernelGGL(k_check_range, dim3(grid), dim3(256), 0, st,
          c->ind[m], nnz, dims[m], bad);
      if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
    }
  }
  if(rc == SPD_OK) rc = spd_malloc((void **) &c->vals, nnz * sizeof(double));
  if(rc == SPD_OK) rc = spd_h2d(c->vals, vals, nnz * sizeof(double));
  uint32_t hbad = 0;
  if(rc == SPD_OK) rc = spd_d2h(&hbad, bad, sizeof(uint32_t));
  spd_free(bad);
  if(rc == SPD_OK && hbad) {
    spd_set_error(SPD_EINPUT, "coordinate out of range of its mode's dim");
    rc = SPD_EINPUT;
  }
  if(rc != SPD_OK) {
    spd_coo_free(c);
    return rc;
  }
  *out = c;
  return SPD_OK;
}

extern "C" void spd_coo_free(spd_coo * c)
{
  if(c == nullptr) return;
  for(int m=0; m < c->nmodes; ++m) spd_free(c->ind[m]);
  spd_free(c->vals);
  free(c);
}

// ---- keys -----------------------------------------------------------------

struct KeySpec {
  int nlev;                 // levels in this key
  int lev[SPD_MAX_MODES];   // which levels (most significant first)
  uint64_t mul[SPD_MAX_MODES];   // mixed radix dims, or 1<<bits for packed
  int bits;                 // key bits to sort
};

__global__ void k_make_key(KeySpec ks, DevPtrs ind, uint64_t n,
    uint32_t const * __restrict__ order, uint64_t * __restrict__ key,
    uint32_t * __restrict__ idx)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    uint32_t x = order ? order[i] : (uint32_t) i;
    if(x >= n) x = 0;   // only reachable if the permutation were corrupt
    uint64_t k = 0;
    for(int q=0; q < ks.nlev; ++q) {
      k = k * ks.mul[q] + ind.p[ks.lev[q]][x];
    }
    key[i] = k;
    idx[i] = x;
  }
}

static int bits_for(unsigned __int128 v)
{
  int b = 0;
  while(b < 128 && (((unsigned __int128) 1) << b) < v) ++b;
  return b;
}

// ---- tiling keys ------------------------------------------------------------

struct TileSpec {
  int nmodes;
  uint64_t tdim[SPD_MAX_MODES];
  uint64_t tsize[SPD_MAX_MODES];
};

// Tile id of nonzero order[i], last mode least significant (get_tile_id,
// src/tile.c:326-342); coordinate capped at tile_dims-1 (src/tile.c:283).
__global__ void k_tile_key(TileSpec ts, DevPtrs mind, uint64_t n,
    uint32_t const * __restrict__ order, uint64_t * __restrict__ key,
    uint32_t * __restrict__ idx)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    uint32_t x = order[i];
    if(x >= n) x = 0;
    uint64_t id = 0;
    for(int m=0; m < ts.nmodes; ++m) {
      uint64_t c = mind.p[m][x] / ts.tsize[m];
      if(c > ts.tdim[m] - 1) c = ts.tdim[m] - 1;
      id = id * ts.tdim[m] + c;
    }
    key[i] = id;
    idx[i] = x;
  }
}

// ---- gather into level order -------------------------------------------------

__global__ void k_gather_levels(DevPtrs ind, DevPtrs32 lind, int nlev,
    int const * perm_unused, double const * __restrict__ vals,
    uint32_t const * __restrict__ order, uint64_t n, double * __restrict__ svals)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    uint32_t x = order[i];
    if(x >= n) x = 0;
    for(int l=0; l < nlev; ++l) {
      lind.p[l][i] = (uint32_t) ind.p[l][x];
    }
    svals[i] = vals[x];
  }
}

// ---- compression ---------------------------------------------------------------

#define CP_THREADS 256
#define CP_WAVES 4
#define CP_STEPS 16
#define CP_TILE (CP_THREADS * CP_STEPS)

// Level of first difference with the previous nonzero; tiles (if given)
// restart the tree.
__global__ void k_first_diff(DevPtrs32 lind, int nlev, uint64_t n,
    uint64_t const * __restrict__ tile, uint8_t * __restrict__ lev)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    uint8_t d = (uint8_t) (nlev - 1);
    if(i == 0 || (tile != nullptr && tile[i] != tile[i-1])) {
      d = 0;
    } else {
      for(int l=0; l < nlev - 1; ++l) {
        if(lind.p[l][i] != lind.p[l][i-1]) {
          d = (uint8_t) l;
          break;
        }
      }
    }
    lev[i] = d;
  }
}

// Per block and level l < nlev-1: number of nonzeros with lev <= l.
__global__ void __launch_bounds__(CP_THREADS)
k_level_counts(uint8_t const * __restrict__ lev, uint64_t n, int nlev,
    uint32_t * __restrict__ bc, uint32_t nb)
{
  __shared__ uint32_t cnt[SPD_MAX_MODES];
  if(threadIdx.x < SPD_MAX_MODES) cnt[threadIdx.x] = 0;
  __syncthreads();
  uint32_t mine[SPD_MAX_MODES];
  for(int l=0; l < SPD_MAX_MODES; ++l) mine[l] = 0;
  uint64_t const base = (uint64_t) blockIdx.x * CP_TILE;
  for(int k=0; k < CP_STEPS; ++k) {
    uint64_t const i = base + (uint64_t) k * CP_THREADS + threadIdx.x;
    if(i < n) {
      int const d = lev[i];
      for(int l=0; l < nlev - 1; ++l) mine[l] += (d <= l);
    }
  }
  for(int l=0; l < nlev - 1; ++l) {
    uint32_t v = mine[l];
    for(int o=32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if((threadIdx.x & 63) == 0) atomicAdd(&cnt[l], v);
  }
  __syncthreads();
  if(threadIdx.x < (unsigned) (nlev - 1)) {
    bc[(uint64_t) threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
  }
}

// Emit fptr/fids. Wave w of block b owns [b*4096 + w*1024, +1024) in 16 steps
// of 64; node ids = block offset + wave offset + running count + rank.
__global__ void __launch_bounds__(CP_THREADS)
k_emit(uint8_t const * __restrict__ lev, DevPtrs32 lind, uint64_t n, int nlev,
    uint32_t const * __restrict__ boff, uint32_t nb, DevPtrs32 fptr,
    DevPtrs32 fids, int store_root_ids, uint64_t const * __restrict__ tile,
    uint32_t * __restrict__ tile_off, LevelCounts nfibs, uint64_t ntiles)
{
  __shared__ uint32_t wsum[CP_WAVES][SPD_MAX_MODES];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const L = nlev - 1;
  uint64_t const lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint64_t const base = (uint64_t) blockIdx.x * CP_TILE + (uint64_t) wave * (CP_TILE / CP_WAVES);

  // pass 1: per-wave totals
  uint32_t tot[SPD_MAX_MODES];
  for(int l=0; l < SPD_MAX_MODES; ++l) tot[l] = 0;
  for(int j=0; j < CP_STEPS; ++j) {
    uint64_t const i = base + (uint64_t) j * 64 + lane;
    int const d = (i < n) ? (int) lev[i] : 127;
    for(int l=0; l < L; ++l) {
      tot[l] += (uint32_t) __popcll(__ballot(d <= l));
    }
  }
  if(lane == 0) {
    for(int l=0; l < L; ++l) wsum[wave][l] = tot[l];
  }
  __syncthreads();
  uint32_t run[SPD_MAX_MODES];
  for(int l=0; l < L; ++l) {
    uint32_t o = boff[(uint64_t) l * nb + blockIdx.x];
    for(int w=0; w < wave; ++w) o += wsum[w][l];
    run[l] = o;
  }
  // pass 2: ranks and writes
  for(int j=0; j < CP_STEPS; ++j) {
    uint64_t const i = base + (uint64_t) j * 64 + lane;
    bool const ok = i < n;
    int const d = ok ? (int) lev[i] : 127;
    uint32_t node[SPD_MAX_MODES + 1];
    for(int l=0; l < L; ++l) {
      uint64_t const bal = __ballot(d <= l);
      node[l] = run[l] + (uint32_t) __popcll(bal & lt);
      run[l] += (uint32_t) __popcll(bal);
    }
    node[L] = (uint32_t) i;
    if(ok) {
      for(int l=d; l < L; ++l) {
        if(node[l] >= nfibs.v[l]) continue;   // unreachable with consistent counts
        if(l > 0 || store_root_ids) {
          fids.p[l][node[l]] = lind.p[l][i];
        }
        fptr.p[l][node[l]] = node[l + 1];
      }
      if(tile != nullptr && d == 0 && (i == 0 || tile[i] != tile[i-1]) &&
          tile[i] < ntiles) {
        uint32_t * o = tile_off + (uint64_t) tile[i] * (uint64_t) nlev;
        for(int l=0; l <= L; ++l) o[l] = node[l];
      }
    }
  }
}

// ---- the build --------------------------------------------------------------

static void free_struct(CsfArrays * a, int nlev)
{
  for(int l=0; l < nlev; ++l) {
    spd_free(a->fptr[l]); a->fptr[l] = nullptr;
    spd_free(a->fids[l]); a->fids[l] = nullptr;
  }
  spd_free(a->vals); a->vals = nullptr;
}

// Compress nonzeros already in final order (lind[l], svals) into `a`.
static int compress(CsfArrays * a, DevPtrs32 lind, double * svals, uint64_t n,
    int nlev, uint64_t const * tile, uint32_t * tile_off, uint64_t ntiles,
    int store_root_ids, hipStream_t st)
{
  uint64_t const nb = spd_cdiv(n, CP_TILE);
  uint8_t * lev = nullptr;
  uint32_t * bc = nullptr;
  SPD_TRY(spd_malloc((void **) &lev, n));
  int rc = spd_malloc((void **) &bc, (uint64_t) (nlev - 1) * nb * sizeof(uint32_t));
  if(rc != SPD_OK) { spd_free(lev); return rc; }
  unsigned grid = (unsigned) (spd_cdiv(n, 256) < 16384 ? spd_cdiv(n, 256) : 16384);
  hipLaunchKernelGGL(k_first_diff, dim3(grid), dim3(256), 0, st, lind, nlev, n, tile, lev);
  hipLaunchKernelGGL(k_level_counts, dim3((unsigned) nb), dim3(CP_THREADS), 0, st,
      lev, n, nlev, bc, (uint32_t) nb);
  if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
  a->nfibs[nlev - 1] = n;
  for(int l=0; l < nlev - 1 && rc == SPD_OK; ++l) {
    uint32_t tot = 0;
    rc = spd_scan_u32(bc + (uint64_t) l * nb, bc + (uint64_t) l * nb, nb, &tot, st);
    a->nfibs[l] = tot;
  }
  DevPtrs32 fp = {}, fi = {};
  for(int l=0; l < nlev - 1 && rc == SPD_OK; ++l) {
    rc = spd_malloc((void **) &a->fptr[l], (a->nfibs[l] + 1) * sizeof(uint32_t));
    if(rc == SPD_OK) rc = spd_malloc((void **) &a->fids[l], (a->nfibs[l] ? a->nfibs[l] : 1) * sizeof(uint32_t));
    fp.p[l] = a->fptr[l];
    fi.p[l] = a->fids[l];
  }
  if(rc == SPD_OK) {
    LevelCounts lc = {};
    for(int l=0; l < nlev; ++l) lc.v[l] = a->nfibs[l];
    hipLaunchKernelGGL(k_emit, dim3((unsigned) nb), dim3(CP_THREADS), 0, st,
        lev, lind, n, nlev, bc, (uint32_t) nb, fp, fi, store_root_ids, tile, tile_off,
        lc, ntiles);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
  }
  // closing entries fptr[l][nfibs[l]] = nfibs[l+1]
  for(int l=0; l < nlev - 1 && rc == SPD_OK; ++l) {
    uint32_t const v = (uint32_t) a->nfibs[l + 1];
    if(hipMemcpyAsync(a->fptr[l] + a->nfibs[l], &v, sizeof(uint32_t),
        hipMemcpyHostToDevice, st) != hipSuccess) rc = SPD_EHIP;
    if(rc == SPD_OK && hipStreamSynchronize(st) != hipSuccess) rc = SPD_EHIP;
  }
  spd_free(lev);
  spd_free(bc);
  if(rc == SPD_EHIP) spd_set_error(SPD_EHIP, "compress: %s", hipGetErrorString(hipGetLastError()));
  return rc;
}

// Lexicographic order of n coordinates under the level order given by
// lind64 (level l -> index array) with per-level sizes dims_l: the stable
// LSD radix sort over one or more packed keys. On return *order points at
// the permutation (one of v0/v1); k0/k1/v0/v1 are n-element work buffers.
static int lex_order(DevPtrs const & lind64, uint64_t const * dims_l, int N, uint64_t n,
    uint64_t * k0, uint64_t * k1, uint32_t * v0, uint32_t * v1, uint32_t ** order_out,
    hipStream_t st)
{
  unsigned grid = (unsigned) (spd_cdiv(n, 256) < 16384 ? spd_cdiv(n, 256) : 16384);
  int rc = SPD_OK;
  std::vector<KeySpec> groups;
  {
    int packed_bits = 0;
    unsigned __int128 prod = 1;
    int bits_l[SPD_MAX_MODES];
    for(int l=0; l < N; ++l) {
      uint64_t const d = dims_l[l] ? dims_l[l] : 1;
      bits_l[l] = bits_for(d);
      packed_bits += bits_l[l];
      if(prod <= ((unsigned __int128) 1 << 64)) prod *= d;
    }
    if(packed_bits <= 64) {
      KeySpec ks = {};
      ks.nlev = N;
      for(int l=0; l < N; ++l) { ks.lev[l] = l; ks.mul[l] = 1ull << bits_l[l]; }
      ks.bits = packed_bits;
      groups.push_back(ks);
    } else if(prod < ((unsigned __int128) 1 << 64)) {
      KeySpec ks = {};
      ks.nlev = N;
      for(int l=0; l < N; ++l) { ks.lev[l] = l; ks.mul[l] = dims_l[l]; }
      ks.bits = bits_for(prod);
      groups.push_back(ks);
    } else {
      int l = N - 1;
      while(l >= 0) {
        int acc = 0;
        int first = l;
        while(first >= 0 && acc + bits_l[first] <= 64) { acc += bits_l[first]; --first; }
        if(first == l) { acc = bits_l[l]; first = l - 1; } // a single level wider than 64 bits cannot happen
        KeySpec ks = {};
        ks.nlev = l - first;
        for(int q=0; q < ks.nlev; ++q) {
          ks.lev[q] = first + 1 + q;
          ks.mul[q] = 1ull << bits_l[first + 1 + q];
        }
        ks.bits = acc;
        groups.push_back(ks);
        l = first;
      }
    }
  }
  uint32_t * order = nullptr;   // current permutation (device)
  for(size_t g=0; g < groups.size() && rc == SPD_OK; ++g) {
    KeySpec ks = groups[g];
    hipLaunchKernelGGL(k_make_key, dim3(grid), dim3(256), 0, st, ks, lind64, n,
        order, k0, v0);
    if(hipGetLastError() != hipSuccess) { rc = SPD_EHIP; break; }
    int which = 0;
    rc = spd_radix_sort(k0, v0, k1, v1, n, ks.bits, &which, st);
    if(rc != SPD_OK) break;
    if(which) { uint64_t * t = k0; k0 = k1; k1 = t; uint32_t * tv = v0; v0 = v1; v1 = tv; }
    // v0 holds the permutation; the next group writes its payload into v1
    order = v0;
    uint32_t * t = v0; v0 = v1; v1 = t;
  }
  *order_out = order;
  return rc;
}

extern "C" int spd_csf_build(spd_coo const * coo, int const * perm,
    uint64_t const * tile_dims, spd_csf ** out)
{
  *out = nullptr;
  SPD_TRY(spd_init());
  hipStream_t st = spd_hs(nullptr);
  int const N = coo->nmodes;
  uint64_t const n = coo->nnz;
  if(n == 0) {
    spd_set_error(SPD_EINPUT, "empty tensor");
    return SPD_EINPUT;
  }
  spd_csf * c = (spd_csf *) calloc(1, sizeof(spd_csf));
  if(c == nullptr) {
    spd_set_error(SPD_ENOMEM, "host allocation of the CSF handle failed");
    return SPD_ENOMEM;
  }
  c->nmodes = N;
  c->nnz = n;
  for(int m=0; m < N; ++m) {
    c->dims[m] = coo->dims[m];
    c->perm[m] = perm[m];
    c->tile_dims[m] = tile_dims ? tile_dims[m] : 1;
  }
  c->ntiles = 1;
  for(int m=0; m < N; ++m) c->ntiles *= c->tile_dims[m];

  // level-ordered device index pointers
  DevPtrs lind64 = {};
  DevPtrs mind64 = {};
  for(int l=0; l < N; ++l) lind64.p[l] = coo->ind[perm[l]];
  for(int m=0; m < N; ++m) mind64.p[m] = coo->ind[m];

  uint64_t * k0 = nullptr; uint64_t * k1 = nullptr;
  uint32_t * v0 = nullptr; uint32_t * v1 = nullptr;
  int rc = spd_malloc((void **) &k0, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &k1, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v0, n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v1, n * 4);
  unsigned grid = (unsigned) (spd_cdiv(n, 256) < 16384 ? spd_cdiv(n, 256) : 16384);

  uint32_t * order = nullptr;
  {
    uint64_t dl[SPD_MAX_MODES];
    for(int l=0; l < N; ++l) dl[l] = c->dims[perm[l]];
    if(rc == SPD_OK) rc = lex_order(lind64, dl, N, n, k0, k1, v0, v1, &order, st);
  }
  // canonical order is in `order`
  DevPtrs32 lind = {};
  double * svals = nullptr;
  for(int l=0; l < N && rc == SPD_OK; ++l) rc = spd_malloc((void **) &lind.p[l], n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &svals, n * 8);
  if(rc == SPD_OK) {
    hipLaunchKernelGGL(k_gather_levels, dim3(grid), dim3(256), 0, st, lind64, lind, N,
        (int const *) nullptr, coo->vals, order, n, svals);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
  }
  if(rc == SPD_OK) {
    rc = compress(&c->u, lind, svals, n, N, nullptr, nullptr, 1, 1, st);
    if(rc == SPD_OK) {
      c->u.fids[N-1] = lind.p[N-1];
      lind.p[N-1] = nullptr;
      c->u.vals = svals;
      svals = nullptr;
    }
  }

  // dense tiling for the host mirror
  if(rc == SPD_OK && c->ntiles > 1) {
    TileSpec ts = {};
    ts.nmodes = N;
    for(int m=0; m < N; ++m) {
      ts.tdim[m] = c->tile_dims[m];
      uint64_t s = c->dims[m] / c->tile_dims[m];
      ts.tsize[m] = s < 1 ? 1 : s;
    }
    hipLaunchKernelGGL(k_tile_key, dim3(grid), dim3(256), 0, st, ts, mind64, n,
        order, k0, v0);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
    int which = 0;
    int const tbits = bits_for((unsigned __int128) c->ntiles);
    if(rc == SPD_OK) rc = spd_radix_sort(k0, v0, k1, v1, n, tbits, &which, st);
    uint64_t * tkey = which ? k1 : k0;
    uint32_t * torder = which ? v1 : v0;
    DevPtrs32 tl = {};
    double * tv = nullptr;
    for(int l=0; l < N && rc == SPD_OK; ++l) rc = spd_malloc((void **) &tl.p[l], n * 4);
    if(rc == SPD_OK) rc = spd_malloc((void **) &tv, n * 8);
    if(rc == SPD_OK) rc = spd_malloc((void **) &c->tile_off, c->ntiles * N * sizeof(uint32_t));
    if(rc == SPD_OK) rc = spd_zero_async(c->tile_off, c->ntiles * N * sizeof(uint32_t), st);
    if(rc == SPD_OK) {
      hipLaunchKernelGGL(k_gather_levels, dim3(grid), dim3(256), 0, st, lind64, tl, N,
          (int const *) nullptr, coo->vals, torder, n, tv);
      if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
    }
    if(rc == SPD_OK) {
      rc = compress(&c->t, tl, tv, n, N, tkey, c->tile_off, c->ntiles, 1, st);
      if(rc == SPD_OK) {
        c->t.fids[N-1] = tl.p[N-1];
        tl.p[N-1] = nullptr;
        c->t.vals = tv;
        tv = nullptr;
      }
    }
    // per-tile nnz counts from the sorted tile ids
    if(rc == SPD_OK) {
      std::vector<uint64_t> hkey(n);
      rc = spd_d2h(hkey.data(), tkey, n * 8);
      c->tile_nnz = (uint64_t *) calloc(c->ntiles + 1, sizeof(uint64_t));
      if(c->tile_nnz == nullptr) {
        spd_set_error(SPD_ENOMEM, "host allocation of the tile counts failed");
        rc = SPD_ENOMEM;
      }
      for(uint64_t i=0; i < n && rc == SPD_OK; ++i) c->tile_nnz[hkey[i] + 1]++;
      for(uint64_t t=0; t < c->ntiles && rc == SPD_OK; ++t) c->tile_nnz[t+1] += c->tile_nnz[t];
    }
    for(int l=0; l < N; ++l) spd_free(tl.p[l]);
    spd_free(tv);
  }

  for(int l=0; l < N; ++l) spd_free(lind.p[l]);
  spd_free(svals);
  spd_free(k0); spd_free(k1); spd_free(v0); spd_free(v1);
  if(rc == SPD_OK) rc = (hipStreamSynchronize(st) == hipSuccess) ? SPD_OK : SPD_EHIP;
  if(rc != SPD_OK) {
    spd_csf_free(c);
    return rc;
  }
  *out = c;
  return SPD_OK;
}

extern "C" void spd_csf_free(spd_csf * c)
{
  if(c == nullptr) return;
  free_struct(&c->u, c->nmodes);
  free_struct(&c->t, c->nmodes);
  spd_free(c->vals32);
  spd_free(c->tile_off);
  spd_csf_drop_plans(c);
  free(c->tile_nnz);
  free(c);
}

extern "C" uint64_t spd_csf_ntiles(spd_csf const * c) { return c->ntiles; }

// Node offsets of every tile at every level, from the tile-start records;
// empty tiles take the offsets of the next non-empty tile (or the totals).
static int tile_offsets(spd_csf const * c, std::vector<uint64_t> & off)
{
  int const N = c->nmodes;
  uint64_t const T = c->ntiles;
  off.assign((T + 1) * N, 0);
  if(T == 1) {
    for(int l=0; l < N; ++l) off[N + l] = c->u.nfibs[l];
    return SPD_OK;
  }
  std::vector<uint32_t> raw(T * N);
  SPD_TRY(spd_d2h(raw.data(), c->tile_off, T * N * sizeof(uint32_t)));
  for(int l=0; l < N; ++l) off[T * N + l] = c->t.nfibs[l];
  for(uint64_t t=T; t-- != 0; ) {
    bool const empty = c->tile_nnz[t+1] == c->tile_nnz[t];
    for(int l=0; l < N; ++l) {
      off[t * N + l] = empty ? off[(t + 1) * N + l] : raw[t * N + l];
    }
  }
  return SPD_OK;
}

extern "C" int spd_csf_tile_nfibs(spd_csf const * c, uint64_t * nfibs)
{
  int const N = c->nmodes;
  std::vector<uint64_t> off;
  SPD_TRY(tile_offsets(c, off));
  for(uint64_t t=0; t < c->ntiles; ++t) {
    for(int l=0; l < N; ++l) {
      nfibs[t * N + l] = off[(t + 1) * N + l] - off[t * N + l];
    }
  }
  return SPD_OK;
}

extern "C" int spd_csf_export_tile(spd_csf const * c, uint64_t t, uint64_t ** fptr,
    uint64_t ** fids, double * vals)
{
  int const N = c->nmodes;
  CsfArrays const * a = (c->ntiles > 1) ? &c->t : &c->u;
  std::vector<uint64_t> off;
  SPD_TRY(tile_offsets(c, off));
  uint64_t const * o0 = &off[t * N];
  uint64_t const * o1 = &off[(t + 1) * N];
  // 4-byte device entries widened to the mirror's 8-byte ones (less the
  // tile's first child for fptr) chunk by chunk, on the host threads, while
  // the next chunk copies
  auto widen = [](uint32_t const * dev, uint64_t cnt, uint64_t * dst, uint64_t base) -> int {
    return spd_d2h_chunks(dev, cnt * sizeof(uint32_t),
        [&](void const * h, size_t off, size_t nbytes) {
          uint32_t const * const s = (uint32_t const *) h;
          uint64_t * const d = dst + off / sizeof(uint32_t);
          spd_par_range(nbytes / sizeof(uint32_t), [&](uint64_t lo, uint64_t hi) {
            for(uint64_t f=lo; f < hi; ++f) d[f] = (uint64_t) s[f] - base;
          });
        });
  };
  for(int l=0; l < N; ++l) {
    uint64_t const cnt = o1[l] - o0[l];
    if(cnt == 0) continue;
    if(l < N - 1 && fptr != nullptr && fptr[l] != nullptr) {
      uint64_t const base = o0[l + 1];
      SPD_TRY(widen(a->fptr[l] + o0[l], cnt, fptr[l], base));
      fptr[l][cnt] = o1[l + 1] - base;
    }
    if(fids != nullptr && fids[l] != nullptr) {
      SPD_TRY(widen(a->fids[l] + o0[l], cnt, fids[l], 0));
    }
  }
  uint64_t const nn = o1[N - 1] - o0[N - 1];
  if(vals != nullptr && nn > 0) {
    SPD_TRY(spd_d2h_chunks(a->vals + o0[N - 1], nn * sizeof(double),
        [&](void const * h, size_t off, size_t nbytes) {
          double * const d = vals + off / sizeof(double);
          double const * const s = (double const *) h;
          spd_par_range(nbytes / sizeof(double), [&](uint64_t lo, uint64_t hi) {
            if(hi > lo) memcpy(d + lo, s + lo, (hi - lo) * sizeof(double));
          });
        }));
  }
  return SPD_OK;
}

__global__ void k_sumsq(double const * __restrict__ v, uint64_t n, double * __restrict__ part)
{
  __shared__ double s[256];
  double acc = 0;
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) acc += v[i] * v[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for(int o=128; o > 0; o >>= 1) {
    if((int) threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[blockIdx.x] = s[0];
}

extern "C" int spd_csf_frobsq(spd_csf const * c, double * out)
{
  unsigned const nb = 1024;
  void * p = nullptr;
  SPD_TRY(spd_scratch(1, nb * sizeof(double), &p));
  hipStream_t st = spd_hs(nullptr);
  hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, st, c->u.vals, c->nnz, (double *) p);
  SPD_LAUNCH_CHECK();
  std::vector<double> h(nb);
  SPD_HIP(hipMemcpyAsync(h.data(), p, nb * sizeof(double), hipMemcpyDeviceToHost, st));
  SPD_HIP(hipStreamSynchronize(st));
  double s = 0;
  for(unsigned b=0; b < nb; ++b) s += h[b];
  *out = s;
  return SPD_OK;
}

extern "C" int spd_coo_lexsort(spd_coo const * coo, int const * perm, uint32_t * order_host)
{
  SPD_TRY(spd_init());
  hipStream_t st = spd_hs(nullptr);
  int const N = coo->nmodes;
  uint64_t const n = coo->nnz;
  if(n == 0) return SPD_OK;
  DevPtrs lind64 = {};
  uint64_t dl[SPD_MAX_MODES];
  for(int l=0; l < N; ++l) {
    lind64.p[l] = coo->ind[perm[l]];
    dl[l] = coo->dims[perm[l]];
  }
  uint64_t * k0 = nullptr; uint64_t * k1 = nullptr;
  uint32_t * v0 = nullptr; uint32_t * v1 = nullptr;
  int rc = spd_malloc((void **) &k0, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &k1, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v0, n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v1, n * 4);
  uint32_t * order = nullptr;
  if(rc == SPD_OK) rc = lex_order(lind64, dl, N, n, k0, k1, v0, v1, &order, st);
  if(rc == SPD_OK) rc = spd_d2h(order_host, order, n * sizeof(uint32_t));
  spd_free(k0); spd_free(k1); spd_free(v0); spd_free(v1);
  return rc;
}
