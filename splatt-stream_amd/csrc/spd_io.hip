// spd_io.hip -- the text tensor reader on the device, and empty-index
// removal on a device COO.
//
// The text format is the reference's (src/io.c:62-108, :273-348): one
// nonzero per line, "i_1 ... i_N value"; a line is data when it is longer
// than its '\n' alone and does not start with '#'; each field is converted
// with strtoull (indices) and strtod (the value); a mode is 1-indexed unless
// its smallest index is 0, and its dim is the largest index (+1 when
// 0-indexed).
//
// The whole text is copied to the device once. Line starts come from a
// count of newlines per 4 KiB tile and a scan; each line is then classified
// and converted by one thread (spd_decimal.hpp: the exact results of
// strtoull and strtod for the usual field shapes). Lines outside those
// shapes are handed back to the host, which converts them with the C library
// (the caller's callback) and writes them into place, so every line ends up
// converted exactly as the reference converts it.
#include "spd_internal.hpp"
#include "spd_csf_dev.hpp"
#include "spd_decimal.hpp"
#include <vector>

static uint64_t const h_pow5[] = {
#include "spd_pow5.inc"
};
__constant__ uint64_t d_pow5[] = {
#include "spd_pow5.inc"
};

#define IO_THREADS 256
#define IO_BYTES 16                         // bytes per thread (one 16-byte word)
#define IO_TILE (IO_THREADS * IO_BYTES)     // bytes per block

struct IoInd { uint64_t * p[SPD_MAX_MODES]; };

// The device copy of the text is padded with zeros to a multiple of 16
// bytes plus 16, so that every 16-byte word that holds a byte of the text
// can be loaded whole.
#define IO_PAD 16

// newlines in one thread's 16 bytes (one aligned load; the padding holds
// no '\n')
__device__ __forceinline__ uint32_t io_nl_mask(char const * txt, uint64_t b0)
{
  uint4 const w = *(uint4 const *) (txt + b0);
  uint32_t const x[4] = {w.x, w.y, w.z, w.w};
  uint32_t m = 0;
  #pragma unroll
  for(int q=0; q < 4; ++q) {
    #pragma unroll
    for(int b=0; b < 4; ++b) {
      if(((x[q] >> (8 * b)) & 0xFFu) == (uint32_t) '\n') m |= 1u << (4 * q + b);
    }
  }
  return m;
}

// newlines per tile
__global__ void __launch_bounds__(IO_THREADS)
k_nl_count(char const * __restrict__ txt, uint64_t len, uint32_t * __restrict__ cnt)
{
  __shared__ uint32_t s[IO_THREADS / 64];
  uint64_t const b0 = (uint64_t) blockIdx.x * IO_TILE + (uint64_t) threadIdx.x * IO_BYTES;
  uint32_t c = (b0 < len) ? (uint32_t) __builtin_popcount(io_nl_mask(txt, b0)) : 0u;
  #pragma unroll
  for(int o=32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if(threadIdx.x == 0) cnt[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// line starts: line k + 1 starts after the (k+1)-th newline, line 0 at 0
__global__ void __launch_bounds__(IO_THREADS)
k_nl_pos(char const * __restrict__ txt, uint64_t len, uint32_t const * __restrict__ base,
    uint64_t * __restrict__ lstart)
{
  __shared__ uint32_t s[IO_THREADS / 64];
  uint64_t const b0 = (uint64_t) blockIdx.x * IO_TILE + (uint64_t) threadIdx.x * IO_BYTES;
  uint32_t const mask = (b0 < len) ? io_nl_mask(txt, b0) : 0u;
  uint32_t const c = (uint32_t) __builtin_popcount(mask);
  // exclusive scan over the block: wave scan, then the waves before
  int const lane = threadIdx.x & 63;
  uint32_t inc = c;
  #pragma unroll
  for(int o=1; o < 64; o <<= 1) {
    uint32_t const t = __shfl_up(inc, o, 64);
    if(lane >= o) inc += t;
  }
  if(lane == 63) s[threadIdx.x >> 6] = inc;
  __syncthreads();
  uint32_t woff = 0;
  for(int w=0; w < (int) (threadIdx.x >> 6); ++w) woff += s[w];
  uint64_t k = (uint64_t) base[blockIdx.x] + woff + inc - c;
  if(blockIdx.x == 0 && threadIdx.x == 0) lstart[0] = 0;
  for(uint32_t m = mask; m != 0; m &= m - 1) {
    ++k;
    lstart[k] = b0 + (uint64_t) __builtin_ctz(m) + 1;
  }
}

// line k spans [lstart[k], end): end is its '\n' (k < nl) or the text's end
__device__ __forceinline__ uint64_t io_line_end(uint64_t const * lstart, uint64_t k,
    uint64_t nl, uint64_t len)
{
  return (k < nl) ? lstart[k + 1] - 1 : len;
}

// data lines as the reference's getline loop sees them: more than the
// '\n' alone, first character other than '#'
__global__ void k_line_class(char const * __restrict__ txt, uint64_t len,
    uint64_t const * __restrict__ lstart, uint64_t nl, uint64_t nlines,
    uint32_t * __restrict__ flag)
{
  uint64_t const k = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= nlines) return;
  uint64_t const s = lstart[k];
  uint64_t const e = io_line_end(lstart, k, nl, len);
  uint64_t const rd = e - s + ((k < nl) ? 1 : 0);
  flag[k] = (rd > 1 && txt[s] != '#') ? 1u : 0u;
}

struct ParseArgs {
  char const * txt;
  uint64_t len;
  uint64_t const * lstart;
  uint64_t nl;
  uint64_t nlines;
  uint32_t const * flag;       // data line
  uint32_t const * pos;        // exclusive scan of flag: the line's nonzero
  IoInd ind;
  double * vals;
  int nmodes;
  uint64_t * host;             // lines left to the host: (start, end, nonzero) triples
  uint64_t host_max;
  unsigned long long * nhost;
};

// One line: each index is written as soon as it is converted (a line that
// then fails is rewritten whole by the host conversion).
__device__ __forceinline__ void io_parse_line(ParseArgs const & a, char const * p,
    char const * pe, uint64_t s, uint64_t e, uint64_t n)
{
  bool ok = true;
  for(int m=0; m < a.nmodes; ++m) {
    uint64_t x = 0;
    if(!spd_parse_index(&p, pe, &x)) {
      ok = false;
      break;
    }
    a.ind.p[m][n] = x;
  }
  double val = 0.;
  if(ok) ok = spd_parse_value(&p, pe, d_pow5, &val);
  if(ok) {
    a.vals[n] = val;
  } else {
    unsigned long long const slot = atomicAdd(a.nhost, 1ull);
    if(slot < a.host_max) {
      a.host[3 * slot + 0] = s;
      a.host[3 * slot + 1] = e;
      a.host[3 * slot + 2] = n;
    }
  }
}

// One thread per line. The block's lines are contiguous text: when they
// fit IO_STAGE bytes they are copied to LDS with 16-byte loads and parsed
// from there (one byte at a time from LDS instead of from global memory).
#define IO_STAGE 16384
__global__ void __launch_bounds__(IO_THREADS)
k_parse_lines(ParseArgs a)
{
  __shared__ __align__(16) char s_txt[IO_STAGE];
  uint64_t const k0 = (uint64_t) blockIdx.x * IO_THREADS;
  uint64_t const k1 = (k0 + IO_THREADS < a.nlines) ? k0 + IO_THREADS : a.nlines;
  uint64_t const bs = a.lstart[k0];
  uint64_t const be = io_line_end(a.lstart, k1 - 1, a.nl, a.len);
  uint64_t const w0 = bs & ~(uint64_t) 15;
  uint64_t const w1 = (be + 15) & ~(uint64_t) 15;
  bool const staged = (w1 - w0) <= IO_STAGE;
  if(staged) {
    uint4 const * const src = (uint4 const *) (a.txt + w0);
    uint4 * const dst = (uint4 *) s_txt;
    uint32_t const nw = (uint32_t) ((w1 - w0) / 16);
    for(uint32_t i=threadIdx.x; i < nw; i += IO_THREADS) dst[i] = src[i];
    __syncthreads();
  }
  uint64_t const k = k0 + threadIdx.x;
  if(k >= k1 || !a.flag[k]) return;
  uint64_t const s = a.lstart[k];
  uint64_t const e = io_line_end(a.lstart, k, a.nl, a.len);
  uint64_t const n = a.pos[k];
  if(staged) {
    io_parse_line(a, s_txt + (s - w0), s_txt + (e - w0), s, e, n);
  } else {
    io_parse_line(a, a.txt + s, a.txt + e, s, e, n);
  }
}

// Largest and smallest index of one mode (the smallest starts at 1, as the
// reference's reader): wave, block, then one atomic each per block.
__global__ void __launch_bounds__(256)
k_ind_extremes(uint64_t const * __restrict__ ind, uint64_t n, unsigned long long * mx,
    unsigned long long * mn)
{
  __shared__ unsigned long long s_x[4], s_y[4];
  unsigned long long x = 0ull, y = 1ull;
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    unsigned long long const v = ind[i];
    x = (v > x) ? v : x;
    y = (v < y) ? v : y;
  }
  #pragma unroll
  for(int o=32; o > 0; o >>= 1) {
    unsigned long long const x2 = __shfl_xor(x, o, 64);
    unsigned long long const y2 = __shfl_xor(y, o, 64);
    x = (x2 > x) ? x2 : x;
    y = (y2 < y) ? y2 : y;
  }
  if((threadIdx.x & 63) == 0) {
    s_x[threadIdx.x >> 6] = x;
    s_y[threadIdx.x >> 6] = y;
  }
  __syncthreads();
  if(threadIdx.x == 0) {
    for(int w=1; w < 4; ++w) {
      x = (s_x[w] > x) ? s_x[w] : x;
      y = (s_y[w] < y) ? s_y[w] : y;
    }
    atomicMax(mx, x);
    atomicMin(mn, y);
  }
}

// lines converted on the host, written into place
__global__ void k_place_lines(uint64_t nh, uint64_t const * __restrict__ at,
    uint64_t const * __restrict__ inds, double const * __restrict__ vals, int nmodes,
    IoInd dst, double * __restrict__ out_vals)
{
  uint64_t const j = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= nh) return;
  uint64_t const n = at[j];
  #pragma unroll
  for(int m=0; m < SPD_MAX_MODES; ++m) {
    if(m < nmodes) dst.p[m][n] = inds[j * nmodes + m];
  }
  out_vals[n] = vals[j];
}

// subtract 1 from the indices of a 1-indexed mode
__global__ void k_shift_down(uint64_t * __restrict__ ind, uint64_t n)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) ind[i] -= 1;
}

// one thread per item
static unsigned io_grid(uint64_t n, uint64_t per)
{
  return (unsigned) spd_cdiv(n ? n : 1, per);
}

// grid-stride loops
static unsigned io_grid_capped(uint64_t n)
{
  uint64_t const g = spd_cdiv(n ? n : 1, 256);
  return (unsigned) (g < 8192 ? g : 8192);
}

namespace {
// device buffers of one parse, freed on every exit
struct TnsBufs {
  char * txt = nullptr;
  uint32_t * cnt = nullptr;
  uint32_t * base = nullptr;
  uint64_t * lstart = nullptr;
  uint32_t * flag = nullptr;
  uint32_t * pos = nullptr;
  uint64_t * host = nullptr;
  unsigned long long * stats = nullptr;     // mx[8], mn[8], nhost
  uint64_t * at = nullptr;
  uint64_t * hinds = nullptr;
  double * hvals = nullptr;
  ~TnsBufs()
  {
    spd_sync();
    spd_free(txt);
    spd_free(cnt);
    spd_free(base);
    spd_free(lstart);
    spd_free(flag);
    spd_free(pos);
    spd_free(host);
    spd_free(stats);
    spd_free(at);
    spd_free(hinds);
    spd_free(hvals);
  }
};

int tns_parse(char const * text, size_t len, int nmodes, spd_line_fn line_fn, spd_coo * c,
    uint64_t * nhost_out)
{
  hipStream_t const st = spd_hs(nullptr);
  TnsBufs b;
  uint64_t const ntile = spd_cdiv(len ? len : 1, IO_TILE);
  uint64_t const tlen = (len + 15) / 16 * 16 + IO_PAD;
  SPD_TRY(spd_malloc((void **) &b.txt, tlen));
  uint64_t const zs = len / 16 * 16;          // the word holding byte len, and the pad
  SPD_TRY(spd_zero_async(b.txt + zs, tlen - zs, st));
  if(len) SPD_TRY(spd_h2d(b.txt, text, len));
  SPD_TRY(spd_malloc((void **) &b.cnt, ntile * sizeof(uint32_t)));
  SPD_TRY(spd_malloc((void **) &b.base, ntile * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_nl_count, dim3((unsigned) ntile), dim3(IO_THREADS), 0, st, b.txt,
      (uint64_t) len, b.cnt);
  SPD_LAUNCH_CHECK();
  if(len >= (1ull << 32)) {
    // the scan's 32-bit total must hold the newline count
    std::vector<uint32_t> hc(ntile);
    SPD_TRY(spd_d2h(hc.data(), b.cnt, ntile * sizeof(uint32_t)));
    uint64_t tot = 0;
    for(uint32_t x : hc) tot += x;
    if(tot >= (1ull << 32) - 1) {
      spd_set_error(SPD_EINPUT, "tns: %llu lines exceed the 32-bit line index",
          (unsigned long long) tot);
      return SPD_EINPUT;
    }
  }
  uint32_t nl32 = 0;
  SPD_TRY(spd_scan_u32(b.cnt, b.base, ntile, &nl32, st));
  uint64_t const nl = nl32;
  // a last line without '\n' counts; an empty one after the last '\n' does not
  uint64_t const nlines = nl + ((len > 0 && text[len - 1] != '\n') ? 1 : 0);
  SPD_TRY(spd_malloc((void **) &b.lstart, (nl + 2) * sizeof(uint64_t)));
  hipLaunchKernelGGL(k_nl_pos, dim3((unsigned) ntile), dim3(IO_THREADS), 0, st, b.txt,
      (uint64_t) len, b.base, b.lstart);
  SPD_LAUNCH_CHECK();
  SPD_TRY(spd_malloc((void **) &b.flag, (nlines ? nlines : 1) * sizeof(uint32_t)));
  SPD_TRY(spd_malloc((void **) &b.pos, (nlines ? nlines : 1) * sizeof(uint32_t)));
  uint32_t nnz32 = 0;
  if(nlines > 0) {
    hipLaunchKernelGGL(k_line_class, dim3(io_grid(nlines, 256)), dim3(256), 0, st, b.txt,
        (uint64_t) len, b.lstart, nl, nlines, b.flag);
    SPD_LAUNCH_CHECK();
    SPD_TRY(spd_scan_u32(b.flag, b.pos, nlines, &nnz32, st));
  }
  uint64_t const nnz = nnz32;
  c->nmodes = nmodes;
  c->nnz = nnz;
  for(int m=0; m < nmodes; ++m) {
    SPD_TRY(spd_malloc((void **) &c->ind[m], (nnz ? nnz : 1) * sizeof(uint64_t)));
  }
  SPD_TRY(spd_malloc((void **) &c->vals, (nnz ? nnz : 1) * sizeof(double)));
  uint64_t const host_max = nlines < SPD_TNS_HOST_MAX ? nlines : SPD_TNS_HOST_MAX;
  SPD_TRY(spd_malloc((void **) &b.host, (host_max ? host_max : 1) * 3 * sizeof(uint64_t)));
  SPD_TRY(spd_malloc((void **) &b.stats, 17 * sizeof(unsigned long long)));
  unsigned long long hs[17];
  for(int m=0; m < 8; ++m) {
    hs[m] = 0ull;
    hs[8 + m] = 1ull;
  }
  hs[16] = 0ull;
  SPD_TRY(spd_h2d(b.stats, hs, sizeof(hs)));
  if(nlines > 0) {
    ParseArgs a = {};
    a.txt = b.txt;
    a.len = len;
    a.lstart = b.lstart;
    a.nl = nl;
    a.nlines = nlines;
    a.flag = b.flag;
    a.pos = b.pos;
    for(int m=0; m < SPD_MAX_MODES; ++m) a.ind.p[m] = c->ind[m < nmodes ? m : 0];
    a.vals = c->vals;
    a.nmodes = nmodes;
    a.host = b.host;
    a.host_max = host_max;
    a.nhost = b.stats + 16;
    hipLaunchKernelGGL(k_parse_lines, dim3(io_grid(nlines, IO_THREADS)), dim3(IO_THREADS), 0,
        st, a);
    SPD_LAUNCH_CHECK();
  }
  SPD_TRY(spd_d2h(hs, b.stats, sizeof(hs)));
  uint64_t const nh = hs[16];
  *nhost_out = nh;
  if(nh > host_max) {
    spd_set_error(SPD_EHOST, "tns: %llu lines need the host conversion",
        (unsigned long long) nh);
    return SPD_EHOST;
  }
  // lines outside the device's field shapes: the caller's C-library
  // conversion, then written into place
  if(nh > 0) {
    if(line_fn == nullptr) {
      spd_set_error(SPD_EINPUT, "tns: %llu lines need a host conversion and none was given",
          (unsigned long long) nh);
      return SPD_EINPUT;
    }
    std::vector<uint64_t> rec(3 * nh), at(nh), inds(nh * nmodes);
    std::vector<double> vals(nh);
    SPD_TRY(spd_d2h(rec.data(), b.host, 3 * nh * sizeof(uint64_t)));
    for(uint64_t j=0; j < nh; ++j) {
      uint64_t const s = rec[3 * j], e = rec[3 * j + 1];
      at[j] = rec[3 * j + 2];
      if(line_fn(text + s, (size_t) (e - s), nmodes, &inds[j * nmodes], &vals[j]) != 0) {
        spd_set_error(SPD_EINPUT, "tns: the line at byte %llu could not be converted",
            (unsigned long long) s);
        return SPD_EINPUT;
      }
    }
    SPD_TRY(spd_malloc((void **) &b.at, nh * sizeof(uint64_t)));
    SPD_TRY(spd_malloc((void **) &b.hinds, nh * nmodes * sizeof(uint64_t)));
    SPD_TRY(spd_malloc((void **) &b.hvals, nh * sizeof(double)));
    SPD_TRY(spd_h2d(b.at, at.data(), nh * sizeof(uint64_t)));
    SPD_TRY(spd_h2d(b.hinds, inds.data(), nh * nmodes * sizeof(uint64_t)));
    SPD_TRY(spd_h2d(b.hvals, vals.data(), nh * sizeof(double)));
    IoInd dst = {};
    for(int m=0; m < SPD_MAX_MODES; ++m) dst.p[m] = c->ind[m < nmodes ? m : 0];
    hipLaunchKernelGGL(k_place_lines, dim3(io_grid(nh, 256)), dim3(256), 0, st, nh, b.at,
        b.hinds, b.hvals, nmodes, dst, c->vals);
    SPD_LAUNCH_CHECK();
  }
  // 0- or 1-indexed per mode, dims (src/io.c:62-108)
  if(nnz > 0) {
    for(int m=0; m < nmodes; ++m) {
      hipLaunchKernelGGL(k_ind_extremes, dim3(io_grid_capped(nnz)), dim3(256), 0, st, c->ind[m],
          nnz, b.stats + m, b.stats + 8 + m);
      SPD_LAUNCH_CHECK();
    }
    SPD_TRY(spd_d2h(hs, b.stats, sizeof(hs)));
  }
  for(int m=0; m < nmodes; ++m) {
    unsigned long long const off = hs[8 + m];
    if(off != 0 && off != 1) {
      fprintf(stderr, "SPLATT: ERROR tensors must be 0 or 1 indexed. Mode %d is %llu "
          "indexed.\n", m, off);
      spd_set_error(SPD_EINPUT, "tns: mode %d is %llu-indexed", m, off);
      return SPD_EINPUT;
    }
    c->dims[m] = hs[m] + (off == 0 ? 1 : 0);
    if(off == 1 && nnz > 0) {
      hipLaunchKernelGGL(k_shift_down, dim3(io_grid_capped(nnz)), dim3(256), 0, st, c->ind[m],
          nnz);
      SPD_LAUNCH_CHECK();
    }
  }
  return spd_sync();
}
}  // namespace

extern "C" int spd_tns_parse(char const * text, size_t len, int nmodes, spd_line_fn line_fn,
    spd_coo ** out, uint64_t * nhost)
{
  *out = nullptr;
  uint64_t nh = 0;
  if(nhost) *nhost = 0;
  SPD_TRY(spd_init());
  if(nmodes < 1 || nmodes > SPD_MAX_MODES) {
    spd_set_error(SPD_EINPUT, "tns: %d modes (1 .. %d)", nmodes, SPD_MAX_MODES);
    return SPD_EINPUT;
  }
  if(spd_cdiv(len ? len : 1, IO_TILE) >= 0x7FFFFFFFull) {
    spd_set_error(SPD_EINPUT, "tns: text of %zu bytes too long", len);
    return SPD_EINPUT;
  }
  spd_coo * c = (spd_coo *) calloc(1, sizeof(spd_coo));
  if(c == nullptr) {
    spd_set_error(SPD_ENOMEM, "host allocation of the COO handle failed");
    return SPD_ENOMEM;
  }
  int const rc = tns_parse(text, len, nmodes, line_fn, c, &nh);
  if(nhost) *nhost = nh;
  if(rc != SPD_OK) {
    spd_coo_free(c);
    return rc;
  }
  *out = c;
  return SPD_OK;
}

// ---- empty-index removal (src/sptensor.c:190-252) -----------------------

__global__ void k_mark_used(uint64_t const * __restrict__ ind, uint64_t n,
    uint32_t * __restrict__ used)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) used[ind[i]] = 1u;
}

__global__ void k_relabel(uint64_t * __restrict__ ind, uint64_t n,
    uint32_t const * __restrict__ rank)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) ind[i] = rank[ind[i]];
}

// Per mode: the indices that occur keep their order and are renumbered
// 0..k-1 (an index's new number is the count of occurring indices below it:
// an exclusive scan of the occurrence marks); dims become k.
extern "C" int spd_coo_remove_empty(spd_coo * c, uint64_t * removed)
{
  if(removed) *removed = 0;
  hipStream_t const st = spd_hs(nullptr);
  uint64_t maxdim = 1;
  for(int m=0; m < c->nmodes; ++m) if(c->dims[m] > maxdim) maxdim = c->dims[m];
  if(maxdim >= (1ull << 32) - 1) {
    spd_set_error(SPD_EINPUT, "remove_empty: dim %llu exceeds 32-bit numbering",
        (unsigned long long) maxdim);
    return SPD_EINPUT;
  }
  uint32_t * used = nullptr;
  uint32_t * rank = nullptr;
  int rc = spd_malloc((void **) &used, maxdim * sizeof(uint32_t));
  if(rc == SPD_OK) rc = spd_malloc((void **) &rank, maxdim * sizeof(uint32_t));
  uint64_t tot = 0;
  for(int m=0; m < c->nmodes && rc == SPD_OK; ++m) {
    uint64_t const dim = c->dims[m];
    if(dim == 0) continue;
    rc = spd_zero_async(used, dim * sizeof(uint32_t), st);
    if(rc == SPD_OK && c->nnz > 0) {
      hipLaunchKernelGGL(k_mark_used, dim3(io_grid_capped(c->nnz)), dim3(256), 0, st,
          c->ind[m], c->nnz, used);
      if(hipGetLastError() != hipSuccess) {
        spd_set_error(SPD_EHIP, "remove_empty: launch failed");
        rc = SPD_EHIP;
      }
    }
    uint32_t k = 0;
    if(rc == SPD_OK) rc = spd_scan_u32(used, rank, dim, &k, st);
    if(rc == SPD_OK && k < dim) {
      tot += dim - k;
      if(c->nnz > 0) {
        hipLaunchKernelGGL(k_relabel, dim3(io_grid_capped(c->nnz)), dim3(256), 0, st,
            c->ind[m], c->nnz, rank);
        if(hipGetLastError() != hipSuccess) {
          spd_set_error(SPD_EHIP, "remove_empty: launch failed");
          rc = SPD_EHIP;
        }
      }
      c->dims[m] = k;
    }
  }
  spd_sync();
  spd_free(used);
  spd_free(rank);
  if(removed && rc == SPD_OK) *removed = tot;
  return rc;
}

extern "C" void spd_coo_dims(spd_coo const * c, int * nmodes, uint64_t * nnz, uint64_t * dims)
{
  *nmodes = c->nmodes;
  *nnz = c->nnz;
  for(int m=0; m < c->nmodes; ++m) dims[m] = c->dims[m];
}

extern "C" int spd_coo_download(spd_coo const * c, uint64_t * const * inds, double * vals)
{
  for(int m=0; m < c->nmodes; ++m) {
    SPD_TRY(spd_d2h(inds[m], c->ind[m], c->nnz * sizeof(uint64_t)));
  }
  SPD_TRY(spd_d2h(vals, c->vals, c->nnz * sizeof(double)));
  return SPD_OK;
}

extern "C" int spd_decimal_convert(char const * buf, uint64_t const * off, uint64_t n, int kind,
    double * vals, uint64_t * idx, int * ok)
{
  for(uint64_t i=0; i < n; ++i) {
    char const * p = buf + off[i];
    char const * const e = buf + off[i + 1];
    bool r;
    if(kind == 0) {
      double v = 0.;
      r = spd_parse_value(&p, e, h_pow5, &v);
      vals[i] = v;
    } else {
      uint64_t v = 0;
      r = spd_parse_index(&p, e, &v);
      idx[i] = v;
    }
    ok[i] = r ? 1 : 0;
  }
  return SPD_OK;
}
