// spd_mttkrp.hip -- MTTKRP over a device CSF (root, internal and leaf output
// levels), the GPU counterpart of the reference's mttkrp_csf
// (src/mttkrp.c:1287-1343 and the kernels at :390-1278).
//
// Work decomposition. The nonzeros (leaf level, in CSF order) are cut into
// chunks of 256 consecutive nonzeros without regard to fiber or slice
// boundaries, so a giant Zipf slice is spread over many chunks and every
// chunk is the same amount of leaf work. A wavefront owns a chunk; it is
// split between 64/L groups of L lanes, each lane holding V consecutive
// columns of a factor row (a second grid dimension tiles wider ranks).
// Before walking, the wave stages the chunk's structure (node rows, node
// starts, values) into LDS with coalesced loads, so the walk never waits
// on a dependent global load: it reads its records from LDS and keeps a
// batch of factor-row gathers in flight.
//
// 3-mode root output: k_root3h (spd_root3h.hpp) follows the reference's
// CSF factoring (the leaf sum of a fiber is multiplied by the fiber's row
// once) with the hottest leaf rows in LDS; k_root3 multiplies both rows per
// nonzero and serves the layouts k_root3h is not built for (ranks below 16,
// small tensors). Every other output (k_walks) multiplies, per nonzero, the
// rows of all levels but the output level and sums runs of equal output
// rows (contiguous in CSF order) in registers. Root rows are unique per
// node: runs closed inside a group's range are stored, runs cut by range
// edges are added atomically. Internal and leaf outputs scatter to rows
// shared by many nodes; the rows with the most contributions (top-K,
// chosen once per tensor and level) are accumulated per block in LDS with
// LDS atomics and written once per block, the remaining (cold) rows take
// global atomics.
#include "spd_internal.hpp"
#include "spd_csf_dev.hpp"
#include <stdlib.h>
#include <vector>
#include <mutex>

#define MK_THREADS 256
#define KIND_ROOT 0
#define KIND_INTL 1
#define KIND_LEAF 2

template<typename T, int V> struct VecT;
template<> struct VecT<float, 4> { typedef float __attribute__((ext_vector_type(4))) type; };
template<> struct VecT<float, 1> { typedef float type; };
template<> struct VecT<double, 2> { typedef double __attribute__((ext_vector_type(2))) type; };
template<> struct VecT<double, 1> { typedef double type; };

template<typename VT> __device__ __forceinline__ VT vzero() { VT z = (VT) 0; return z; }

template<typename T, int V>
__device__ __forceinline__ typename VecT<T, V>::type
vload(T const * __restrict__ base, uint32_t row, uint32_t R, uint32_t col, bool ok)
{
  typedef typename VecT<T, V>::type VT;
  if(!ok) return vzero<VT>();
  return *(VT const *) (base + (uint64_t) row * R + col);
}

template<typename T, int V>
__device__ __forceinline__ void
vstore(T * __restrict__ base, uint32_t row, uint32_t R, uint32_t col, bool ok,
    typename VecT<T, V>::type v)
{
  typedef typename VecT<T, V>::type VT;
  if(ok) *(VT *) (base + (uint64_t) row * R + col) = v;
}

__device__ __forceinline__ void vatomic(float * p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void vatomic(double * p, double v) { unsafeAtomicAdd(p, v); }

template<typename T, int V>
__device__ __forceinline__ void
vadd_atomic(T * base, uint32_t row, uint32_t R, uint32_t col, bool ok,
    typename VecT<T, V>::type v)
{
  if(!ok) return;
  T * p = base + (uint64_t) row * R + col;
  if constexpr (V == 1) {
    vatomic(p, v);
  } else {
    #pragma unroll
    for(int k=0; k < V; ++k) vatomic(p + k, v[k]);
  }
}

#define CARRY_NONE 0xFFFFFFFFu

// ---- 3-mode root kernel with LDS staging ---------------------------------
//
// The hot path of an all-mode (or two-mode) CSF: out[s] = sum over fibers f
// of slice s of A[fid(f)] o sum_{nnz j in f} v_j B[leaf(j)].
//
// A wave owns 256 consecutive nonzeros, split into G = 64/L items of 256/G
// nonzeros (one per group of L lanes). Before walking, the wave stages into
// LDS, with coalesced loads: the values and leaf ids of its nonzeros and the
// id of every fiber and output row of every slice that overlaps its range.
// Where fibers and slices start comes from two bitmaps built once per CSF
// (one bit per nonzero: a fiber starts, a slice starts); a position's local
// fiber and slice index is a population count over the wave's 256 bits. A group then walks its item in batches of R3_BATCH
// nonzeros: the B row and the fiber's A row of every nonzero of a batch are
// gathered together (2 * R3_BATCH gathers in flight), and every boundary
// test reads LDS. k_root3h (spd_root3h.hpp) is the persistent variant that
// gathers A once per fiber and keeps the hottest B rows in LDS.
// Slices completed inside an item are stored; the partial sums at item
// edges are combined per wave and then per block (across lane groups in
// registers when they all belong to one slice, through LDS otherwise), and
// only a block's first and last slice (when they cross the block's edges)
// reach the carry buffer.

#define R3_WAVES 4
#define R3_WN 256                  // nonzeros per wave
#define R3_BATCH 6
#define R3_FST 0x80000000u         // record slice field: first nonzero of a fiber

// Makes v opaque to the optimiser once it is loaded. A walk that gathers A
// rows at fiber starts only selects the current fiber's row (avc = new
// fiber ? av[j] : avc); without the barrier the compiler folds that select
// into the conditional gathers (each then starts from a copy of the
// previous one, still in flight, and waits for it). Placed after every
// gather of a batch is issued.
template<typename VT>
__device__ __forceinline__ void r3_opaque(VT & v)
{
  asm volatile("" : "+v"(v));
}

template<typename T>
struct Root3Args {
  uint32_t const * fptr0;          // slice -> first fiber
  uint32_t const * fids0;          // slice -> output row (tiled host CSF)
  uint32_t const * fptr1;          // fiber -> first nonzero
  uint32_t const * fids1;          // fiber -> A row
  uint32_t const * leaf;           // nonzero -> B row
  T const * vals;
  T const * A;
  T const * B;
  T * out;
  uint32_t const * plan;           // per item: slice, fiber containing its first nnz
  uint8_t const * head;            // per item: first slice began earlier
  uint64_t const * bits;           // per 64 nonzeros: slice-start word, fiber-start word
  T * carry;                       // 2 rows per block
  uint32_t * carry_row;
  uint64_t nnz;
  uint32_t nslices;
  uint32_t nfibers;
  uint32_t nitems;
  uint32_t R;
  int rows_are_positions;
};

// Inclusive prefix sum over a wave of 64 lanes.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  int const lane = threadIdx.x & 63;
  #pragma unroll
  for(int o=1; o < 64; o <<= 1) {
    uint32_t const t = __shfl_up(v, o, 64);
    if(lane >= o) v += t;
  }
  return v;
}


// Sum runs of equal rows over entries [e0, e1) (rows[e] == CARRY_NONE are
// empty). The first run is "open" if open_first, the run containing entry
// e_last is open at the end. Open runs go to slot 1 (open at the end) or
// slot 0 (open at the start only) of (orow, oval) -- or, when oval is null,
// to rows of carry (cv + slot*R, the caller's row range); closed runs are
// stored to out. Empty slots report CARRY_NONE in orow.
template<typename T, int V>
__device__ __forceinline__ void
combine_runs(uint32_t const * rows, T const * vals, int ld, uint32_t e0, uint32_t e1,
    bool open_first, uint32_t e_last, uint32_t * orow, T * oval, int old_,
    T * out, uint32_t R, uint32_t col, uint32_t c, bool cok, T * cv = nullptr)
{
  typedef typename VecT<T, V>::type VT;
  uint32_t cur = CARRY_NONE;
  bool cur_first = false, cur_last = false, seen = false;
  uint32_t r0 = CARRY_NONE, r1 = CARRY_NONE;
  VT acc = vzero<VT>();
  for(uint32_t e=e0; e <= e1; ++e) {
    uint32_t const r = (e < e1) ? rows[e] : CARRY_NONE;
    if(e < e1 && r == CARRY_NONE) continue;
    if(r != cur || e == e1) {
      if(cur != CARRY_NONE) {
        int slot = -1;
        if(cur_last) slot = 1;
        else if(cur_first && open_first) slot = 0;
        if(slot < 0) {
          vstore<T, V>(out, cur, R, col, cok, acc);
        } else {
          if(slot == 0) r0 = cur; else r1 = cur;
          if(cok) {
            if(oval != nullptr) *(VT *) (oval + slot * old_ + c * V) = acc;
            else vstore<T, V>(cv, (uint32_t) slot, R, col, cok, acc);
          }
        }
      }
      if(e == e1) break;
      cur = r;
      cur_first = !seen;
      cur_last = false;
      seen = true;
      acc = vzero<VT>();
    }
    if(cok) acc += *(VT const *) (vals + (uint64_t) e * ld + c * V);
    if(e == e_last) cur_last = true;
  }
  orow[0] = r0;
  orow[1] = r1;
}

// Sum of v over the G = 64/L groups of a wave, returned to every lane (lanes
// of one column position c in different groups are L apart). Addition is
// commutative, so every group gets the same bits.
template<typename T, int V, int L>
__device__ __forceinline__ typename VecT<T, V>::type
group_allsum(typename VecT<T, V>::type v)
{
  #pragma unroll
  for(int o=L; o < 64; o <<= 1) {
    if constexpr (V == 1) {
      v += __shfl_xor(v, o, 64);
    } else {
      #pragma unroll
      for(int k=0; k < V; ++k) v[k] += __shfl_xor(v[k], o, 64);
    }
  }
  return v;
}

// Root fix-up: sums the partial rows of slices that spanned several work
// ranges (carry entries, in position order: the entries of one row are
// consecutive). One wave owns a chunk of FIX_CHUNK entries; group g of L
// lanes loads entries g*FIX_CHUNK/G .. (g+1)*FIX_CHUNK/G - 1 at once and sums
// their runs. A run closed on both sides inside a group is added to its row;
// the groups' first and last runs are combined across the wave as in the
// MTTKRP kernels (registers when they all belong to one row, LDS otherwise),
// and the runs that reach the chunk's edges are added with one atomic
// row-add each (a slice spanning P entries costs ~2P/FIX_CHUNK atomics on
// its row). With a second-level buffer (carry2, crow2 not null) the chunk's
// first and last runs go to its two entries there instead (the first run,
// or none, then the last), and a second pass over those entries adds them:
// a heavy row then sees ~4P/FIX_CHUNK^2 atomics instead of ~2P/FIX_CHUNK,
// which long carry buffers need (atomics on one row serialise).
#define FIX_CHUNK 64
// carry entries from which the fix-up takes two passes
static uint64_t g_fix_two_level = 131072;

extern "C" int spd_mttkrp_set_fix_two_level(uint64_t entries)
{
  g_fix_two_level = entries;
  return SPD_OK;
}

template<typename T, int V, int L>
__global__ void __launch_bounds__(MK_THREADS)
k_root_fixup(T const * __restrict__ carry, uint32_t const * __restrict__ crow,
    uint32_t nentries, T * __restrict__ out, uint32_t R, T * __restrict__ carry2,
    uint32_t * __restrict__ crow2)
{
  typedef typename VecT<T, V>::type VT;
  constexpr int G = 64 / L;
  constexpr int EPG = FIX_CHUNK / G;           // entries per group
  constexpr int CW = L * V;
  constexpr int NWB = MK_THREADS / 64;
  __shared__ uint32_t s_erow[NWB][2 * G];
  __shared__ __align__(16) T s_eval[NWB][2 * G][CW];
  __shared__ __align__(16) T s_slot[NWB][2][CW];
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  uint32_t const c = lane % L;
  uint32_t const g = lane / L;
  uint64_t const wc = (uint64_t) blockIdx.x * NWB + wave;
  uint64_t const w0 = wc * FIX_CHUNK;
  if(w0 >= nentries) return;                  // the whole wave leaves together
  bool const two = carry2 != nullptr;
  bool const rowst = lane == 0 && blockIdx.y == 0;
  uint32_t const col = blockIdx.y * CW + c * V;
  bool const cok = col < R;
  uint64_t const e0 = w0 + (uint64_t) g * EPG;
  uint32_t rw[EPG];
  VT v[EPG];
  #pragma unroll
  for(int j=0; j < EPG; ++j) {
    bool const in = e0 + j < nentries;
    rw[j] = in ? crow[e0 + j] : CARRY_NONE;
    // entries past the end read entry w0's row (in bounds, unused)
    v[j] = vload<T, V>(carry, (uint32_t) (in ? e0 + j : w0), R, col, cok);
  }
  // this group's runs: the first (r0, v0) and last (r1, v1) may continue
  // in the neighbouring groups; the ones between are complete
  uint32_t r0 = CARRY_NONE, r1 = CARRY_NONE;
  VT v0 = vzero<VT>(), acc = vzero<VT>();
  bool first = true;
  #pragma unroll
  for(int j=0; j < EPG; ++j) {
    if(rw[j] == CARRY_NONE) continue;
    if(rw[j] != r1) {
      if(r1 != CARRY_NONE) {
        if(first) {
          r0 = r1;
          v0 = acc;
        } else {
          vadd_atomic<T, V>(out, r1, R, col, cok, acc);
        }
        first = false;
      }
      r1 = rw[j];
      acc = vzero<VT>();
    }
    acc += v[j];
  }
  VT const v1 = acc;
  // the chunk's runs across the groups
  uint32_t const x = (r0 != CARRY_NONE) ? r0 : r1;
  uint64_t const m = __ballot(x != CARRY_NONE);
  if(m == 0) {                                // wave-uniform
    if(two && rowst) {
      crow2[2 * wc] = CARRY_NONE;
      crow2[2 * wc + 1] = CARRY_NONE;
    }
    return;
  }
  uint32_t const ref = __shfl(x, __ffsll((unsigned long long) m) - 1, 64);
  bool const same = __all((r0 == CARRY_NONE || r0 == ref) && (r1 == CARRY_NONE || r1 == ref));
  if(same) {
    VT t = vzero<VT>();
    if(r0 != CARRY_NONE) t += v0;
    if(r1 != CARRY_NONE) t += v1;
    t = group_allsum<T, V, L>(t);
    if(two) {
      // one run: it may continue on both sides
      if(g == 0) vstore<T, V>(carry2, (uint32_t) (2 * wc + 1), R, col, cok, t);
      if(rowst) {
        crow2[2 * wc] = CARRY_NONE;
        crow2[2 * wc + 1] = ref;
      }
    } else if(g == 0) {
      vadd_atomic<T, V>(out, ref, R, col, cok, t);
    }
    return;
  }
  s_erow[wave][2 * g] = r0;
  s_erow[wave][2 * g + 1] = r1;
  if(cok) {
    *(VT *) &s_eval[wave][2 * g][c * V] = v0;
    *(VT *) &s_eval[wave][2 * g + 1][c * V] = v1;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if(g == 0) {
    // runs closed inside the chunk are stored (they hold every partial of
    // their row); the chunk's first run and the run of its last entry (it
    // may continue past empty entries into the next chunk) go to the slots
    uint32_t el = 0;
    for(int e=2 * G - 1; e >= 0; --e) {
      if(s_erow[wave][e] != CARRY_NONE) {
        el = (uint32_t) e;
        break;
      }
    }
    uint32_t rows[2];
    combine_runs<T, V>(&s_erow[wave][0], &s_eval[wave][0][0], CW, 0, 2 * G, true, el,
        rows, &s_slot[wave][0][0], CW, out, R, col, c, cok);
    #pragma unroll
    for(int q=0; q < 2; ++q) {
      if(rows[q] != CARRY_NONE && cok) {
        VT const sv = *(VT const *) &s_slot[wave][q][c * V];
        if(two) vstore<T, V>(carry2, (uint32_t) (2 * wc + q), R, col, cok, sv);
        else vadd_atomic<T, V>(out, rows[q], R, col, cok, sv);
      }
    }
    if(two && rowst) {
      crow2[2 * wc] = rows[0];
      crow2[2 * wc + 1] = rows[1];
    }
  }
}

// Entries of the fix-up's second-level buffer for n carry entries (kept
// after the n entries in the carry buffers).
static inline uint64_t fix_level2_entries(uint64_t n)
{
  return 2 * ((n + FIX_CHUNK - 1) / FIX_CHUNK);
}

// One wave per FIX_CHUNK entries; from g_fix_two_level entries on, the
// chunks' edge runs pass through the second-level buffer (carry and crow then hold
// n + fix_level2_entries(n) entries).
template<typename T, int V, int L>
static int launch_fixup(T * carry, uint32_t * crow, uint64_t n, T * out, uint32_t R,
    uint32_t ny, hipStream_t st)
{
  if(n == 0) return SPD_OK;
  uint64_t const waves = spd_cdiv(n, FIX_CHUNK);
  dim3 grid((unsigned) spd_cdiv(waves, MK_THREADS / 64), ny);
  if(n >= g_fix_two_level) {
    T * const carry2 = carry + n * R;
    uint32_t * const crow2 = crow + n;
    hipLaunchKernelGGL((k_root_fixup<T, V, L>), grid, dim3(MK_THREADS), 0, st,
        (T const *) carry, (uint32_t const *) crow, (uint32_t) n, out, R, carry2, crow2);
    SPD_LAUNCH_CHECK();
    carry = carry2;
    crow = crow2;
    n = fix_level2_entries(n);
    dim3 g2((unsigned) spd_cdiv(spd_cdiv(n, FIX_CHUNK), MK_THREADS / 64), ny);
    grid = g2;
  }
  hipLaunchKernelGGL((k_root_fixup<T, V, L>), grid, dim3(MK_THREADS), 0, st,
      (T const *) carry, (uint32_t const *) crow, (uint32_t) n, out, R, (T *) nullptr,
      (uint32_t *) nullptr);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

// HOIST: the fiber's A row is gathered once per fiber (and at the start of
// each item) and multiplied into the slice sum at the fiber's end, as the
// reference factors it; otherwise both rows are gathered per nonzero.
template<typename T, int V, int L, bool HOIST>
__global__ void __launch_bounds__(R3_WAVES * 64, sizeof(T) == 4 ? 5 : 1)
k_root3(Root3Args<T> a)
{
  typedef typename VecT<T, V>::type VT;
  constexpr int G = 64 / L;
  constexpr int S = R3_WN / G;
  constexpr int CW = L * V;
  constexpr int NE = R3_WAVES * G * 2;        // combine entries per block
  constexpr int PER = R3_WN / 64;             // staged nonzeros per lane
  // per nonzero: byte offset of its B row, byte offset of its fiber's A row,
  // local slice index, and (FP32) the value's bits; FP64 values in s_v64
  // records of group g start at g * (S + 1): the groups' concurrent reads of
  // their own records fall in different LDS banks
  __shared__ uint4 s_rec[R3_WAVES][R3_WN + G];
  __shared__ T s_v64[sizeof(T) == 8 ? R3_WAVES : 1][sizeof(T) == 8 ? R3_WN + G : 1];
  __shared__ uint32_t s_srow[R3_WAVES][R3_WN + 1];   // output row of slice S0 + k
  __shared__ uint32_t s_crow[NE];
  // the staging-only array s_arow (A row offset of local fiber k) shares its
  // LDS with the walk's partial rows s_cval (two per item; the block barrier
  // between staging and walking separates the two uses). With the batch of
  // 6 this keeps five blocks (20 waves) per CU.
  constexpr int SZ_AROW = R3_WAVES * (R3_WN + 1) * 4;
  constexpr int SZ_STAGE = SZ_AROW;
  constexpr int SZ_WALK = NE * CW * (int) sizeof(T);
  __shared__ __align__(16) unsigned char s_u[SZ_STAGE > SZ_WALK ? SZ_STAGE : SZ_WALK];
  auto s_arow = reinterpret_cast<uint32_t (*)[R3_WN + 1]>(s_u);
  auto s_cval = reinterpret_cast<T (*)[CW]>(s_u);
  __shared__ uint32_t s_wrow[R3_WAVES][2];
  __shared__ __align__(16) T s_wval[R3_WAVES][2][CW];

  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  uint32_t const c = lane % L;
  uint32_t const g = lane / L;
  uint32_t const R = a.R;
  uint32_t const rowb = R * (uint32_t) sizeof(T);
  uint32_t const col = blockIdx.y * CW + c * V;
  bool const cok = col < R;
  // lanes past the last column read column 0 (in bounds, never stored)
  uint32_t const colb = cok ? col * (uint32_t) sizeof(T) : 0u;

  uint64_t const wid = (uint64_t) blockIdx.x * R3_WAVES + wave;
  uint64_t const W0 = wid * R3_WN;
  uint64_t const W1 = (W0 + R3_WN < a.nnz) ? W0 + R3_WN : a.nnz;
  uint32_t const item0 = (uint32_t) (wid * G);
  uint32_t const wn = (W0 < a.nnz) ? (uint32_t) (W1 - W0) : 0u;

  // ---- stage this wave's range ----
  // Each lane owns PER consecutive positions p = lane*PER + q, inside word
  // lane/16 of the wave's four 64-position bitmap words. The local index of
  // a position's fiber (slice) is the number of fiber (slice) starts at
  // positions 1..p; local index 0 is fiber F0 (slice S0), the one holding
  // the wave's first nonzero. Each position then gets its record.
  uint32_t F0 = 0, S0 = 0;
  if(wn > 0) {
    S0 = a.plan[2ull * item0 + 0];
    F0 = a.plan[2ull * item0 + 1];
    // this lane's PER consecutive leaf ids and values, issued first: one
    // 16-byte load each (two for FP64 values) when all PER lie in the range
    static_assert(PER == 4, "the vector staging loads assume 4 positions per lane");
    uint32_t lf[PER];
    T lv[PER];
    if(lane * PER + PER <= wn) {
      uint4 const l4 = *(uint4 const *) (a.leaf + W0 + lane * PER);
      lf[0] = l4.x;
      lf[1] = l4.y;
      lf[2] = l4.z;
      lf[3] = l4.w;
      if constexpr (sizeof(T) == 4) {
        float4 const v4 = *(float4 const *) (a.vals + W0 + lane * PER);
        lv[0] = v4.x;
        lv[1] = v4.y;
        lv[2] = v4.z;
        lv[3] = v4.w;
      } else {
        double2 const v0 = *(double2 const *) (a.vals + W0 + lane * PER);
        double2 const v1 = *(double2 const *) (a.vals + W0 + lane * PER + 2);
        lv[0] = v0.x;
        lv[1] = v0.y;
        lv[2] = v1.x;
        lv[3] = v1.y;
      }
    } else {
      #pragma unroll
      for(int q=0; q < PER; ++q) {
        uint32_t const p = lane * PER + q;
        lf[q] = (p < wn) ? a.leaf[W0 + p] : 0u;
        lv[q] = (p < wn) ? a.vals[W0 + p] : (T) 0;
      }
    }
    // this lane's slice-start and fiber-start words (16 lanes share them)
    uint4 const bw = *(uint4 const *) (a.bits + 2 * (wid * 4 + (uint64_t) (lane >> 4)));
    uint64_t sw = (uint64_t) bw.x | ((uint64_t) bw.y << 32);
    uint64_t fw = (uint64_t) bw.z | ((uint64_t) bw.w << 32);
    if(lane < 16) {
      // position 0 belongs to F0 and S0 whether or not it starts them
      fw &= ~1ull;
      sw &= ~1ull;
    }
    uint32_t const fcnt = (uint32_t) __popcll(fw);
    uint32_t const scnt = (uint32_t) __popcll(sw);
    uint32_t const fc0 = __builtin_amdgcn_readlane(fcnt, 0), fc1 = __builtin_amdgcn_readlane(fcnt, 16);
    uint32_t const fc2 = __builtin_amdgcn_readlane(fcnt, 32), fc3 = __builtin_amdgcn_readlane(fcnt, 48);
    uint32_t const sc0 = __builtin_amdgcn_readlane(scnt, 0), sc1 = __builtin_amdgcn_readlane(scnt, 16);
    uint32_t const sc2 = __builtin_amdgcn_readlane(scnt, 32), sc3 = __builtin_amdgcn_readlane(scnt, 48);
    uint32_t const wq = (uint32_t) lane >> 4;
    uint32_t const sh = ((uint32_t) lane & 15u) * PER;
    uint64_t const below = (1ull << sh) - 1ull;
    // starts at positions before this lane's first one
    uint32_t fpre = (wq > 0 ? fc0 : 0u) + (wq > 1 ? fc1 : 0u) + (wq > 2 ? fc2 : 0u) +
                    (uint32_t) __popcll(fw & below);
    uint32_t spre = (wq > 0 ? sc0 : 0u) + (wq > 1 ? sc1 : 0u) + (wq > 2 ? sc2 : 0u) +
                    (uint32_t) __popcll(sw & below);
    uint32_t const f4 = (uint32_t) (fw >> sh) & 15u;
    uint32_t const s4 = (uint32_t) (sw >> sh) & 15u;
    // fibers F0 .. F0+nfw-1 and slices S0 .. S0+nsw-1 overlap the range
    uint32_t const nfw = 1u + fc0 + fc1 + fc2 + fc3;
    uint32_t const nsw = 1u + sc0 + sc1 + sc2 + sc3;
    for(uint32_t k=lane; k < nsw; k += 64) {
      s_srow[wave][k] = a.rows_are_positions ? S0 + k : a.fids0[S0 + k];
    }
    for(uint32_t k=lane; k < nfw; k += 64) {
      s_arow[wave][k] = a.fids1[F0 + k] * rowb;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t fm[PER];
    #pragma unroll
    for(int q=0; q < PER; ++q) fm[q] = (f4 >> q) & 1u;
    #pragma unroll
    for(int q=0; q < PER; ++q) {
      fpre += fm[q];
      spre += (s4 >> q) & 1u;
      uint32_t const p = lane * PER + q;
      uint32_t const leaf = lf[q];
      T const v = lv[q];
      uint4 r;
      r.x = leaf * rowb;
      r.y = s_arow[wave][fpre];
      r.z = spre | ((HOIST && fm[q]) ? R3_FST : 0u);
      if constexpr (sizeof(T) == 4) {
        r.w = __float_as_uint(v);
      } else {
        r.w = 0;
        s_v64[wave][p + p / S] = v;
      }
      s_rec[wave][p + p / S] = r;
    }
  }
  __syncthreads();

  // ---- walk this group's item ----
  // Every nonzero contributes v * B[leaf] * A[fiber row] to its slice's
  // accumulator; a change of local slice flushes the finished slice.
  uint32_t const item = item0 + g;
  uint64_t const n0 = (uint64_t) item * S;
  uint32_t const e0 = (wave * G + g) * 2;
  char const * const Bb = (char const *) a.B;
  char const * const Ab = (char const *) a.A;
  // the item's partial slices: H, the first slice when it began before the
  // item and ends inside it or at its end (its sum goes to LDS entry e0 when
  // it ends); T, the last slice when it continues past the item's end
  uint32_t hrow = CARRY_NONE, trow = CARRY_NONE;
  VT tval = vzero<VT>();
  if(n0 < a.nnz) {
    uint64_t const n1 = (n0 + S < a.nnz) ? n0 + S : a.nnz;
    bool head_partial = a.head[item] != 0;
    uint32_t const k0 = (uint32_t) (n0 - W0);
    uint32_t const k1 = (uint32_t) (n1 - W0);
    uint32_t cur_s = s_rec[wave][k0 + k0 / S].z & ~R3_FST;
    VT accs = vzero<VT>();
    VT accf = vzero<VT>();
    VT avc = vzero<VT>();
    auto rec_val = [&](uint4 const & r, uint32_t k) -> T {
      if constexpr (sizeof(T) == 4) {
        return __uint_as_float(r.w);
      } else {
        return s_v64[wave][k + k / S];
      }
    };
    auto flush = [&]() {
      // slice cur_s ended before this nonzero
      uint32_t const orow = s_srow[wave][cur_s];
      if(head_partial) {
        hrow = orow;
        if(cok) *(VT *) &s_cval[e0][c * V] = accs;
      } else {
        vstore<T, V>(a.out, orow, R, col, cok, accs);
      }
      head_partial = false;
      accs = vzero<VT>();
    };
    auto step = [&](uint4 const & r, T v, VT b, VT av, bool nf) {
      if constexpr (HOIST) {
        if(nf) {
          accs += accf * avc;
          accf = vzero<VT>();
          avc = av;
        }
      }
      uint32_t const sl = r.z & ~R3_FST;
      if(sl != cur_s) {
        flush();
        cur_s = sl;
      }
      if constexpr (HOIST) {
        accf += v * b;
      } else {
        accs += v * (b * av);
      }
    };
    for(uint32_t k=k0; k < k1; k += R3_BATCH) {
      if(k + R3_BATCH <= k1) {
        uint4 r[R3_BATCH];
        VT b[R3_BATCH];
        VT av[R3_BATCH];
        bool nf[R3_BATCH];
        #pragma unroll
        for(int j=0; j < R3_BATCH; ++j) r[j] = s_rec[wave][k + j + (k + j) / S];
        #pragma unroll
        for(int j=0; j < R3_BATCH; ++j) {
          b[j] = *(VT const *) (Bb + (r[j].x + colb));
          nf[j] = !HOIST || (r[j].z & R3_FST) != 0u || (k + j == k0);
          av[j] = nf[j] ? *(VT const *) (Ab + (r[j].y + colb)) : vzero<VT>();
        }
        if constexpr (HOIST) {
          #pragma unroll
          for(int j=0; j < R3_BATCH; ++j) r3_opaque(av[j]);
        }
        #pragma unroll
        for(int j=0; j < R3_BATCH; ++j) step(r[j], rec_val(r[j], k + j), b[j], av[j], nf[j]);
      } else {
        for(uint32_t kk=k; kk < k1; ++kk) {
          uint4 const r = s_rec[wave][kk + kk / S];
          bool const nf = !HOIST || (r.z & R3_FST) != 0u || kk == k0;
          VT const b = *(VT const *) (Bb + (r.x + colb));
          VT const av = nf ? *(VT const *) (Ab + (r.y + colb)) : vzero<VT>();
          step(r, rec_val(r, kk), b, av, nf);
        }
      }
    }
    if constexpr (HOIST) accs += accf * avc;
    // is the last slice complete at n1? It is if n1 is the start of a new
    // slice (or the tensor's end)
    bool complete = (n1 == a.nnz);
    if(!complete && n1 < W1) {
      complete = (s_rec[wave][k1 + k1 / S].z & ~R3_FST) != cur_s;
    } else if(!complete) {
      // the item ends at the wave's end: the slice is complete iff the
      // next item's first slice starts there
      complete = a.head[item + 1] == 0;
    }
    uint32_t const orow = s_srow[wave][cur_s];
    if(complete && !head_partial) {
      vstore<T, V>(a.out, orow, R, col, cok, accs);
    } else if(complete) {
      hrow = orow;
      if(cok) *(VT *) &s_cval[e0][c * V] = accs;
    } else {
      trow = orow;
      tval = accs;
    }
  }
  // ---- combine partial slices: per wave, then per block ----
  // Entries (H and T of every item) are in nonzero order, so one slice's
  // entries are adjacent. A run (entries of one row) that may continue past
  // an edge is kept open; a run closed on both sides is complete and stored.
  // Open runs of a wave go to its two wave slots (slot 1 if the run reaches
  // the wave's end, slot 0 if it began before the wave), and the block
  // combines the wave slots the same way into its two carry rows. When all
  // of a level's entries belong to one row (a wave inside one long slice, the
  // common case) they form a single run, summed across groups in registers;
  // otherwise the entries go through LDS to a serial walk over the runs.
  {
    uint32_t const x = (hrow != CARRY_NONE) ? hrow : trow;
    uint64_t const m = __ballot(x != CARRY_NONE);
    uint32_t ref = CARRY_NONE;
    bool same = true;
    if(m != 0) {
      ref = __shfl(x, __ffsll((unsigned long long) m) - 1, 64);
      same = __all((hrow == CARRY_NONE || hrow == ref) && (trow == CARRY_NONE || trow == ref));
    }
    bool const wopen_first = (item0 < a.nitems) && (wn > 0) && a.head[item0] != 0;
    uint32_t last_item = item0 + G - 1;
    if(last_item >= a.nitems) last_item = a.nitems - 1;
    if(same) {
      VT v = vzero<VT>();
      if(hrow != CARRY_NONE && cok) v += *(VT const *) &s_cval[e0][c * V];
      if(trow != CARRY_NONE) v += tval;
      v = group_allsum<T, V, L>(v);
      // the run reaches the wave's end if the last item's T entry is in it
      uint32_t const tlast = (wn > 0) ? __shfl(trow, (int) ((last_item - item0) * L), 64) : CARRY_NONE;
      uint32_t r0 = CARRY_NONE, r1 = CARRY_NONE;
      if(m != 0) {
        if(tlast != CARRY_NONE) r1 = ref;
        else if(wopen_first) r0 = ref;
      }
      if(g == 0) {
        if(r1 != CARRY_NONE) {
          if(cok) *(VT *) &s_wval[wave][1][c * V] = v;
        } else if(r0 != CARRY_NONE) {
          if(cok) *(VT *) &s_wval[wave][0][c * V] = v;
        } else if(m != 0) {
          vstore<T, V>(a.out, ref, R, col, cok, v);
        }
        s_wrow[wave][0] = r0;
        s_wrow[wave][1] = r1;
      }
    } else {
      s_crow[e0] = hrow;
      s_crow[e0 + 1] = trow;
      if(cok) *(VT *) &s_cval[e0 + 1][c * V] = tval;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if(g == 0) {
        uint32_t const e_begin = wave * G * 2;
        uint32_t const e_lastp1 = e_begin + (last_item - item0) * 2 + 1;
        combine_runs<T, V>(s_crow, &s_cval[0][0], CW, e_begin, e_begin + 2 * G,
            wn > 0 ? wopen_first : false, wn > 0 ? e_lastp1 : 0xFFFFFFFFu,
            &s_wrow[wave][0], &s_wval[wave][0][0], CW, a.out, R, col, c, cok);
      }
    }
  }
  __syncthreads();
  if(wave == 0) {
    uint32_t const bitem0 = (uint32_t) blockIdx.x * R3_WAVES * G;
    bool const block_head = (bitem0 < a.nitems) && a.head[bitem0] != 0;
    // the last wave of the block with nonzeros
    uint64_t const bw0 = (uint64_t) blockIdx.x * R3_WAVES;
    uint64_t const nw_total = (a.nnz + R3_WN - 1) / R3_WN;
    uint32_t const lastw = (uint32_t) (((bw0 + R3_WAVES <= nw_total) ? R3_WAVES : nw_total - bw0) - 1);
    T * cv = a.carry + (uint64_t) (2 * blockIdx.x) * R;
    // the block's 2 * R3_WAVES wave slots, entry e = 2 * wave + slot; group g
    // holds entries g, g + G, ...
    constexpr int NB = 2 * R3_WAVES;
    uint32_t x = CARRY_NONE;
    #pragma unroll
    for(int e=(int) g; e < NB; e += G) {
      uint32_t const r = s_wrow[e >> 1][e & 1];
      if(r != CARRY_NONE) x = r;
    }
    uint64_t const m = __ballot(x != CARRY_NONE);
    uint32_t ref = CARRY_NONE;
    bool same_g = true;
    if(m != 0) {
      ref = __shfl(x, __ffsll((unsigned long long) m) - 1, 64);
      #pragma unroll
      for(int e=(int) g; e < NB; e += G) {
        uint32_t const r = s_wrow[e >> 1][e & 1];
        if(r != CARRY_NONE && r != ref) same_g = false;
      }
    }
    bool const same = __all(same_g);
    uint32_t rows[2] = {CARRY_NONE, CARRY_NONE};
    if(same) {
      VT v = vzero<VT>();
      #pragma unroll
      for(int e=(int) g; e < NB; e += G) {
        if(s_wrow[e >> 1][e & 1] != CARRY_NONE) v += *(VT const *) &s_wval[e >> 1][e & 1][c * V];
      }
      v = group_allsum<T, V, L>(v);
      if(m != 0) {
        if(s_wrow[lastw][1] != CARRY_NONE) rows[1] = ref;
        else if(block_head) rows[0] = ref;
      }
      if(g == 0) {
        if(rows[1] != CARRY_NONE) vstore<T, V>(cv, 1u, R, col, cok, v);
        else if(rows[0] != CARRY_NONE) vstore<T, V>(cv, 0u, R, col, cok, v);
        else if(m != 0) vstore<T, V>(a.out, ref, R, col, cok, v);
      }
    } else if(g == 0) {
      combine_runs<T, V>(&s_wrow[0][0], &s_wval[0][0][0], CW, 0, 2 * R3_WAVES,
          block_head, 2 * lastw + 1, rows, nullptr, 0, a.out, R, col, c, cok, cv);
    }
    if(g == 0 && c == 0 && blockIdx.y == 0) {
      a.carry_row[2 * blockIdx.x] = rows[0];
      a.carry_row[2 * blockIdx.x + 1] = rows[1];
    }
  }
}

#include "spd_root3h.hpp"

// ---- staged walk: internal and leaf outputs, and roots of 4+ modes --------
//
// out[r] = sum over nonzeros n whose output-level node has row r of
//          v_n * prod_{levels l != output} U_l[row of n's level-l node].
// CSF order keeps the nonzeros of one output-level node contiguous, so a
// group sums a run of equal output rows in registers and flushes it once.
//
// A wave owns chunks of 256 consecutive nonzeros (persistent blocks stride
// over the chunks). Per chunk it stages into LDS, with coalesced loads, the
// row ids and child pointers of every node of every level that overlaps the
// chunk. Node starts are resolved bottom-up (a node starts where its first
// child starts), marked per position and turned into local node indices with
// wave prefix sums; each position then gets a record: the row of each
// gathered level, the output row (and its LDS slot, if privatised) and the
// value. Groups of L lanes then walk S = 256 / (64/L) positions each in
// batches of B: B records are read from LDS, B x (N-1) factor-row gathers
// are in flight at once, and no step waits on a dependent global load.
//
// Root outputs (rows unique per node): runs closed inside a group's range
// are stored; the first and last runs of the groups of a chunk are combined
// across the wave, and the chunk's first and last run go to a carry buffer
// (two rows per chunk of 256 nonzeros) that k_root_fixup sums in order, so a
// heavy row cut into thousands of chunks costs a few dozen atomics. Internal and leaf outputs:
// rows repeat across nodes, so every run is added, in LDS for the most
// frequent rows (flushed once per block) and with global atomics for the
// rest.
//
// Row-sorted copies. For an internal or leaf output the nonzeros can also
// be walked in the order of their output row: KIND_EXPAND writes each
// position's record (gathered rows and output row) to HBM, a stable radix
// sort by output row orders them, and KIND_COPY walks that copy with the
// root rules (runs are whole output rows: stores and carries, no atomics on
// the output rows).
#define KIND_EXPAND 3
#define KIND_COPY 4
#define WS_WAVES 4
#define WS_WN 256
#define WS_SENT 0x7FFFFFFFu
#define WS_BLOCKS 1024
#define WS_HOT_BYTES (32 * 1024)

template<typename T>
struct WsArgs {
  uint32_t const * fptr[SPD_MAX_MODES];
  uint32_t const * fids[SPD_MAX_MODES];
  uint32_t nfibs[SPD_MAX_MODES];
  T const * vals;
  T const * gm[SPD_MAX_MODES];    // factor of the q-th gathered level, levels in order
  T * out;
  uint32_t const * plan;          // per chunk: node at levels 0..N-2 holding its first nonzero
  uint64_t const * bits;          // node starts: per 64 nonzeros, one word per level < N-1
  int32_t const * hot_slot;       // internal/leaf: LDS slot of an output row (>= nhot: cold)
  uint32_t const * hot_rows;      // internal/leaf: output row of each LDS slot
  T * carry;                      // root/copy: 2 partial rows per range (R wide)
  uint32_t * carry_row;           // root/copy: output row of each partial, or CARRY_NONE
  uint32_t * xg[SPD_MAX_MODES];   // expand: gathered rows per position (q-th level)
  uint32_t * xrow;                // expand: output row per position
  uint32_t const * cg[SPD_MAX_MODES];   // copy: gathered rows in copy order
  uint32_t const * crow;          // copy: output rows (ascending)
  uint32_t nhot;
  uint64_t nnz;
  uint32_t nchunks;
  uint32_t R;
  int depth;                      // output level (internal, leaf and expand kernels)
  int rows_are_positions;         // level-0 row = node index
};

template<typename T, int V>
__device__ __forceinline__ typename VecT<T, V>::type
vload_c(T const * __restrict__ base, uint32_t row, uint32_t R, uint32_t colc)
{
  typedef typename VecT<T, V>::type VT;
  return *(VT const *) (base + (uint64_t) row * R + colc);
}

// Walk positions [k0, k1) of the staged chunk: s_gid[q][k] gathered rows,
// s_orow[k] output rows, s_osl[k] LDS slots (internal/leaf), s_val[k]
// values. KIND is KIND_ROOT (runs unique per row: stores, and the range's
// first and last runs returned in (r0, v0) and (r1, v1), as they may
// continue in the neighbouring ranges; a range with one run returns it as
// the last) or KIND_INTL/KIND_LEAF (every run added).
template<typename T, int NG, int KIND, int V, int L, int B>
__device__ __forceinline__ void
walk_range(WsArgs<T> const & a, uint32_t const (*s_gid)[WS_WN], uint32_t const * s_orow,
    int32_t const * s_osl, T const * s_val, T * hot, uint32_t nhot, uint32_t k0, uint32_t k1,
    uint32_t c, uint32_t col, uint32_t colc, bool cok,
    uint32_t & r0, typename VecT<T, V>::type & v0, uint32_t & r1, typename VecT<T, V>::type & v1)
{
  typedef typename VecT<T, V>::type VT;
  constexpr int CW = L * V;
  uint32_t const R = a.R;
  r0 = CARRY_NONE;
  r1 = CARRY_NONE;
  if(k0 < k1) {
    uint32_t cur = s_orow[k0];
    int32_t csl = -1;
    if constexpr (KIND != KIND_ROOT) csl = s_osl[k0];
    VT acc = vzero<VT>();
    bool first = true;
    auto flush = [&]() {
      if constexpr (KIND == KIND_ROOT) {
        if(first) {
          // may have begun in an earlier range
          r0 = cur;
          v0 = acc;
        } else {
          vstore<T, V>(a.out, cur, R, col, cok, acc);
        }
      } else {
        if(csl >= 0 && (uint32_t) csl < nhot) {
          if(cok) {
            T * p = hot + (uint32_t) csl * CW + c * V;
            if constexpr (V == 1) {
              atomicAdd(p, acc);
            } else {
              #pragma unroll
              for(int k=0; k < V; ++k) atomicAdd(p + k, acc[k]);
            }
          }
        } else {
          vadd_atomic<T, V>(a.out, cur, R, col, cok, acc);
        }
      }
    };
    auto step = [&](uint32_t orow, int32_t osl, VT prod) {
      if(orow != cur) {
        flush();
        first = false;
        cur = orow;
        csl = osl;
        acc = vzero<VT>();
      }
      acc += prod;
    };
    // FP32 root and row-sorted-copy walks of 4+ modes: the rows of the upper
    // gathered levels repeat over consecutive positions; a group reuses the previous
    // position's row when the id is unchanged, and the wave skips that
    // level's gather when no group needs a new row (the FP64 walks, at two
    // waves per SIMD, measured slower with it)
    constexpr bool SKIP = (KIND == KIND_ROOT) && NG >= 3 && sizeof(T) == 4;
    constexpr int NS = SKIP ? NG - 1 : 1;
    uint32_t lgid[NS];
    VT lb[NS];
    #pragma unroll
    for(int q=0; q < NS; ++q) {
      lgid[q] = 0xFFFFFFFFu;
      lb[q] = vzero<VT>();
    }
    uint32_t k = k0;
    for(; k + B <= k1; k += B) {
      uint32_t gid[B][NG];
      uint32_t orw[B];
      int32_t osl[B];
      T v[B];
      #pragma unroll
      for(int j=0; j < B; ++j) {
        #pragma unroll
        for(int q=0; q < NG; ++q) gid[j][q] = s_gid[q][k + j];
        orw[j] = s_orow[k + j];
        osl[j] = -1;
        if constexpr (KIND != KIND_ROOT) osl[j] = s_osl[k + j];
        v[j] = s_val[k + j];
      }
      VT b[B][NG];
      #pragma unroll
      for(int j=0; j < B; ++j) {
        #pragma unroll
        for(int q=0; q < NG; ++q) {
          if constexpr (SKIP) {
            if(q < NS) {
              uint32_t const pg = (j == 0) ? lgid[q] : gid[j - 1][q];
              bool const nd = gid[j][q] != pg;
              b[j][q] = vzero<VT>();
              if(__any(nd)) {
                if(nd) b[j][q] = vload_c<T, V>(a.gm[q], gid[j][q], R, colc);
              }
              continue;
            }
          }
          b[j][q] = vload_c<T, V>(a.gm[q], gid[j][q], R, colc);
        }
      }
      if constexpr (SKIP) {
        // resolve the reused rows once every gather of the batch is issued
        #pragma unroll
        for(int j=0; j < B; ++j) {
          #pragma unroll
          for(int q=0; q < NS; ++q) r3_opaque(b[j][q]);
        }
        #pragma unroll
        for(int j=0; j < B; ++j) {
          #pragma unroll
          for(int q=0; q < NS; ++q) {
            uint32_t const pg = (j == 0) ? lgid[q] : gid[j - 1][q];
            if(gid[j][q] == pg) b[j][q] = (j == 0) ? lb[q] : b[j - 1][q];
          }
        }
        #pragma unroll
        for(int q=0; q < NS; ++q) {
          lgid[q] = gid[B - 1][q];
          lb[q] = b[B - 1][q];
        }
      }
      #pragma unroll
      for(int j=0; j < B; ++j) {
        VT pr = v[j] * b[j][0];
        #pragma unroll
        for(int q=1; q < NG; ++q) pr *= b[j][q];
        step(orw[j], osl[j], pr);
      }
    }
    for(; k < k1; ++k) {
      VT pr = s_val[k] * vload_c<T, V>(a.gm[0], s_gid[0][k], R, colc);
      #pragma unroll
      for(int q=1; q < NG; ++q) pr *= vload_c<T, V>(a.gm[q], s_gid[q][k], R, colc);
      int32_t osl = -1;
      if constexpr (KIND != KIND_ROOT) osl = s_osl[k];
      step(s_orow[k], osl, pr);
    }
    if constexpr (KIND == KIND_ROOT) {
      // the last run may continue in the next range
      r1 = cur;
      v1 = acc;
    } else {
      flush();
    }
  }
}

// FP32 root and copy walks of up to 4 modes run four waves per SIMD (batch
// of 4, no spills); FP64, higher orders and the LDS-heavy internal and leaf
// walks keep their registers.
template<typename T, int N, int KIND, int V, int L, int B>
__global__ void __launch_bounds__(WS_WAVES * 64,
    (sizeof(T) == 4 && N <= 4 && (KIND == KIND_ROOT || KIND == KIND_COPY)) ? 4 : 1)
k_walks(WsArgs<T> a)
{
  constexpr int NL = N - 1;                   // internal levels 0..N-2
  constexpr int NG = N - 1;                   // gathered levels per nonzero
  constexpr int G = 64 / L;
  constexpr int S = WS_WN / G;
  constexpr int CW = L * V;
  constexpr int PER = WS_WN / 64;
  constexpr bool STRUCT = (KIND != KIND_COPY);          // stages the CSF structure
  constexpr bool PRIV = (KIND == KIND_INTL || KIND == KIND_LEAF);
  constexpr int WALK = (KIND == KIND_COPY) ? KIND_ROOT : KIND;
  constexpr int NLS = STRUCT ? NL : 1;
  constexpr int HS = PRIV ? WS_WN : 1;
  __shared__ uint32_t s_row[WS_WAVES][NLS][WS_WN];  // row id of local node k
  // gathered row ids per position
  __shared__ __align__(16) uint32_t s_tmp[WS_WAVES][NL][WS_WN];
  __shared__ __align__(16) uint32_t s_orow[WS_WAVES][WS_WN];   // output row per position
  __shared__ int32_t s_osl[WS_WAVES][HS];           // its LDS slot
  __shared__ int32_t s_nsl[WS_WAVES][HS];           // slot of output-level node k
  __shared__ __align__(16) T s_val[WS_WAVES][KIND == KIND_EXPAND ? 1 : WS_WN];
  extern __shared__ __align__(16) unsigned char ws_dyn[];
  T * hot = (T *) ws_dyn;

  int const lane = threadIdx.x & 63;
  int const wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint32_t const c = lane % L;
  uint32_t const g = lane / L;
  uint32_t const col = blockIdx.y * CW + c * V;
  bool const cok = col < a.R;
  uint32_t const colc = cok ? col : 0u;      // lanes past the last column read column 0
  int const d = (KIND == KIND_ROOT) ? 0 : ((KIND == KIND_LEAF) ? NL : a.depth);
  uint32_t const nhot = PRIV ? a.nhot : 0u;

  if constexpr (PRIV) {
    for(uint32_t x=threadIdx.x; x < nhot * CW; x += WS_WAVES * 64) hot[x] = (T) 0;
    __syncthreads();
  }

  for(uint32_t ch = blockIdx.x * WS_WAVES + wave; ch < a.nchunks; ch += gridDim.x * WS_WAVES) {
    uint64_t const W0 = (uint64_t) ch * WS_WN;
    uint64_t const W1 = (W0 + WS_WN < a.nnz) ? W0 + WS_WN : a.nnz;
    uint32_t const wn = (uint32_t) (W1 - W0);
    if constexpr (!STRUCT) {
      // a row-sorted copy: records are stored as they are walked; each lane
      // copies PER consecutive positions with 16-byte loads and stores when
      // all of them lie in the chunk
      static_assert(PER == 4, "the vector copy assumes 4 positions per lane");
      uint32_t const p0 = (uint32_t) lane * PER;
      if(p0 + PER <= wn) {
        #pragma unroll
        for(int q=0; q < NG; ++q) {
          *(uint4 *) &s_tmp[wave][q][p0] = *(uint4 const *) (a.cg[q] + W0 + p0);
        }
        *(uint4 *) &s_orow[wave][p0] = *(uint4 const *) (a.crow + W0 + p0);
        if constexpr (sizeof(T) == 4) {
          *(float4 *) &s_val[wave][p0] = *(float4 const *) (a.vals + W0 + p0);
        } else {
          *(double2 *) &s_val[wave][p0] = *(double2 const *) (a.vals + W0 + p0);
          *(double2 *) &s_val[wave][p0 + 2] = *(double2 const *) (a.vals + W0 + p0 + 2);
        }
      } else {
        for(uint32_t p=p0; p < wn && p < p0 + PER; ++p) {
          #pragma unroll
          for(int q=0; q < NG; ++q) s_tmp[wave][q][p] = a.cg[q][W0 + p];
          s_orow[wave][p] = a.crow[W0 + p];
          s_val[wave][p] = a.vals[W0 + p];
        }
      }
    } else {
      // Nodes overlapping the chunk, per level l: local node 0 is the node
      // holding W0 (from the plan), local node k >= 1 the k-th node starting
      // at a position 1..wn-1. A position's local node index is the number
      // of starts at positions 1..p, counted in the level's bitmap words
      // (lane owns positions 4*lane..4*lane+3, inside word lane/16).
      uint32_t a0[NL], cnt[NL];
      uint32_t idx[NL][PER];
      {
        uint32_t const wq = (uint32_t) lane >> 4;
        uint32_t const sh = ((uint32_t) lane & 15u) * PER;
        uint64_t const below = (1ull << sh) - 1ull;
        uint64_t const wbase = ((uint64_t) ch * 4 + wq) * NL;
        #pragma unroll
        for(int l=0; l < NL; ++l) {
          a0[l] = a.plan[(uint64_t) ch * NL + l];
          uint2 const b2 = *(uint2 const *) (a.bits + wbase + l);
          uint64_t bw = (uint64_t) b2.x | ((uint64_t) b2.y << 32);
          if(lane < 16) bw &= ~1ull;     // position 0 belongs to local node 0
          uint32_t const bc = (uint32_t) __popcll(bw);
          uint32_t const c0 = __builtin_amdgcn_readlane(bc, 0), c1 = __builtin_amdgcn_readlane(bc, 16);
          uint32_t const c2 = __builtin_amdgcn_readlane(bc, 32), c3 = __builtin_amdgcn_readlane(bc, 48);
          uint32_t pre = (wq > 0 ? c0 : 0u) + (wq > 1 ? c1 : 0u) + (wq > 2 ? c2 : 0u) +
                         (uint32_t) __popcll(bw & below);
          uint32_t const b4 = (uint32_t) (bw >> sh) & 15u;
          #pragma unroll
          for(int q=0; q < PER; ++q) {
            pre += (b4 >> q) & 1u;
            idx[l][q] = pre;
          }
          cnt[l] = 1u + c0 + c1 + c2 + c3;
        }
      }
      #pragma unroll
      for(int l=0; l < NL; ++l) {
        for(uint32_t k=lane; k < cnt[l]; k += 64) {
          uint32_t const node = a0[l] + k;
          uint32_t const row = (l == 0 && a.rows_are_positions) ? node : a.fids[l][node];
          s_row[wave][l][k] = row;
          if constexpr (KIND == KIND_INTL) {
            if(l == d) s_nsl[wave][k] = (nhot > 0) ? a.hot_slot[row] : -1;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // records; each lane's PER leaf ids (and values) with one 16-byte
      // load each when all of them lie in the chunk
      bool const leaf_out = (KIND == KIND_LEAF) || (KIND == KIND_EXPAND && d == NL);
      uint32_t lf[PER];
      T lv[PER];
      {
        uint32_t const p0 = (uint32_t) lane * PER;
        if(p0 + PER <= wn) {
          uint4 const l4 = *(uint4 const *) (a.fids[N - 1] + W0 + p0);
          lf[0] = l4.x;
          lf[1] = l4.y;
          lf[2] = l4.z;
          lf[3] = l4.w;
          if constexpr (KIND != KIND_EXPAND) {
            if constexpr (sizeof(T) == 4) {
              float4 const v4 = *(float4 const *) (a.vals + W0 + p0);
              lv[0] = v4.x;
              lv[1] = v4.y;
              lv[2] = v4.z;
              lv[3] = v4.w;
            } else {
              double2 const v0 = *(double2 const *) (a.vals + W0 + p0);
              double2 const v1 = *(double2 const *) (a.vals + W0 + p0 + 2);
              lv[0] = v0.x;
              lv[1] = v0.y;
              lv[2] = v1.x;
              lv[3] = v1.y;
            }
          }
        } else {
          #pragma unroll
          for(int q=0; q < PER; ++q) {
            uint32_t const p = p0 + q;
            lf[q] = (p < wn) ? a.fids[N - 1][W0 + p] : 0u;
            if constexpr (KIND != KIND_EXPAND) lv[q] = (p < wn) ? a.vals[W0 + p] : (T) 0;
          }
        }
      }
      #pragma unroll
      for(int q=0; q < PER; ++q) {
        uint32_t const p = lane * PER + q;
        if(p < wn) {
          uint32_t const leaf = lf[q];
          uint32_t orow = leaf;
          int32_t osl = -1;
          #pragma unroll
          for(int l=0; l < NL; ++l) {
            uint32_t const r = s_row[wave][l][idx[l][q]];
            int const gq = l - (l > d ? 1 : 0);
            if(l == d) {
              orow = r;
              if constexpr (KIND == KIND_INTL) osl = s_nsl[wave][idx[l][q]];
            } else if constexpr (KIND == KIND_EXPAND) {
              a.xg[gq][W0 + p] = r;
            } else {
              s_tmp[wave][gq][p] = r;
            }
          }
          if(leaf_out) {
            if constexpr (KIND == KIND_LEAF) osl = (nhot > 0) ? a.hot_slot[leaf] : -1;
          } else if constexpr (KIND == KIND_EXPAND) {
            a.xg[NG - 1][W0 + p] = leaf;
          } else {
            s_tmp[wave][NG - 1][p] = leaf;
          }
          if constexpr (KIND == KIND_EXPAND) {
            a.xrow[W0 + p] = orow;
          } else {
            s_orow[wave][p] = orow;
            if constexpr (PRIV) s_osl[wave][p] = osl;
            s_val[wave][p] = lv[q];
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    if constexpr (KIND != KIND_EXPAND) {
      typedef typename VecT<T, V>::type VT;
      uint32_t const k0 = g * S;
      uint32_t const k1 = (k0 + S < wn) ? k0 + S : wn;
      uint32_t r0, r1;
      VT v0 = vzero<VT>(), v1 = vzero<VT>();
      walk_range<T, NG, WALK, V, L, B>(a, s_tmp[wave], s_orow[wave], s_osl[wave], s_val[wave],
          hot, nhot, k0, k1, c, col, colc, cok, r0, v0, r1, v1);
      if constexpr (WALK == KIND_ROOT) {
        // The chunk's edge runs, 2 per group in position order: runs of one
        // row are summed; a run closed on both sides inside the chunk is
        // complete and stored; the chunk's first run (it may have begun in
        // the previous chunk) and its last run (it may continue) go to the
        // chunk's two carry rows for k_root_fixup. Every group with
        // positions ends with a last run, so the chunk's last run is always
        // carried.
        uint32_t const x = (r0 != CARRY_NONE) ? r0 : r1;
        uint64_t const m = __ballot(x != CARRY_NONE);
        uint32_t ref = CARRY_NONE;
        bool same = true;
        if(m != 0) {
          ref = __shfl(x, __ffsll((unsigned long long) m) - 1, 64);
          same = __all((r0 == CARRY_NONE || r0 == ref) && (r1 == CARRY_NONE || r1 == ref));
        }
        T * cv = a.carry + (uint64_t) (2 * ch) * a.R;
        if(same) {
          VT v = vzero<VT>();
          if(r0 != CARRY_NONE) v += v0;
          if(r1 != CARRY_NONE) v += v1;
          v = group_allsum<T, V, L>(v);
          if(g == 0) {
            if(m != 0) vstore<T, V>(cv, 1u, a.R, col, cok, v);
            if(c == 0 && blockIdx.y == 0) {
              a.carry_row[2 * ch] = CARRY_NONE;
              a.carry_row[2 * ch + 1] = ref;
            }
          }
        } else {
          // through LDS: the records are dead once every group has walked
          constexpr int NE = 2 * G;
          static_assert(NE * CW * sizeof(T) <= NL * WS_WN * sizeof(uint32_t),
              "the edge runs fit in the record space");
          uint32_t * s_erow = s_orow[wave];
          T * s_eval = reinterpret_cast<T *>(&s_tmp[wave][0][0]);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          s_erow[2 * g] = r0;
          s_erow[2 * g + 1] = r1;
          if(cok) {
            *(VT *) (s_eval + (2 * g) * CW + c * V) = v0;
            *(VT *) (s_eval + (2 * g + 1) * CW + c * V) = v1;
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if(g == 0) {
            // the last group with positions holds the chunk's last run
            uint32_t const glast = (wn + S - 1) / S - 1;
            uint32_t rows[2];
            combine_runs<T, V>(s_erow, s_eval, CW, 0, NE, true, 2 * glast + 1, rows, nullptr, 0,
                a.out, a.R, col, c, cok, cv);
            if(c == 0 && blockIdx.y == 0) {
              a.carry_row[2 * ch] = rows[0];
              a.carry_row[2 * ch + 1] = rows[1];
            }
          }
        }
      }
    }
    // the next chunk's staging overwrites these records
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  if constexpr (PRIV) {
    __syncthreads();
    // privatised rows accumulated by this block go to HBM once per block
    for(uint32_t x=threadIdx.x; x < nhot * CW; x += WS_WAVES * 64) {
      uint32_t const slot = x / CW;
      uint32_t const cc = x % CW;
      uint32_t const colx = blockIdx.y * CW + cc;
      T const v = hot[x];
      if(colx < a.R && v != (T) 0) {
        vatomic(a.out + (uint64_t) a.hot_rows[slot] * a.R + colx, v);
      }
    }
  }
}

// Row-sorted copy: position i of the copy is CSF position perm[i].
__global__ void k_copy_gather(uint32_t const * __restrict__ perm, uint64_t n, int ng,
    DevPtrs32 xg, DevPtrs32 cg, double const * __restrict__ vals,
    double * __restrict__ cv64, float * __restrict__ cv32)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    uint32_t const p = perm[i];
    for(int q=0; q < ng; ++q) cg.p[q][i] = xg.p[q][p];
    double const v = vals[p];
    cv64[i] = v;
    cv32[i] = (float) v;
  }
}

__global__ void k_u64_to_u32(uint64_t const * __restrict__ k, uint64_t n, uint32_t * __restrict__ o)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) o[i] = (uint32_t) k[i];
}

// ---- plans ----------------------------------------------------------------

struct MttkrpPlan {
  uint32_t S;
  uint32_t nitems;
  uint32_t * nodes;     // nitems x (N-1)
  uint8_t * head;       // nitems
  MttkrpPlan * next;
};

__device__ __forceinline__ uint32_t upper_bound_u32(uint32_t const * a, uint32_t n, uint64_t x)
{
  // first index i in [0, n) with a[i] > x, or n
  uint32_t lo = 0, hi = n;
  while(lo < hi) {
    uint32_t const mid = lo + ((hi - lo) >> 1);
    if((uint64_t) a[mid] > x) hi = mid; else lo = mid + 1;
  }
  return lo;
}

struct PlanArgs {
  uint32_t const * fptr[SPD_MAX_MODES];
  uint32_t nfibs[SPD_MAX_MODES];
  int N;
  uint64_t nnz;
  uint32_t S;
  uint32_t nitems;
  uint32_t * nodes;
  uint8_t * head;
};

__global__ void k_plan(PlanArgs p)
{
  uint32_t const it = blockIdx.x * blockDim.x + threadIdx.x;
  if(it >= p.nitems) return;
  int const NL = p.N - 1;
  uint64_t const n0 = (uint64_t) it * p.S;
  // leaf fiber containing n0: fptr[NL-1][f] <= n0 < fptr[NL-1][f+1]
  uint64_t x = n0;
  uint32_t nd[SPD_MAX_MODES];
  for(int l=NL - 1; l >= 0; --l) {
    uint32_t const f = upper_bound_u32(p.fptr[l], p.nfibs[l] + 1, x) - 1;
    nd[l] = f;
    x = f;
  }
  for(int l=0; l < NL; ++l) p.nodes[(uint64_t) it * NL + l] = nd[l];
  // first nonzero of the slice
  uint64_t s = nd[0];
  for(int l=0; l < NL; ++l) s = p.fptr[l][s];
  p.head[it] = (s < n0) ? 1 : 0;
}

// Node-start bitmaps of a CSF: for the 64 nonzeros of word w, bit j of
// bits[w * (N-1) + l] is set when a node of level l starts at nonzero
// 64w + j (its first nonzero: follow first children down the levels). One
// thread per word; nodes are never empty, so a word holds at most 64 starts
// per level, and the words past the last nonzero are zero.
__global__ void k_start_bits(PlanArgs p, uint64_t nwords, uint64_t * __restrict__ bits)
{
  uint64_t const w = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if(w >= nwords) return;
  int const NL = p.N - 1;
  uint64_t const p0 = w * 64, p1 = p0 + 64;
  for(int l=0; l < NL; ++l) {
    auto start = [&](uint32_t n) -> uint64_t {
      uint64_t x = n;
      for(int m=l; m < NL; ++m) x = p.fptr[m][x];
      return x;
    };
    // first node of level l starting at or after p0
    uint32_t lo = 0, hi = p.nfibs[l];
    while(lo < hi) {
      uint32_t const mid = lo + ((hi - lo) >> 1);
      if(start(mid) < p0) lo = mid + 1; else hi = mid;
    }
    uint64_t b = 0;
    for(uint32_t n=lo; n < p.nfibs[l]; ++n) {
      uint64_t const st = start(n);
      if(st >= p1) break;
      b |= 1ull << (st - p0);
    }
    bits[w * NL + l] = b;
  }
}

// The node-start bitmaps of c (built once, kept with the CSF), padded to
// whole 256-nonzero chunks.
static int get_startbits(spd_csf * c, hipStream_t st)
{
  if(c->startbits != nullptr) return SPD_OK;
  int const NL = c->nmodes - 1;
  uint64_t const nwords = 4 * spd_cdiv(c->nnz, 256);
  if(nwords == 0) return SPD_OK;
  uint64_t * b = nullptr;
  SPD_TRY(spd_malloc((void **) &b, nwords * NL * sizeof(uint64_t)));
  PlanArgs pa = {};
  for(int l=0; l < NL; ++l) {
    pa.fptr[l] = c->u.fptr[l];
    pa.nfibs[l] = (uint32_t) c->u.nfibs[l];
  }
  pa.N = c->nmodes;
  pa.nnz = c->nnz;
  hipLaunchKernelGGL(k_start_bits, dim3((unsigned) spd_cdiv(nwords, 256)), dim3(256), 0, st,
      pa, nwords, b);
  if(hipGetLastError() != hipSuccess) {
    spd_free(b);
    spd_set_error(SPD_EHIP, "node-start bitmap launch failed");
    return SPD_EHIP;
  }
  c->startbits = b;
  return SPD_OK;
}

// Rows of one output level ranked by how many nodes of that level carry
// them (descending): rank[row] < nhot selects the rows a block privatises.
struct HotPlan {
  int level;
  uint32_t nrows;
  int32_t * rank;        // device, nrows
  uint32_t * rows;       // device, nrows, most contributions first
  std::vector<uint32_t> counts;   // host, counts in rank order (first 4096)
  HotPlan * next;
};

// A row-sorted copy of a CSF for one output level: the nonzeros ordered by
// output row (CSF order within a row), with the row of every other level.
struct RowCopy {
  int level;
  int rows_pos;                  // level-0 rows were node positions
  uint32_t * cg[SPD_MAX_MODES];  // gathered rows, levels other than `level` in order
  uint32_t * crow;               // output rows, ascending
  double * v64;
  float * v32;
  RowCopy * next;
};

// Leaf ids of a CSF with its most frequent leaf rows re-encoded for
// k_root3h (RH_HOTF | rank), built once per CSF.
struct RootHot {
  uint32_t * leafx;
  uint32_t nhot;
};

void spd_csf_drop_plans(spd_csf * c)
{
  spd_free(c->startbits);
  c->startbits = nullptr;
  if(c->roothot != nullptr) {
    spd_free(c->roothot->leafx);
    delete c->roothot;
    c->roothot = nullptr;
  }
  MttkrpPlan * p = c->plans;
  while(p) {
    MttkrpPlan * nx = p->next;
    spd_free(p->nodes);
    spd_free(p->head);
    free(p);
    p = nx;
  }
  c->plans = nullptr;
  HotPlan * h = c->hot;
  while(h) {
    HotPlan * nx = h->next;
    spd_free(h->rank);
    spd_free(h->rows);
    delete h;
    h = nx;
  }
  c->hot = nullptr;
  RowCopy * r = c->copies;
  while(r) {
    RowCopy * nx = r->next;
    for(int q=0; q < SPD_MAX_MODES; ++q) spd_free(r->cg[q]);
    spd_free(r->crow);
    spd_free(r->v64);
    spd_free(r->v32);
    delete r;
    r = nx;
  }
  c->copies = nullptr;
}

__global__ void k_count_rows(uint32_t const * __restrict__ ids, uint64_t n,
    uint32_t * __restrict__ cnt)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) atomicAdd(&cnt[ids[i]], 1u);
}

__global__ void k_rank_keys(uint32_t const * __restrict__ cnt, uint32_t n,
    uint64_t * __restrict__ key, uint32_t * __restrict__ idx)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) {
    key[i] = (uint64_t) (0xFFFFFFFFu - cnt[i]);
    idx[i] = i;
  }
}

__global__ void k_rank_scatter(uint32_t const * __restrict__ rows, uint32_t n,
    int32_t * __restrict__ rank)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n && rows[i] < n) rank[rows[i]] = (int32_t) i;
}

static int get_hot(spd_csf * c, int level, HotPlan ** out, hipStream_t st)
{
  for(HotPlan * h = c->hot; h; h = h->next) {
    if(h->level == level) { *out = h; return SPD_OK; }
  }
  uint32_t const nrows = (uint32_t) c->dims[c->perm[level]];
  uint64_t const nids = c->u.nfibs[level];
  uint32_t * cnt = nullptr;
  uint64_t * k0 = nullptr; uint64_t * k1 = nullptr;
  uint32_t * v0 = nullptr; uint32_t * v1 = nullptr;
  HotPlan * h = new HotPlan();
  h->level = level;
  h->nrows = nrows;
  int rc = spd_malloc((void **) &cnt, nrows * 4ull);
  if(rc == SPD_OK) rc = spd_malloc((void **) &k0, nrows * 8ull);
  if(rc == SPD_OK) rc = spd_malloc((void **) &k1, nrows * 8ull);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v0, nrows * 4ull);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v1, nrows * 4ull);
  if(rc == SPD_OK) rc = spd_malloc((void **) &h->rank, nrows * 4ull);
  if(rc == SPD_OK) rc = spd_zero_async(cnt, nrows * 4ull, st);
  if(rc == SPD_OK) {
    unsigned grid = (unsigned) (spd_cdiv(nids, 256) < 8192 ? spd_cdiv(nids, 256) : 8192);
    hipLaunchKernelGGL(k_count_rows, dim3(grid), dim3(256), 0, st, c->u.fids[level], nids, cnt);
    hipLaunchKernelGGL(k_rank_keys, dim3((unsigned) spd_cdiv(nrows, 256)), dim3(256), 0, st,
        cnt, nrows, k0, v0);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
  }
  int which = 0;
  if(rc == SPD_OK) rc = spd_radix_sort(k0, v0, k1, v1, nrows, 32, &which, st);
  uint64_t * sk = which ? k1 : k0;
  uint32_t * sv = which ? v1 : v0;
  if(rc == SPD_OK) {
    hipLaunchKernelGGL(k_rank_scatter, dim3((unsigned) spd_cdiv(nrows, 256)), dim3(256), 0, st,
        sv, nrows, h->rank);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
  }
  if(rc == SPD_OK) {
    uint32_t const keep = nrows < 4096 ? nrows : 4096;
    std::vector<uint64_t> hk(keep);
    rc = spd_d2h(hk.data(), sk, keep * 8ull);
    h->counts.resize(keep);
    for(uint32_t i=0; i < keep; ++i) h->counts[i] = 0xFFFFFFFFu - (uint32_t) hk[i];
  }
  spd_free(cnt);
  spd_free(k0); spd_free(k1);
  if(rc == SPD_OK) {
    h->rows = sv;                  // keep the sorted row ids
    spd_free(which ? v0 : v1);
  } else {
    spd_free(v0); spd_free(v1);
    spd_free(h->rank);
    delete h;
    return rc == SPD_EHIP ? (spd_set_error(SPD_EHIP, "hot-row ranking failed"), rc) : rc;
  }
  h->next = c->hot;
  c->hot = h;
  *out = h;
  return SPD_OK;
}

static int get_plan(spd_csf * c, uint32_t S, MttkrpPlan ** out, hipStream_t st)
{
  for(MttkrpPlan * p = c->plans; p; p = p->next) {
    if(p->S == S) { *out = p; return SPD_OK; }
  }
  MttkrpPlan * p = (MttkrpPlan *) calloc(1, sizeof(MttkrpPlan));
  if(p == nullptr) {
    spd_set_error(SPD_ENOMEM, "host allocation of an MTTKRP plan failed");
    return SPD_ENOMEM;
  }
  p->S = S;
  p->nitems = (uint32_t) spd_cdiv(c->nnz, S);
  int const NL = c->nmodes - 1;
  int rc = spd_malloc((void **) &p->nodes, (uint64_t) p->nitems * NL * sizeof(uint32_t));
  if(rc == SPD_OK) rc = spd_malloc((void **) &p->head, p->nitems);
  if(rc != SPD_OK) {
    spd_free(p->nodes); spd_free(p->head); free(p);
    return rc;
  }
  PlanArgs pa = {};
  for(int l=0; l < NL; ++l) {
    pa.fptr[l] = c->u.fptr[l];
    pa.nfibs[l] = (uint32_t) c->u.nfibs[l];
  }
  pa.N = c->nmodes;
  pa.nnz = c->nnz;
  pa.S = S;
  pa.nitems = p->nitems;
  pa.nodes = p->nodes;
  pa.head = p->head;
  hipLaunchKernelGGL(k_plan, dim3((unsigned) spd_cdiv(p->nitems, 256)), dim3(256), 0, st, pa);
  if(hipGetLastError() != hipSuccess) {
    spd_free(p->nodes); spd_free(p->head); free(p);
    spd_set_error(SPD_EHIP, "plan launch failed");
    return SPD_EHIP;
  }
  p->next = c->plans;
  c->plans = p;
  *out = p;
  return SPD_OK;
}

// ---- dispatch -------------------------------------------------------------

template<typename T, int N, int KIND, int V, int L>
static int launch_walks(WsArgs<T> a, hipStream_t st)
{
  constexpr int CW = L * V;
  constexpr int B = (N <= 4 && sizeof(T) == 8) ? 8 : 4;
  uint32_t const ny = (uint32_t) spd_cdiv(a.R, (uint64_t) CW);
  uint64_t nb = spd_cdiv(a.nchunks, WS_WAVES);
  if(nb > WS_BLOCKS) nb = WS_BLOCKS;
  if(nb == 0) return SPD_OK;
  size_t shmem = 0;
  if constexpr (KIND == KIND_INTL || KIND == KIND_LEAF) {
    uint32_t const cap = WS_HOT_BYTES / (CW * sizeof(T));
    if(a.nhot > cap) a.nhot = cap;
    shmem = (size_t) a.nhot * CW * sizeof(T);
  }
  hipLaunchKernelGGL((k_walks<T, N, KIND, V, L, B>), dim3((unsigned) nb, ny),
      dim3(WS_WAVES * 64), shmem, st, a);
  SPD_LAUNCH_CHECK();
  if constexpr (KIND == KIND_ROOT || KIND == KIND_COPY) {
    SPD_TRY((launch_fixup<T, V, L>(a.carry, a.carry_row, 2ull * a.nchunks, a.out, a.R, ny, st)));
  }
  return SPD_OK;
}

template<typename T, int N, int KIND>
static int dispatch_vl(WsArgs<T> const & a, hipStream_t st)
{
  uint32_t const R = a.R;
  if constexpr (sizeof(T) == 4) {
    if(R % 4 == 0 && R >= 8) {
      uint32_t const lanes = R / 4;
      if(lanes <= 2) return launch_walks<T, N, KIND, 4, 2>(a, st);
      if(lanes <= 4) return launch_walks<T, N, KIND, 4, 4>(a, st);
      if(lanes <= 8) return launch_walks<T, N, KIND, 4, 8>(a, st);
      return launch_walks<T, N, KIND, 4, 16>(a, st);
    }
  } else {
    if(R % 2 == 0 && R >= 8) {
      uint32_t const lanes = R / 2;
      if(lanes <= 4) return launch_walks<T, N, KIND, 2, 4>(a, st);
      if(lanes <= 8) return launch_walks<T, N, KIND, 2, 8>(a, st);
      if(lanes <= 16) return launch_walks<T, N, KIND, 2, 16>(a, st);
      return launch_walks<T, N, KIND, 2, 32>(a, st);
    }
  }
  if(R <= 16) return launch_walks<T, N, KIND, 1, 16>(a, st);
  return launch_walks<T, N, KIND, 1, 64>(a, st);
}

template<typename T, int V, int L, bool HOIST>
static int launch_root3(Root3Args<T> const & a, uint32_t nblocks, hipStream_t st,
    hipEvent_t ev_main_end)
{
  constexpr int CW = L * V;
  dim3 grid(nblocks, (unsigned) spd_cdiv(a.R, (uint64_t) CW));
  hipLaunchKernelGGL((k_root3<T, V, L, HOIST>), grid, dim3(R3_WAVES * 64), 0, st, a);
  SPD_LAUNCH_CHECK();
  if(ev_main_end) SPD_HIP(hipEventRecord(ev_main_end, st));
  return launch_fixup<T, V, L>(a.carry, a.carry_row, 2ull * nblocks, a.out, a.R, grid.y, st);
}

// Lane layout (V columns per lane, L lanes per row chunk) for a rank R.
static void lane_layout(uint32_t R, int dtype, int * V, int * L)
{
  if(dtype == SPD_F32 && R % 4 == 0 && R >= 8) {
    uint32_t const lanes = R / 4;
    *V = 4;
    *L = lanes <= 2 ? 2 : lanes <= 4 ? 4 : lanes <= 8 ? 8 : 16;
  } else if(dtype == SPD_F64 && R % 2 == 0 && R >= 8) {
    uint32_t const lanes = R / 2;
    *V = 2;
    *L = lanes <= 4 ? 4 : lanes <= 8 ? 8 : lanes <= 16 ? 16 : 32;
  } else {
    *V = 1;
    *L = R <= 16 ? 16 : 64;
  }
}

template<typename T, bool HOIST>
static int dispatch_root3_h(Root3Args<T> const & a, uint32_t nblocks, int V, int L,
    hipStream_t st, hipEvent_t ev)
{
  if constexpr (sizeof(T) == 4) {
    if(V == 4) {
      switch(L) {
      case 2: return launch_root3<T, 4, 2, HOIST>(a, nblocks, st, ev);
      case 4: return launch_root3<T, 4, 4, HOIST>(a, nblocks, st, ev);
      case 8: return launch_root3<T, 4, 8, HOIST>(a, nblocks, st, ev);
      default: return launch_root3<T, 4, 16, HOIST>(a, nblocks, st, ev);
      }
    }
  } else {
    if(V == 2) {
      switch(L) {
      case 4: return launch_root3<T, 2, 4, HOIST>(a, nblocks, st, ev);
      case 8: return launch_root3<T, 2, 8, HOIST>(a, nblocks, st, ev);
      case 16: return launch_root3<T, 2, 16, HOIST>(a, nblocks, st, ev);
      default: return launch_root3<T, 2, 32, HOIST>(a, nblocks, st, ev);
      }
    }
  }
  if(L == 16) return launch_root3<T, 1, 16, HOIST>(a, nblocks, st, ev);
  return launch_root3<T, 1, 64, HOIST>(a, nblocks, st, ev);
}

template<typename T>
static int dispatch_root3(Root3Args<T> const & a, uint32_t nblocks, int V, int L,
    hipStream_t st, hipEvent_t ev, bool hoist)
{
  if(hoist) return dispatch_root3_h<T, true>(a, nblocks, V, L, st, ev);
  return dispatch_root3_h<T, false>(a, nblocks, V, L, st, ev);
}

template<typename T, int KIND>
static int dispatch_n(int N, WsArgs<T> const & a, hipStream_t st)
{
  switch(N) {
  case 3: return dispatch_vl<T, 3, KIND>(a, st);
  case 4: return dispatch_vl<T, 4, KIND>(a, st);
  case 5: return dispatch_vl<T, 5, KIND>(a, st);
  case 6: return dispatch_vl<T, 6, KIND>(a, st);
  case 7: return dispatch_vl<T, 7, KIND>(a, st);
  case 8: return dispatch_vl<T, 8, KIND>(a, st);
  default: break;
  }
  spd_set_error(SPD_EINPUT, "unsupported number of modes %d", N);
  return SPD_EINPUT;
}

// Kernel timing: when enabled, a pair of HIP events is recorded around every
// walker launch on the stream it runs on, into a ring of SPD_NEV pairs.
#define SPD_NEV 4096
static int g_timing = 0;
static hipEvent_t g_ev[SPD_NEV][2];
static int g_ev_made = 0;
static uint64_t g_ev_count = 0;

extern "C" int spd_mttkrp_timing(int enable)
{
  if(enable && !g_ev_made) {
    for(int i=0; i < SPD_NEV; ++i) {
      SPD_HIP(hipEventCreate(&g_ev[i][0]));
      SPD_HIP(hipEventCreate(&g_ev[i][1]));
    }
    g_ev_made = 1;
  }
  g_timing = enable;
  g_ev_count = 0;
  return SPD_OK;
}

static double ev_ms(uint64_t k)
{
  float ms = -1.f;
  int const s = (int) (k % SPD_NEV);
  if(hipEventSynchronize(g_ev[s][1]) != hipSuccess) return -1.0;
  if(hipEventElapsedTime(&ms, g_ev[s][0], g_ev[s][1]) != hipSuccess) return -1.0;
  return (double) ms;
}

extern "C" double spd_mttkrp_last_kernel_ms(void)
{
  if(g_ev_count == 0) return -1.0;
  return ev_ms(g_ev_count - 1);
}

// Durations (ms) of the last min(max, recorded, SPD_NEV) timed launches, in
// launch order. Returns how many were written.
extern "C" int spd_mttkrp_kernel_times(double * ms, int max)
{
  uint64_t n = g_ev_count < (uint64_t) SPD_NEV ? g_ev_count : (uint64_t) SPD_NEV;
  if((uint64_t) max < n) n = (uint64_t) max;
  uint64_t const first = g_ev_count - n;
  for(uint64_t k=0; k < n; ++k) ms[k] = ev_ms(first + k);
  return (int) n;
}

static int ensure_vals32(spd_csf * c)
{
  if(c->vals32 != nullptr) return SPD_OK;
  SPD_TRY(spd_malloc((void **) &c->vals32, c->nnz * sizeof(float)));
  SPD_TRY(spd_f64_to_f32(c->u.vals, c->vals32, c->nnz, nullptr));
  return SPD_OK;
}

// ---- fixed scratch (the allocation-free entry point) ----------------------

// A caller-side scratch region: when a run_* function gets one, it builds
// nothing (plans, hot-row tables and row-sorted copies must exist) and
// carves its partial-row buffers from the region instead of allocating.
struct MkFixed {
  char * p;
  size_t cap;
};

static MttkrpPlan * find_plan(spd_csf * c, uint32_t S)
{
  for(MttkrpPlan * p = c->plans; p; p = p->next) if(p->S == S) return p;
  return nullptr;
}

static HotPlan * find_hot(spd_csf * c, int level)
{
  for(HotPlan * h = c->hot; h; h = h->next) if(h->level == level) return h;
  return nullptr;
}

static RowCopy * find_copy(spd_csf * c, int level, int rows_pos)
{
  for(RowCopy * r = c->copies; r; r = r->next) {
    if(r->level == level && r->rows_pos == rows_pos) return r;
  }
  return nullptr;
}

static int not_prepared(char const * what)
{
  spd_set_error(SPD_EINPUT, "allocation-free MTTKRP: the CSF has no %s for this mode, rank "
      "and precision yet (one splatt_amd_mttkrp_dev call with the same arguments builds it)",
      what);
  return SPD_EINPUT;
}

// partial rows (cv_bytes) and their row ids (cr_bytes) from the region
static int fixed_carve(MkFixed const * fx, size_t cv_bytes, size_t cr_bytes, void ** cv,
    void ** cr)
{
  size_t const a = (cv_bytes + 255) & ~(size_t) 255;
  if(a + cr_bytes > fx->cap) {
    spd_set_error(SPD_EINPUT, "allocation-free MTTKRP: its partial rows need %zu bytes, "
        "above the %zu bytes of its scratch", a + cr_bytes, fx->cap);
    return SPD_EINPUT;
  }
  *cv = fx->p;
  *cr = fx->p + a;
  return SPD_OK;
}

// ---- persistent root kernel with hot leaf rows (spd_root3h.hpp) ----------

#define RH_NHOT_ENC 1024           // leaf ranks encoded per CSF

// 0: k_root3; 1: k_root3h with hot leaf rows in LDS; 2: k_root3h with
// every row from HBM; 3: k_root3 gathering A once per fiber
static int g_root_variant = 0;

extern "C" int spd_mttkrp_set_root_variant(int v)
{
  g_root_variant = v;
  return SPD_OK;
}

static int cu_count(void)
{
  static int n = 0;
  if(n == 0) {
    int dev = 0, v = 0;
    if(hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0) {
      n = v;
    } else {
      n = 256;
    }
  }
  return n;
}

// The CSF's leaf ids with the RH_NHOT_ENC most frequent leaf rows (those
// with at least two nonzeros) encoded as RH_HOTF | rank.
static int get_roothot(spd_csf * c, RootHot ** out, hipStream_t st)
{
  if(c->roothot != nullptr) {
    *out = c->roothot;
    return SPD_OK;
  }
  HotPlan * hp = nullptr;
  SPD_TRY(get_hot(c, c->nmodes - 1, &hp, st));
  RootHot * rh = new RootHot();
  uint32_t nhot = 0;
  while(nhot < hp->counts.size() && nhot < RH_NHOT_ENC && hp->counts[nhot] >= 2) ++nhot;
  rh->nhot = nhot;
  rh->leafx = nullptr;
  int rc = spd_malloc((void **) &rh->leafx, c->nnz * sizeof(uint32_t));
  if(rc == SPD_OK) {
    unsigned const grid = (unsigned) (spd_cdiv(c->nnz, 256) < 8192 ? spd_cdiv(c->nnz, 256) : 8192);
    hipLaunchKernelGGL(k_encode_hot, dim3(grid), dim3(256), 0, st, c->u.fids[c->nmodes - 1],
        c->nnz, hp->rank, nhot, rh->leafx);
    if(hipGetLastError() != hipSuccess) {
      spd_set_error(SPD_EHIP, "hot leaf encoding failed");
      rc = SPD_EHIP;
    }
  }
  if(rc != SPD_OK) {
    spd_free(rh->leafx);
    delete rh;
    return rc;
  }
  c->roothot = rh;
  *out = rh;
  return SPD_OK;
}

template<typename T, int V, int L, int NW, bool HOT>
static int launch_root3h(Root3hArgs<T> h, hipStream_t st, hipEvent_t ev_main_end)
{
  constexpr int CW = L * V;
  uint64_t const nw_need = spd_cdiv(h.nchunks, 1);
  uint64_t nb = spd_cdiv(nw_need, NW);
  if(nb > (uint64_t) cu_count()) nb = (uint64_t) cu_count();
  if(nb < 1) nb = 1;
  h.cpw = (uint32_t) spd_cdiv(h.nchunks, nb * NW);
  dim3 grid((unsigned) nb, (unsigned) spd_cdiv(h.b.R, (uint64_t) CW));
  hipLaunchKernelGGL((k_root3h<T, V, L, NW, HOT>), grid, dim3(NW * 64), 0, st, h);
  SPD_LAUNCH_CHECK();
  if(ev_main_end) SPD_HIP(hipEventRecord(ev_main_end, st));
  return launch_fixup<T, V, L>(h.b.carry, h.b.carry_row, 2ull * nb * NW, h.b.out, h.b.R, grid.y,
      st);
}

// waves per block of k_root3h: one block per CU, as many waves as its LDS
// (hot rows + per-wave staging) allows
#define RH_NW_F32 16
#define RH_NW_F64 12

// the (V, L) layouts k_root3h is built for; others use k_root3
static bool root3h_layout(int dtype, int V, int L)
{
  if(dtype == SPD_F32) return V == 4 && (L == 4 || L == 8 || L == 16);
  return V == 2 && (L == 8 || L == 16 || L == 32);
}

template<typename T, bool HOT>
static int dispatch_root3h_h(Root3hArgs<T> const & h, int V, int L, hipStream_t st, hipEvent_t ev)
{
  if constexpr (sizeof(T) == 4) {
    (void) V;
    switch(L) {
    case 4: return launch_root3h<T, 4, 4, RH_NW_F32, HOT>(h, st, ev);
    case 8: return launch_root3h<T, 4, 8, RH_NW_F32, HOT>(h, st, ev);
    default: return launch_root3h<T, 4, 16, RH_NW_F32, HOT>(h, st, ev);
    }
  } else {
    (void) V;
    switch(L) {
    case 8: return launch_root3h<T, 2, 8, RH_NW_F64, HOT>(h, st, ev);
    case 16: return launch_root3h<T, 2, 16, RH_NW_F64, HOT>(h, st, ev);
    default: return launch_root3h<T, 2, 32, RH_NW_F64, HOT>(h, st, ev);
    }
  }
}

template<typename T>
static int dispatch_root3h(Root3hArgs<T> const & h, int V, int L, hipStream_t st, hipEvent_t ev,
    bool hot)
{
  if(hot) return dispatch_root3h_h<T, true>(h, V, L, st, ev);
  return dispatch_root3h_h<T, false>(h, V, L, st, ev);
}

template<typename T>
static int run_root3h(spd_csf * c, uint64_t R, void * const * mats, void * out,
    int rows_pos, hipStream_t st, int V, int L, MkFixed const * fx)
{
  uint32_t const G = 64 / L;
  uint32_t const S = R3_WN / G;
  MttkrpPlan * plan = nullptr;
  RootHot * rh = nullptr;
  HotPlan * hp = nullptr;
  if(fx == nullptr) {
    SPD_TRY(get_plan(c, S, &plan, spd_hs(nullptr)));
    SPD_TRY(get_roothot(c, &rh, spd_hs(nullptr)));
    SPD_TRY(get_hot(c, c->nmodes - 1, &hp, spd_hs(nullptr)));
  } else {
    plan = find_plan(c, S);
    rh = c->roothot;
    hp = find_hot(c, c->nmodes - 1);
    if(plan == nullptr || rh == nullptr || hp == nullptr) {
      return not_prepared("launch plan or hot leaf-row table");
    }
  }
  if(st != spd_hs(nullptr)) {
    SPD_HIP(hipStreamSynchronize(spd_hs(nullptr)));
  }
  uint32_t const NW = (sizeof(T) == 4) ? RH_NW_F32 : RH_NW_F64;
  uint64_t const nchunks = spd_cdiv(c->nnz, R3_WN);
  uint64_t nb = spd_cdiv(nchunks, NW);
  if(nb > (uint64_t) cu_count()) nb = (uint64_t) cu_count();
  if(nb < 1) nb = 1;
  // the carry buffers belong to this call when it runs on a caller's stream
  SpdCallScratch scratch(fx ? spd_hs(nullptr) : st);
  void * cv = nullptr;
  void * cr = nullptr;
  if(fx != nullptr) {
    SPD_TRY(fixed_carve(fx, 2ull * nb * NW * R * sizeof(T), 2ull * nb * NW * sizeof(uint32_t),
        &cv, &cr));
  } else {
    SPD_TRY(spd_scratch(2, 2ull * nb * NW * R * sizeof(T), &cv));
    SPD_TRY(spd_scratch(3, 2ull * nb * NW * sizeof(uint32_t), &cr));
  }
  Root3hArgs<T> h = {};
  Root3Args<T> & a = h.b;
  a.fptr0 = c->u.fptr[0];
  a.fids0 = c->u.fids[0];
  a.fptr1 = c->u.fptr[1];
  a.fids1 = c->u.fids[1];
  a.leaf = c->u.fids[2];
  a.vals = (sizeof(T) == 4) ? (T const *) c->vals32 : (T const *) c->u.vals;
  a.A = (T const *) mats[c->perm[1]];
  a.B = (T const *) mats[c->perm[2]];
  a.out = (T *) out;
  a.plan = plan->nodes;
  a.head = plan->head;
  a.bits = c->startbits;
  a.carry = (T *) cv;
  a.carry_row = (uint32_t *) cr;
  a.nnz = c->nnz;
  a.nslices = (uint32_t) c->u.nfibs[0];
  a.nfibers = (uint32_t) c->u.nfibs[1];
  a.nitems = plan->nitems;
  a.R = (uint32_t) R;
  a.rows_are_positions = rows_pos;
  h.leafx = rh->leafx;
  h.hot_rows = hp->rows;
  h.nhot = rh->nhot;
  h.nchunks = (uint32_t) nchunks;
  SPD_TRY(spd_zero_async(out, c->dims[c->perm[0]] * R * sizeof(T), st));
  int const slot = (int) (g_ev_count % SPD_NEV);
  if(g_timing) SPD_HIP(hipEventRecord(g_ev[slot][0], st));
  int const rc = dispatch_root3h<T>(h, V, L, st, g_timing ? g_ev[slot][1] : (hipEvent_t) nullptr,
      g_root_variant == 1);
  if(g_timing) ++g_ev_count;
  return rc;
}

template<typename T>
static int run_root3(spd_csf * c, uint64_t R, void * const * mats, void * out,
    int rows_pos, hipStream_t st, int dtype, MkFixed const * fx)
{
  int V = 1, L = 16;
  lane_layout((uint32_t) R, dtype, &V, &L);
  // the persistent kernel encodes hot rows in bit 31 of B byte offsets and
  // needs at least one chunk per wave of a full grid to pay for its LDS fill
  if((g_root_variant == 1 || g_root_variant == 2) && root3h_layout(dtype, V, L) &&
      c->dims[c->perm[2]] * R * sizeof(T) < 0x80000000ull && c->nnz < 0xFFFFFFFFull &&
      c->nnz >= (uint64_t) R3_WN * 64) {
    return run_root3h<T>(c, R, mats, out, rows_pos, st, V, L, fx);
  }
  uint32_t const G = 64 / L;
  uint32_t const S = R3_WN / G;
  MttkrpPlan * plan = nullptr;
  if(fx == nullptr) {
    SPD_TRY(get_plan(c, S, &plan, spd_hs(nullptr)));
    SPD_TRY(get_startbits(c, spd_hs(nullptr)));
  } else if((plan = find_plan(c, S)) == nullptr || c->startbits == nullptr) {
    return not_prepared("launch plan");
  }
  if(st != spd_hs(nullptr)) {
    SPD_HIP(hipStreamSynchronize(spd_hs(nullptr)));
  }
  // the carry buffers belong to this call when it runs on a caller's stream
  SpdCallScratch scratch(fx ? spd_hs(nullptr) : st);
  uint64_t const nwaves = spd_cdiv(c->nnz, R3_WN);
  uint32_t const nblocks = (uint32_t) spd_cdiv(nwaves, R3_WAVES);
  // two partial rows per block, then the fix-up's second level
  uint64_t const nent = 2ull * nblocks + fix_level2_entries(2ull * nblocks);
  void * cv = nullptr;
  void * cr = nullptr;
  if(fx != nullptr) {
    SPD_TRY(fixed_carve(fx, nent * R * sizeof(T), nent * sizeof(uint32_t), &cv, &cr));
  } else {
    SPD_TRY(spd_scratch(2, nent * R * sizeof(T), &cv));
    SPD_TRY(spd_scratch(3, nent * sizeof(uint32_t), &cr));
  }
  Root3Args<T> a = {};
  a.fptr0 = c->u.fptr[0];
  a.fids0 = c->u.fids[0];
  a.fptr1 = c->u.fptr[1];
  a.fids1 = c->u.fids[1];
  a.leaf = c->u.fids[2];
  a.vals = (sizeof(T) == 4) ? (T const *) c->vals32 : (T const *) c->u.vals;
  a.A = (T const *) mats[c->perm[1]];
  a.B = (T const *) mats[c->perm[2]];
  a.out = (T *) out;
  a.plan = plan->nodes;
  a.head = plan->head;
  a.bits = c->startbits;
  a.carry = (T *) cv;
  a.carry_row = (uint32_t *) cr;
  a.nnz = c->nnz;
  a.nslices = (uint32_t) c->u.nfibs[0];
  a.nfibers = (uint32_t) c->u.nfibs[1];
  a.nitems = plan->nitems;
  a.R = (uint32_t) R;
  a.rows_are_positions = rows_pos;
  SPD_TRY(spd_zero_async(out, c->dims[c->perm[0]] * R * sizeof(T), st));
  // timed: the main kernel only (the carry fix-up is a separate, small launch)
  int const slot = (int) (g_ev_count % SPD_NEV);
  if(g_timing) SPD_HIP(hipEventRecord(g_ev[slot][0], st));
  int const rc = dispatch_root3<T>(a, nblocks, V, L, st,
      g_timing ? g_ev[slot][1] : (hipEvent_t) nullptr, g_root_variant == 3);
  if(g_timing) ++g_ev_count;
  return rc;
}

static int g_rowcopy = 1;

extern "C" int spd_mttkrp_set_copies(int enable)
{
  g_rowcopy = enable ? 1 : 0;
  return SPD_OK;
}

__global__ void k_sort_init(uint32_t const * __restrict__ row, uint64_t n,
    uint64_t * __restrict__ key, uint32_t * __restrict__ val)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    key[i] = row[i];
    val[i] = (uint32_t) i;
  }
}

template<int N>
static int launch_expand(WsArgs<double> const & a, hipStream_t st)
{
  uint64_t nb = spd_cdiv(a.nchunks, WS_WAVES);
  if(nb > WS_BLOCKS) nb = WS_BLOCKS;
  if(nb == 0) return SPD_OK;
  hipLaunchKernelGGL((k_walks<double, N, KIND_EXPAND, 1, 64, 1>), dim3((unsigned) nb),
      dim3(WS_WAVES * 64), 0, st, a);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

static int bits_of(uint64_t v)
{
  int b = 1;
  while(b < 64 && (v >> b) != 0) ++b;
  return b;
}

// The row-sorted copy of c for output level `level`, built on first use on
// stream st. *out is null (and the call succeeds) when device memory would
// run short: the caller then walks the CSF itself.
static int get_copy(spd_csf * c, int level, int rows_pos, RowCopy ** out, hipStream_t st)
{
  *out = nullptr;
  for(RowCopy * r = c->copies; r; r = r->next) {
    if(r->level == level && r->rows_pos == rows_pos) { *out = r; return SPD_OK; }
  }
  int const N = c->nmodes;
  int const NG = N - 1;
  uint64_t const n = c->nnz;
  if(n == 0 || n >= 0xFFFFFFFFull) return SPD_OK;
  size_t const keep = (size_t) n * (NG * 4 + 4 + 8 + 4);
  size_t const temp = (size_t) n * (NG * 4 + 4 + 2 * 8 + 2 * 4);
  size_t fr = 0, tot = 0;
  SPD_HIP(hipMemGetInfo(&fr, &tot));
  if(fr < 2 * (keep + temp) + (1ull << 30)) return SPD_OK;
  MttkrpPlan * plan = nullptr;
  SPD_TRY(get_plan(c, WS_WN, &plan, st));
  SPD_TRY(get_startbits(c, st));
  RowCopy * r = new RowCopy();
  r->level = level;
  r->rows_pos = rows_pos;
  uint32_t * xg[SPD_MAX_MODES] = {};
  uint32_t * xrow = nullptr;
  uint64_t * k0 = nullptr; uint64_t * k1 = nullptr;
  uint32_t * v0 = nullptr; uint32_t * v1 = nullptr;
  int rc = SPD_OK;
  for(int q=0; q < NG && rc == SPD_OK; ++q) rc = spd_malloc((void **) &xg[q], n * 4);
  for(int q=0; q < NG && rc == SPD_OK; ++q) rc = spd_malloc((void **) &r->cg[q], n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &xrow, n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &r->crow, n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &r->v64, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &r->v32, n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &k0, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &k1, n * 8);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v0, n * 4);
  if(rc == SPD_OK) rc = spd_malloc((void **) &v1, n * 4);
  if(rc == SPD_OK) {
    WsArgs<double> a = {};
    for(int l=0; l < N - 1; ++l) {
      a.fptr[l] = c->u.fptr[l];
      a.nfibs[l] = (uint32_t) c->u.nfibs[l];
    }
    for(int l=0; l < N; ++l) a.fids[l] = c->u.fids[l];
    for(int q=0; q < NG; ++q) a.xg[q] = xg[q];
    a.xrow = xrow;
    a.plan = plan->nodes;
    a.bits = c->startbits;
    a.nnz = n;
    a.nchunks = plan->nitems;
    a.depth = level;
    a.rows_are_positions = rows_pos;
    switch(N) {
    case 3: rc = launch_expand<3>(a, st); break;
    case 4: rc = launch_expand<4>(a, st); break;
    case 5: rc = launch_expand<5>(a, st); break;
    case 6: rc = launch_expand<6>(a, st); break;
    case 7: rc = launch_expand<7>(a, st); break;
    case 8: rc = launch_expand<8>(a, st); break;
    default: rc = SPD_EINPUT; break;
    }
  }
  int which = 0;
  if(rc == SPD_OK) {
    unsigned const grid = (unsigned) (spd_cdiv(n, 256) < 8192 ? spd_cdiv(n, 256) : 8192);
    hipLaunchKernelGGL(k_sort_init, dim3(grid), dim3(256), 0, st, xrow, n, k0, v0);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
    if(rc == SPD_OK) rc = spd_radix_sort(k0, v0, k1, v1, n, bits_of(c->dims[c->perm[level]]),
        &which, st);
  }
  if(rc == SPD_OK) {
    unsigned const grid = (unsigned) (spd_cdiv(n, 256) < 8192 ? spd_cdiv(n, 256) : 8192);
    DevPtrs32 src = {}, dst = {};
    for(int q=0; q < NG; ++q) {
      src.p[q] = xg[q];
      dst.p[q] = r->cg[q];
    }
    hipLaunchKernelGGL(k_u64_to_u32, dim3(grid), dim3(256), 0, st, which ? k1 : k0, n, r->crow);
    hipLaunchKernelGGL(k_copy_gather, dim3(grid), dim3(256), 0, st, which ? v1 : v0, n, NG,
        src, dst, c->u.vals, r->v64, r->v32);
    if(hipGetLastError() != hipSuccess) rc = SPD_EHIP;
  }
  if(rc == SPD_OK) rc = (hipStreamSynchronize(st) == hipSuccess) ? SPD_OK : SPD_EHIP;
  for(int q=0; q < NG; ++q) spd_free(xg[q]);
  spd_free(xrow);
  spd_free(k0); spd_free(k1); spd_free(v0); spd_free(v1);
  if(rc != SPD_OK) {
    for(int q=0; q < SPD_MAX_MODES; ++q) spd_free(r->cg[q]);
    spd_free(r->crow); spd_free(r->v64); spd_free(r->v32);
    delete r;
    // short of device memory after all: walk the CSF instead
    if(rc == SPD_ENOMEM) return SPD_OK;
    if(rc == SPD_EHIP) spd_set_error(SPD_EHIP, "row-sorted copy build failed");
    return rc;
  }
  r->next = c->copies;
  c->copies = r;
  *out = r;
  return SPD_OK;
}

template<typename T>
static int run_mttkrp(spd_csf * c, int mode, uint64_t R, void * const * mats,
    void * out, int rows_pos, hipStream_t st, int dtype, MkFixed const * fx)
{
  int const N = c->nmodes;
  int depth = -1;
  for(int l=0; l < N; ++l) if(c->perm[l] == mode) depth = l;
  if(depth < 0) {
    spd_set_error(SPD_EINPUT, "mode %d not in tensor", mode);
    return SPD_EINPUT;
  }
  if(c->nnz == 0) {
    // no nonzeros: the output is zero and no kernel has work
    return spd_zero_async(out, c->dims[mode] * R * sizeof(T), st);
  }
  // the 3-mode root kernel addresses factor rows with 32-bit byte offsets
  uint64_t const lim = 0xFFFFFFFFull;
  if(N == 3 && depth == 0 && c->dims[c->perm[1]] * R * sizeof(T) <= lim &&
      c->dims[c->perm[2]] * R * sizeof(T) <= lim) {
    return run_root3<T>(c, R, mats, out, rows_pos, st, dtype, fx);
  }
  MttkrpPlan * plan = nullptr;
  // plans are built on the library stream; make the caller's stream wait
  if(fx == nullptr) {
    SPD_TRY(get_plan(c, WS_WN, &plan, spd_hs(nullptr)));
    SPD_TRY(get_startbits(c, spd_hs(nullptr)));
  } else if((plan = find_plan(c, WS_WN)) == nullptr || c->startbits == nullptr) {
    return not_prepared("launch plan");
  }
  WsArgs<T> a = {};
  for(int l=0; l < N - 1; ++l) {
    a.fptr[l] = c->u.fptr[l];
    a.nfibs[l] = (uint32_t) c->u.nfibs[l];
  }
  for(int l=0; l < N; ++l) a.fids[l] = c->u.fids[l];
  // gathered levels: every level but the output level, in level order
  for(int l=0, q=0; l < N; ++l) {
    if(l != depth) a.gm[q++] = (T const *) mats[c->perm[l]];
  }
  a.vals = (sizeof(T) == 4) ? (T const *) c->vals32 : (T const *) c->u.vals;
  a.out = (T *) out;
  a.plan = plan->nodes;
  a.bits = c->startbits;
  a.nnz = c->nnz;
  a.nchunks = plan->nitems;
  a.R = (uint32_t) R;
  a.depth = depth;
  a.rows_are_positions = rows_pos;
  RowCopy * cp = nullptr;
  if(depth > 0 && g_rowcopy) {
    if(fx == nullptr) SPD_TRY(get_copy(c, depth, rows_pos, &cp, spd_hs(nullptr)));
    else cp = find_copy(c, depth, rows_pos);
  }
  if(cp != nullptr) {
    for(int q=0; q < N - 1; ++q) a.cg[q] = cp->cg[q];
    a.crow = cp->crow;
    a.vals = (sizeof(T) == 4) ? (T const *) cp->v32 : (T const *) cp->v64;
  }
  HotPlan * hp = nullptr;
  if(depth > 0 && cp == nullptr) {
    if(fx == nullptr) SPD_TRY(get_hot(c, depth, &hp, spd_hs(nullptr)));
    else if((hp = find_hot(c, depth)) == nullptr) return not_prepared("hot-row table");
  }
  // the carry buffers belong to this call when it runs on a caller's stream
  // (everything above is built on the library stream and kept with the CSF)
  SpdCallScratch scratch(fx ? spd_hs(nullptr) : st);
  if(depth == 0 || cp != nullptr) {
    // two partial rows per chunk of 256 positions, then the fix-up's second
    // level
    uint64_t const nent = 2ull * plan->nitems + fix_level2_entries(2ull * plan->nitems);
    void * cv = nullptr;
    void * cr = nullptr;
    if(fx != nullptr) {
      SPD_TRY(fixed_carve(fx, nent * R * sizeof(T), nent * sizeof(uint32_t), &cv, &cr));
    } else {
      SPD_TRY(spd_scratch(2, nent * R * sizeof(T), &cv));
      SPD_TRY(spd_scratch(3, nent * sizeof(uint32_t), &cr));
    }
    a.carry = (T *) cv;
    a.carry_row = (uint32_t *) cr;
  } else {
    // privatise rows that collect at least two contributions
    uint32_t nhot = 0;
    while(nhot < hp->counts.size() && hp->counts[nhot] >= 2) ++nhot;
    a.hot_slot = hp->rank;
    a.hot_rows = hp->rows;
    a.nhot = nhot;
  }
  if(st != spd_hs(nullptr)) {
    SPD_HIP(hipStreamSynchronize(spd_hs(nullptr)));
  }
  SPD_TRY(spd_zero_async(out, c->dims[mode] * R * sizeof(T), st));
  int const slot = (int) (g_ev_count % SPD_NEV);
  if(g_timing) SPD_HIP(hipEventRecord(g_ev[slot][0], st));
  int rc;
  if(depth == 0) rc = dispatch_n<T, KIND_ROOT>(N, a, st);
  else if(cp != nullptr) rc = dispatch_n<T, KIND_COPY>(N, a, st);
  else if(depth == N - 1) rc = dispatch_n<T, KIND_LEAF>(N, a, st);
  else rc = dispatch_n<T, KIND_INTL>(N, a, st);
  if(g_timing) {
    SPD_HIP(hipEventRecord(g_ev[slot][1], st));
    ++g_ev_count;
  }
  return rc;
}

extern "C" int spd_mttkrp(spd_csf * c, int mode, uint64_t R, void * const * mats,
    void * out, int dtype, int rows_pos, void * stream)
{
  hipStream_t st = spd_hs(stream);
  if(R == 0 || R >= (1ull << 31)) {
    spd_set_error(SPD_EINPUT, "bad rank %llu", (unsigned long long) R);
    return SPD_EINPUT;
  }
  if(dtype == SPD_F32) {
    SPD_TRY(ensure_vals32(c));
    return run_mttkrp<float>(c, mode, R, mats, out, rows_pos, st, dtype, nullptr);
  }
  return run_mttkrp<double>(c, mode, R, mats, out, rows_pos, st, dtype, nullptr);
}

// ---- MTTKRP without allocation ---------------------------------------------
//
// The partial rows of the root and row-copy walks live in one file-scope
// device array per GPU instead of per-call allocations. Host code never
// passes the array's name to a kernel (in host code it is a stand-in, not a
// device address): its address on the current device comes from
// hipGetSymbolAddress, taken once per device on that device's first call.
// Every call on a GPU uses the same array, so calls that overlap on one GPU
// (two host threads, or two streams without an ordering between them) race
// on it. The kernels write every partial row they read in the same launch
// pair (the root walks write both carry rows of every wave or block), so no
// clearing is needed between calls.

#define SPD_NOALLOC_BYTES (16u << 20)
#define SPD_NOALLOC_DEVICES 64

static __device__ __align__(256) unsigned char scratch[SPD_NOALLOC_BYTES];
static void * g_scratch_ptr[SPD_NOALLOC_DEVICES];
static std::mutex g_scratch_mu;

static int noalloc_scratch(void ** out)
{
  int dev = 0;
  SPD_HIP(hipGetDevice(&dev));
  if(dev < 0 || dev >= SPD_NOALLOC_DEVICES) {
    spd_set_error(SPD_EINPUT, "device %d out of range", dev);
    return SPD_EINPUT;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if(g_scratch_ptr[dev] == nullptr) {
    void * p = nullptr;
    hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess || p == nullptr) {
      (void) hipGetLastError();
      spd_set_error(SPD_EHIP, "hipGetSymbolAddress of the MTTKRP scratch failed: %s",
          hipGetErrorString(e));
      return SPD_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SPD_OK;
}

extern "C" size_t spd_mttkrp_noalloc_bytes(void)
{
  return SPD_NOALLOC_BYTES;
}

extern "C" int spd_mttkrp_noalloc(spd_csf * c, int mode, uint64_t R, void * const * mats,
    void * out, int dtype, int rows_pos, void * stream)
{
  hipStream_t st = spd_hs(stream);
  if(R == 0 || R >= (1ull << 31)) {
    spd_set_error(SPD_EINPUT, "bad rank %llu", (unsigned long long) R);
    return SPD_EINPUT;
  }
  if(dtype == SPD_F32 && c->vals32 == nullptr) return not_prepared("FP32 copy of the values");
  void * p = nullptr;
  SPD_TRY(noalloc_scratch(&p));
  MkFixed const fx = {(char *) p, (size_t) SPD_NOALLOC_BYTES};
  if(dtype == SPD_F32) return run_mttkrp<float>(c, mode, R, mats, out, rows_pos, st, dtype, &fx);
  return run_mttkrp<double>(c, mode, R, mats, out, rows_pos, st, dtype, &fx);
}

extern "C" int spd_mttkrp_model(spd_csf const * c, int mode, uint64_t R, int dtype,
    double * bytes, double * flops)
{
  int const N = c->nmodes;
  double const sv = (dtype == SPD_F32) ? 4.0 : 8.0;
  double b = (double) c->nnz * (sv + 4.0);
  double f = 2.0 * (double) R * (double) c->nnz;
  for(int l=1; l <= N - 2; ++l) {
    b += (double) c->u.nfibs[l] * 4.0;
    f += 2.0 * (double) R * (double) c->u.nfibs[l];
  }
  for(int l=0; l <= N - 2; ++l) b += ((double) c->u.nfibs[l] + 1.0) * 4.0;
  for(int m=0; m < N; ++m) b += (double) c->dims[m] * (double) R * sv;
  *bytes = b;
  *flops = f;
  return SPD_OK;
}
