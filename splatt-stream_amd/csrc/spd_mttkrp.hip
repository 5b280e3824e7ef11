// spd_mttkrp.hip -- MTTKRP over a device CSF (root, internal and leaf output
// levels), the GPU counterpart of the reference's mttkrp_csf
// (src/mttkrp.c:1287-1343 and the kernels at :390-1278).
//
// Work decomposition. The nonzeros (leaf level, in CSF order) are cut into
// items of S consecutive nonzeros. A group of L lanes owns one item; each
// lane holds V consecutive columns of a factor row, so one group covers
// L*V columns (a second grid dimension tiles wider ranks). 64/L groups
// share a wavefront. Items are cut without regard to fiber or slice
// boundaries, so a giant Zipf slice is spread over many groups and every
// group does the same amount of leaf work.
//
// A group walks its item like the reference's DFS, but iteratively: it
// keeps the current node and that node's child end at every level; leaf
// nonzeros accumulate v * U_{N-1}[j] into acc[N-2]; when a node completes,
// its accumulator is multiplied by the node's factor row and added to the
// parent's accumulator (root output), or flushed to the output row
// (internal/leaf output). Partial nodes at item edges are folded the same
// way (the products are linear) and flushed with atomics; a root slice that
// starts and ends inside one item is written with a plain store.
#include "spd_internal.hpp"
#include "spd_csf_dev.hpp"
#include <stdlib.h>
#include <vector>

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

template<typename T>
struct WalkArgs {
  uint32_t const * fptr[SPD_MAX_MODES];
  uint32_t const * fids[SPD_MAX_MODES];
  T const * vals;
  T const * um[SPD_MAX_MODES];    // factor of the mode stored at each level
  T * out;
  uint32_t const * plan;          // per item: node at levels 0..N-2
  uint8_t const * head;           // per item: 1 if its first slice began earlier
  uint64_t nnz;
  uint32_t R;
  uint32_t S;
  uint32_t nitems;
  int depth;                      // output level (internal kernels)
  int rows_are_positions;         // root row = slice position (untiled host CSF)
};

template<typename T, int N, int KIND, int V, int L>
__global__ void __launch_bounds__(MK_THREADS)
k_walk(WalkArgs<T> a)
{
  typedef typename VecT<T, V>::type VT;
  constexpr int G = 64 / L;
  constexpr int NL = N - 1;        // internal levels 0..N-2
  int const lane = threadIdx.x & 63;
  int const c = lane % L;
  int const g = lane / L;
  uint32_t const item = ((uint32_t) blockIdx.x * (MK_THREADS / 64) + (threadIdx.x >> 6)) * G + g;
  if(item >= a.nitems) return;
  uint32_t const col = blockIdx.y * (L * V) + c * V;
  bool const cok = col < a.R;
  uint32_t const R = a.R;

  uint64_t const n0 = (uint64_t) item * a.S;
  uint64_t n1 = n0 + a.S;
  if(n1 > a.nnz) n1 = a.nnz;

  uint32_t node[NL];
  uint32_t cend[NL];
  #pragma unroll
  for(int l=0; l < NL; ++l) {
    node[l] = a.plan[(uint64_t) item * NL + l];
  }
  #pragma unroll
  for(int l=0; l < NL; ++l) {
    cend[l] = a.fptr[l][node[l] + 1];
  }
  bool head_partial = a.head[item] != 0;
  int const depth = (KIND == KIND_ROOT) ? 0 : ((KIND == KIND_LEAF) ? N - 1 : a.depth);

  VT acc[NL];
  #pragma unroll
  for(int l=0; l < NL; ++l) acc[l] = vzero<VT>();

  // pre[l] = product of factor rows of the current nodes at levels 0..l,
  // kept for levels below the output level (internal/leaf outputs).
  VT pre[NL];
  if constexpr (KIND != KIND_ROOT) {
    #pragma unroll
    for(int l=0; l < NL; ++l) {
      if(l < depth) {
        uint32_t const row = (l == 0) ? (a.rows_are_positions ? node[0] : a.fids[0][node[0]])
                                      : a.fids[l][node[l]];
        VT const r = vload<T, V>(a.um[l], row, R, col, cok);
        pre[l] = (l == 0) ? r : pre[l > 0 ? l - 1 : 0] * r;
      }
    }
  }

  uint64_t n = n0;
  int stop = NL;   // levels >= stop hold nothing pending at the end
  while(true) {
    uint64_t const e = (cend[NL - 1] < n1) ? cend[NL - 1] : n1;
    if constexpr (KIND == KIND_LEAF) {
      VT const h = pre[NL - 1];
      for(; n < e; ++n) {
        T const v = a.vals[n];
        uint32_t const j = a.fids[N - 1][n];
        vadd_atomic<T, V>(a.out, j, R, col, cok, v * h);
      }
    } else {
      T const * __restrict__ ul = a.um[N - 1];
      uint32_t const * __restrict__ lf = a.fids[N - 1];
      T const * __restrict__ vv = a.vals;
      VT s = acc[NL - 1];
      for(; n + 4 <= e; n += 4) {
        T const v0 = vv[n], v1 = vv[n + 1], v2 = vv[n + 2], v3 = vv[n + 3];
        uint32_t const j0 = lf[n], j1 = lf[n + 1], j2 = lf[n + 2], j3 = lf[n + 3];
        VT const b0 = vload<T, V>(ul, j0, R, col, cok);
        VT const b1 = vload<T, V>(ul, j1, R, col, cok);
        VT const b2 = vload<T, V>(ul, j2, R, col, cok);
        VT const b3 = vload<T, V>(ul, j3, R, col, cok);
        s += v0 * b0;
        s += v1 * b1;
        s += v2 * b2;
        s += v3 * b3;
      }
      for(; n < e; ++n) {
        T const v = vv[n];
        uint32_t const j = lf[n];
        s += v * vload<T, V>(ul, j, R, col, cok);
      }
      acc[NL - 1] = s;
    }
    if(n < cend[NL - 1]) {
      // the item ends inside this leaf fiber
      stop = NL;
      break;
    }
    // the leaf fiber node[NL-1] is complete: walk up through completed nodes
    bool go = true;
    int top = NL - 1;
    #pragma unroll
    for(int l=NL - 1; l >= 0; --l) {
      if(go) {
        top = l;
        if constexpr (KIND == KIND_ROOT) {
          if(l == 0) {
            uint32_t const row = a.rows_are_positions ? node[0] : a.fids[0][node[0]];
            if(head_partial) {
              vadd_atomic<T, V>(a.out, row, R, col, cok, acc[0]);
            } else {
              vstore<T, V>(a.out, row, R, col, cok, acc[0]);
            }
            head_partial = false;
            acc[0] = vzero<VT>();
            node[0] += 1;
            go = false;
          } else {
            VT const r = vload<T, V>(a.um[l], a.fids[l][node[l]], R, col, cok);
            acc[l - 1] += acc[l] * r;
            acc[l] = vzero<VT>();
            node[l] += 1;
            go = (node[l] == cend[l - 1]);
          }
        } else if constexpr (KIND == KIND_INTL) {
          if(l > depth) {
            VT const r = vload<T, V>(a.um[l], a.fids[l][node[l]], R, col, cok);
            acc[l - 1] += acc[l] * r;
            acc[l] = vzero<VT>();
          } else if(l == depth) {
            vadd_atomic<T, V>(a.out, a.fids[l][node[l]], R, col, cok,
                acc[l] * pre[l > 0 ? l - 1 : 0]);
            acc[l] = vzero<VT>();
          }
          node[l] += 1;
          go = (l > 0) && (node[l] == cend[l > 0 ? l - 1 : 0]);
        } else {
          node[l] += 1;
          go = (l > 0) && (node[l] == cend[l > 0 ? l - 1 : 0]);
        }
      }
    }
    if(n >= n1) {
      stop = top;
      break;
    }
    // refresh child ends (and prefix rows) of the levels whose node moved
    #pragma unroll
    for(int l=0; l < NL; ++l) {
      if(l >= top) {
        cend[l] = a.fptr[l][node[l] + 1];
      }
    }
    if constexpr (KIND != KIND_ROOT) {
      #pragma unroll
      for(int l=0; l < NL; ++l) {
        if(l >= top && l < depth) {
          uint32_t const row = (l == 0) ? (a.rows_are_positions ? node[0] : a.fids[0][node[0]])
                                        : a.fids[l][node[l]];
          VT const r = vload<T, V>(a.um[l], row, R, col, cok);
          pre[l] = (l == 0) ? r : pre[l > 0 ? l - 1 : 0] * r;
        }
      }
    }
  }

  // flush what is pending at the item end: levels < stop hold partial sums
  if constexpr (KIND == KIND_ROOT) {
    if(stop >= 1) {
      #pragma unroll
      for(int l=NL - 1; l >= 1; --l) {
        if(l < stop) {
          VT const r = vload<T, V>(a.um[l], a.fids[l][node[l]], R, col, cok);
          acc[l - 1] += acc[l] * r;
        }
      }
      uint32_t const row = a.rows_are_positions ? node[0] : a.fids[0][node[0]];
      vadd_atomic<T, V>(a.out, row, R, col, cok, acc[0]);
    }
  } else if constexpr (KIND == KIND_INTL) {
    if(stop > depth) {
      #pragma unroll
      for(int l=NL - 1; l >= 1; --l) {
        if(l < stop && l > depth) {
          VT const r = vload<T, V>(a.um[l], a.fids[l][node[l]], R, col, cok);
          acc[l - 1] += acc[l] * r;
        }
      }
      #pragma unroll
      for(int l=0; l < NL; ++l) {
        if(l == depth) {
          vadd_atomic<T, V>(a.out, a.fids[l][node[l]], R, col, cok,
              acc[l] * pre[l > 0 ? l - 1 : 0]);
        }
      }
    }
  }
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

void spd_csf_drop_plans(spd_csf * c)
{
  MttkrpPlan * p = c->plans;
  while(p) {
    MttkrpPlan * nx = p->next;
    spd_free(p->nodes);
    spd_free(p->head);
    free(p);
    p = nx;
  }
  c->plans = nullptr;
}

static int get_plan(spd_csf * c, uint32_t S, MttkrpPlan ** out, hipStream_t st)
{
  for(MttkrpPlan * p = c->plans; p; p = p->next) {
    if(p->S == S) { *out = p; return SPD_OK; }
  }
  MttkrpPlan * p = (MttkrpPlan *) calloc(1, sizeof(MttkrpPlan));
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
static int launch_walk(WalkArgs<T> const & a, hipStream_t st)
{
  constexpr int G = 64 / L;
  uint32_t const groups_per_block = (MK_THREADS / 64) * G;
  dim3 grid((unsigned) spd_cdiv(a.nitems, groups_per_block),
            (unsigned) spd_cdiv(a.R, (uint64_t) (L * V)));
  hipLaunchKernelGGL((k_walk<T, N, KIND, V, L>), grid, dim3(MK_THREADS), 0, st, a);
  SPD_LAUNCH_CHECK();
  return SPD_OK;
}

template<typename T, int N, int KIND>
static int dispatch_vl(WalkArgs<T> const & a, hipStream_t st)
{
  uint32_t const R = a.R;
  if constexpr (sizeof(T) == 4) {
    if(R % 4 == 0 && R >= 8) {
      uint32_t const lanes = R / 4;
      if(lanes <= 2) return launch_walk<T, N, KIND, 4, 2>(a, st);
      if(lanes <= 4) return launch_walk<T, N, KIND, 4, 4>(a, st);
      if(lanes <= 8) return launch_walk<T, N, KIND, 4, 8>(a, st);
      return launch_walk<T, N, KIND, 4, 16>(a, st);
    }
  } else {
    if(R % 2 == 0 && R >= 8) {
      uint32_t const lanes = R / 2;
      if(lanes <= 4) return launch_walk<T, N, KIND, 2, 4>(a, st);
      if(lanes <= 8) return launch_walk<T, N, KIND, 2, 8>(a, st);
      if(lanes <= 16) return launch_walk<T, N, KIND, 2, 16>(a, st);
      return launch_walk<T, N, KIND, 2, 32>(a, st);
    }
  }
  if(R <= 16) return launch_walk<T, N, KIND, 1, 16>(a, st);
  return launch_walk<T, N, KIND, 1, 64>(a, st);
}

template<typename T, int KIND>
static int dispatch_n(int N, WalkArgs<T> const & a, hipStream_t st)
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

// Groups to give each CU several waves even at modest nnz, but never
// fewer than 64 nonzeros per item.
static uint32_t choose_S(uint64_t nnz, uint32_t R, int dtype)
{
  uint32_t lanes_per_row = (dtype == SPD_F32) ? (R % 4 == 0 && R >= 8 ? R / 4 : 16)
                                              : (R % 2 == 0 && R >= 8 ? R / 2 : 16);
  if(lanes_per_row > 64) lanes_per_row = 64;
  uint32_t L = 1;
  while(L < lanes_per_row) L <<= 1;
  uint32_t const G = 64 / L;
  uint64_t const target_groups = 256ull * 16 * G;   // 16 waves per CU
  uint64_t S = nnz / target_groups;
  if(S < 64) S = 64;
  if(S > 2048) S = 2048;
  S = (S + 3) & ~3ull;
  return (uint32_t) S;
}

static int g_timing = 0;
static hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
static int g_ev_valid = 0;

extern "C" int spd_mttkrp_timing(int enable)
{
  g_timing = enable;
  if(enable && g_ev0 == nullptr) {
    SPD_HIP(hipEventCreate(&g_ev0));
    SPD_HIP(hipEventCreate(&g_ev1));
  }
  return SPD_OK;
}

extern "C" double spd_mttkrp_last_kernel_ms(void)
{
  if(!g_ev_valid) return -1.0;
  float ms = -1.f;
  if(hipEventSynchronize(g_ev1) != hipSuccess) return -1.0;
  if(hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) return -1.0;
  return (double) ms;
}

static int ensure_vals32(spd_csf * c)
{
  if(c->vals32 != nullptr) return SPD_OK;
  SPD_TRY(spd_malloc((void **) &c->vals32, c->nnz * sizeof(float)));
  SPD_TRY(spd_f64_to_f32(c->u.vals, c->vals32, c->nnz, nullptr));
  return SPD_OK;
}

int spd_zero_async(void * ptr, size_t bytes, hipStream_t st);

template<typename T>
static int run_mttkrp(spd_csf * c, int mode, uint64_t R, void * const * mats,
    void * out, int rows_pos, hipStream_t st, int dtype)
{
  int const N = c->nmodes;
  int depth = -1;
  for(int l=0; l < N; ++l) if(c->perm[l] == mode) depth = l;
  if(depth < 0) {
    spd_set_error(SPD_EINPUT, "mode %d not in tensor", mode);
    return SPD_EINPUT;
  }
  uint32_t const S = choose_S(c->nnz, (uint32_t) R, dtype);
  MttkrpPlan * plan = nullptr;
  // plans are built on the library stream; make the caller's stream wait
  SPD_TRY(get_plan(c, S, &plan, spd_hs(nullptr)));
  if(st != spd_hs(nullptr)) {
    SPD_HIP(hipStreamSynchronize(spd_hs(nullptr)));
  }
  WalkArgs<T> a = {};
  for(int l=0; l < N - 1; ++l) a.fptr[l] = c->u.fptr[l];
  for(int l=0; l < N; ++l) {
    a.fids[l] = c->u.fids[l];
    a.um[l] = (T const *) mats[c->perm[l]];
  }
  a.vals = (sizeof(T) == 4) ? (T const *) c->vals32 : (T const *) c->u.vals;
  a.out = (T *) out;
  a.plan = plan->nodes;
  a.head = plan->head;
  a.nnz = c->nnz;
  a.R = (uint32_t) R;
  a.S = S;
  a.nitems = plan->nitems;
  a.depth = depth;
  a.rows_are_positions = rows_pos;
  SPD_TRY(spd_zero_async(out, c->dims[mode] * R * sizeof(T), st));
  if(g_timing) SPD_HIP(hipEventRecord(g_ev0, st));
  int rc;
  if(depth == 0) rc = dispatch_n<T, KIND_ROOT>(N, a, st);
  else if(depth == N - 1) rc = dispatch_n<T, KIND_LEAF>(N, a, st);
  else rc = dispatch_n<T, KIND_INTL>(N, a, st);
  if(g_timing) {
    SPD_HIP(hipEventRecord(g_ev1, st));
    g_ev_valid = 1;
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
    return run_mttkrp<float>(c, mode, R, mats, out, rows_pos, st, dtype);
  }
  return run_mttkrp<double>(c, mode, R, mats, out, rows_pos, st, dtype);
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
