// spd_root3h.hpp -- persistent 3-mode root MTTKRP with the hottest leaf rows
// held in LDS (included by spd_mttkrp.hip after k_root3, whose staging,
// record layout and run combining it shares).
//
// out[s] = sum over fibers f of slice s of A[fid(f)] o sum_{nnz j in f} v_j B[leaf(j)]
// (the reference's CSF factoring, src/mttkrp.c:390-465: the leaf sum of a
// fiber is scaled by the fiber's A row once).
//
// Differences from k_root3:
// * One block per CU (NW waves), launched once: the block first copies the
//   B rows that carry the most nonzeros (ranked once per CSF) into LDS, then
//   every wave walks a contiguous range of 256-nonzero chunks. The leaf ids
//   of hot rows are re-encoded once per CSF (bit 31 set, rank below), so a
//   record's B reference is either an LDS byte offset or a global one.
// * A row gathers: one per fiber (and one at the start of each group's
//   item), not one per nonzero. A group keeps the fiber's leaf sum and the
//   slice's sum apart and folds the first into the second at fiber ends.
// * The slice open at the end of a chunk is carried to the wave's next
//   chunk in registers; only the wave's first run (when its slice began in
//   an earlier wave) and its last run reach the carry buffer: 2 rows per
//   wave instead of 2 per 1024 nonzeros.

#define RH_HOT_BYTES 32768
#define RH_HOTF 0x80000000u        // leaf id / record B offset: hot row
#define RH_FST 0x80000000u         // record slice field: first nonzero of a fiber
#define RH_BATCH 6

// Makes v opaque to the optimiser once it is loaded. The walk selects the
// current fiber's A row (avc = new fiber ? av[j] : avc); without the barrier
// the compiler folds that select into the conditional gathers (each then
// starts from a copy of the previous one, still in flight, and waits for it).
// Placed after every gather of a batch is issued.
template<typename VT>
__device__ __forceinline__ void rh_opaque(VT & v)
{
  asm volatile("" : "+v"(v));
}

template<typename T>
struct Root3hArgs {
  Root3Args<T> b;
  uint32_t const * leafx;          // leaf ids, hot rows as RH_HOTF | rank
  uint32_t const * hot_rows;       // rank -> B row
  uint32_t nhot;                   // ranks encoded in leafx
  uint32_t nchunks;                // chunks of R3_WN nonzeros
  uint32_t cpw;                    // chunks per wave
};

// LDS bytes of one block of k_root3h
template<typename T, int V, int L, int NW>
constexpr int rh_lds_bytes()
{
  constexpr int G = 64 / L;
  constexpr int CW = L * V;
  constexpr int NE = 2 * G;
  constexpr int stage = (R3_WN + 1) * 4 + 2 * R3_WN * 2;
  constexpr int walk = NE * CW * (int) sizeof(T);
  return RH_HOT_BYTES + NW * ((R3_WN + G) * 16 + (sizeof(T) == 8 ? (R3_WN + G) * 8 : 0) +
      (R3_WN + 1) * 4 + NE * 4 + (stage > walk ? stage : walk));
}

template<typename T, int V, int L, int NW, bool HOT>
__global__ void __launch_bounds__(NW * 64, 1)
k_root3h(Root3hArgs<T> h)
{
  Root3Args<T> const & a = h.b;
  typedef typename VecT<T, V>::type VT;
  constexpr int G = 64 / L;
  constexpr int S = R3_WN / G;
  constexpr int CW = L * V;
  constexpr int NE = 2 * G;                   // combine entries per wave and chunk
  constexpr int PER = R3_WN / 64;
  constexpr int HROWS = RH_HOT_BYTES / (CW * (int) sizeof(T));
  constexpr int SZ_AROW = (R3_WN + 1) * 4;
  constexpr int SZ_MARK = R3_WN * 2;
  constexpr int SZ_STAGE = SZ_AROW + 2 * SZ_MARK;
  constexpr int SZ_WALK = NE * CW * (int) sizeof(T);
  constexpr int SZ_U = SZ_STAGE > SZ_WALK ? SZ_STAGE : SZ_WALK;
  static_assert(rh_lds_bytes<T, V, L, NW>() <= 160 * 1024, "k_root3h LDS");
  __shared__ __align__(16) T s_hot[HROWS * CW];
  __shared__ uint4 s_rec[NW][R3_WN + G];
  __shared__ T s_v64[sizeof(T) == 8 ? NW : 1][sizeof(T) == 8 ? R3_WN + G : 1];
  __shared__ uint32_t s_srow[NW][R3_WN + 1];
  __shared__ uint32_t s_crow[NW][NE];
  // per wave: staging marks and A row offsets, then the walk's partial rows
  __shared__ __align__(16) unsigned char s_u[NW][SZ_U];

  uint32_t const R = a.R;
  // HOT = false: every B row from HBM (hot ranks resolved to their rows)
  uint32_t const nh = !HOT ? 0u : (h.nhot < (uint32_t) HROWS ? h.nhot : (uint32_t) HROWS);
  // ---- hot B rows (this block's column tile) into LDS ----
  for(uint32_t e=threadIdx.x; e < nh * (CW / V); e += NW * 64) {
    uint32_t const slot = e / (CW / V);
    uint32_t const cc = e % (CW / V);
    uint32_t const colh = blockIdx.y * CW + cc * V;
    VT v = vzero<VT>();
    if(colh < R) v = *(VT const *) (a.B + (uint64_t) h.hot_rows[slot] * R + colh);
    *(VT *) &s_hot[slot * CW + cc * V] = v;
  }
  __syncthreads();

  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  uint32_t const c = lane % L;
  uint32_t const g = lane / L;
  uint32_t const rowb = R * (uint32_t) sizeof(T);
  uint32_t const col = blockIdx.y * CW + c * V;
  bool const cok = col < R;
  uint32_t const colb = cok ? col * (uint32_t) sizeof(T) : 0u;
  uint32_t const hcolb = c * V * (uint32_t) sizeof(T);
  auto s_arow = reinterpret_cast<uint32_t *>(&s_u[wave][0]);
  auto s_fi = reinterpret_cast<uint16_t *>(&s_u[wave][SZ_AROW]);
  auto s_si = reinterpret_cast<uint16_t *>(&s_u[wave][SZ_AROW + SZ_MARK]);
  auto s_cval = reinterpret_cast<T (*)[CW]>(&s_u[wave][0]);
  char const * const Bb = (char const *) a.B;
  char const * const Ab = (char const *) a.A;
  T const __attribute__((address_space(3))) * const s_hot_l =
      (T const __attribute__((address_space(3))) *) s_hot;
  char const __attribute__((address_space(1))) * const Bg =
      (char const __attribute__((address_space(1))) *) a.B;

  uint64_t const gw = (uint64_t) blockIdx.x * NW + wave;
  uint64_t const ch0 = gw * h.cpw;
  uint64_t const ch1 = (ch0 + h.cpw < h.nchunks) ? ch0 + h.cpw : h.nchunks;
  // run carried from chunk to chunk (held by group 0's lanes)
  uint32_t crow = CARRY_NONE;
  VT cval = vzero<VT>();
  bool chead = false;
  bool whead = (ch0 < ch1) && a.head[ch0 * G] != 0;
  uint32_t hrow = CARRY_NONE;                   // carry slot 0: the wave's head run
  T * const cv = a.carry + 2 * gw * (uint64_t) R;

  for(uint64_t ch=ch0; ch < ch1; ++ch) {
    uint64_t const W0 = ch * R3_WN;
    uint64_t const W1 = (W0 + R3_WN < a.nnz) ? W0 + R3_WN : a.nnz;
    uint32_t const item0 = (uint32_t) (ch * G);
    uint32_t const wn = (uint32_t) (W1 - W0);
    // ---- stage the chunk (as k_root3) ----
    uint32_t const S0 = a.plan[2ull * item0 + 0];
    uint32_t const F0 = a.plan[2ull * item0 + 1];
    {
      uint32_t const nfw = (a.nfibers - F0 < (uint32_t) (R3_WN + 1)) ? a.nfibers - F0 : (uint32_t) (R3_WN + 1);
      uint32_t const nsw = (a.nslices - S0 < (uint32_t) (R3_WN + 1)) ? a.nslices - S0 : (uint32_t) (R3_WN + 1);
      for(uint32_t k=lane; k < R3_WN; k += 64) {
        s_fi[k] = 0;
        s_si[k] = 0;
      }
      for(uint32_t k=lane; k < nsw; k += 64) {
        s_srow[wave][k] = a.rows_are_positions ? S0 + k : a.fids0[S0 + k];
      }
      for(uint32_t k=lane; k < nfw; k += 64) {
        s_arow[k] = a.fids1[F0 + k] * rowb;
      }
      for(uint32_t e=lane; e < (uint32_t) NE; e += 64) s_crow[wave][e] = CARRY_NONE;
      __builtin_amdgcn_wave_barrier();
      for(uint32_t k=1 + lane; k < nfw; k += 64) {
        uint64_t const st = a.fptr1[F0 + k];
        if(st >= W1) break;
        s_fi[st - W0] = 1;
      }
      for(uint32_t q=1 + lane; q < nsw; q += 64) {
        uint32_t const fs = a.fptr0[S0 + q];
        if(fs >= a.nfibers) break;
        uint64_t const st = a.fptr1[fs];
        if(st >= W1) break;
        s_si[st - W0] = 1;
      }
      __builtin_amdgcn_wave_barrier();
      uint32_t fm[PER], sm[PER];
      uint32_t fsum = 0, ssum = 0;
      #pragma unroll
      for(int q=0; q < PER; ++q) {
        fm[q] = s_fi[lane * PER + q];
        sm[q] = s_si[lane * PER + q];
        fsum += fm[q];
        ssum += sm[q];
      }
      uint32_t fpre = wave_incl_scan(fsum) - fsum;
      uint32_t spre = wave_incl_scan(ssum) - ssum;
      #pragma unroll
      for(int q=0; q < PER; ++q) {
        fpre += fm[q];
        spre += sm[q];
        uint32_t const p = lane * PER + q;
        uint32_t lx = 0;
        T v = (T) 0;
        if(p < wn) {
          lx = h.leafx[W0 + p];
          v = a.vals[W0 + p];
        }
        uint4 r;
        if(lx & RH_HOTF) {
          uint32_t const rank = lx & ~RH_HOTF;
          r.x = (rank < nh) ? (RH_HOTF | (rank * (uint32_t) (CW * sizeof(T))))
                            : h.hot_rows[rank] * rowb;
        } else {
          r.x = lx * rowb;
        }
        r.y = s_arow[fpre];
        r.z = spre | (fm[q] ? RH_FST : 0u);
        if constexpr (sizeof(T) == 4) {
          r.w = __float_as_uint(v);
        } else {
          r.w = 0;
          s_v64[wave][p + p / S] = v;
        }
        s_rec[wave][p + p / S] = r;
      }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- walk this group's item ----
    uint32_t const item = item0 + g;
    uint64_t const n0 = (uint64_t) item * S;
    uint32_t const e0 = g * 2;
    if(n0 < a.nnz) {
      uint64_t const n1 = (n0 + S < a.nnz) ? n0 + S : a.nnz;
      bool head_partial = a.head[item] != 0;
      uint32_t const k0 = (uint32_t) (n0 - W0);
      uint32_t const k1 = (uint32_t) (n1 - W0);
      uint32_t cur_s = s_rec[wave][k0 + k0 / S].z & ~RH_FST;
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
        uint32_t const orow = s_srow[wave][cur_s];
        if(head_partial) {
          s_crow[wave][e0] = orow;
          if(cok) *(VT *) &s_cval[e0][c * V] = accs;
        } else {
          vstore<T, V>(a.out, orow, R, col, cok, accs);
        }
        head_partial = false;
      };
      // a new fiber starts at this nonzero (or the item starts): the last
      // fiber's leaf sum joins its slice; a new slice flushes the old one
      auto step = [&](uint4 const & r, T v, VT b, VT av, bool nf) {
        if(nf) {
          accs += accf * avc;
          accf = vzero<VT>();
          avc = av;
        }
        uint32_t const sl = r.z & ~RH_FST;
        if(sl != cur_s) {
          flush();
          accs = vzero<VT>();
          cur_s = sl;
        }
        accf += v * b;
      };
      // hot rows by an LDS read, the others by a global load: two
      // instructions under complementary lane masks (a generic pointer
      // would make one flat load, whose LDS lanes still pass through the
      // vector memory address path)
      auto loadb = [&](uint32_t x) -> VT {
        VT v;
        if constexpr (HOT) {
          if(x & RH_HOTF) {
            v = *(VT const __attribute__((address_space(3))) *)
                (s_hot_l + (((x & ~RH_HOTF) + hcolb) / (uint32_t) sizeof(T)));
          } else {
            v = *(VT const __attribute__((address_space(1))) *) (Bg + (x + colb));
          }
        } else {
          v = *(VT const __attribute__((address_space(1))) *) (Bg + (x + colb));
        }
        return v;
      };
      for(uint32_t k=k0; k < k1; k += RH_BATCH) {
        if(k + RH_BATCH <= k1) {
          uint4 r[RH_BATCH];
          VT b[RH_BATCH];
          VT av[RH_BATCH];
          bool nf[RH_BATCH];
          #pragma unroll
          for(int j=0; j < RH_BATCH; ++j) r[j] = s_rec[wave][k + j + (k + j) / S];
          #pragma unroll
          for(int j=0; j < RH_BATCH; ++j) b[j] = loadb(r[j].x);
          // A rows at fiber starts only (av[j] is read only where loaded)
          #pragma unroll
          for(int j=0; j < RH_BATCH; ++j) {
            nf[j] = (r[j].z & RH_FST) != 0u || (k + j == k0);
            av[j] = nf[j] ? *(VT const *) (Ab + (r[j].y + colb)) : vzero<VT>();
          }
          #pragma unroll
          for(int j=0; j < RH_BATCH; ++j) rh_opaque(av[j]);
          #pragma unroll
          for(int j=0; j < RH_BATCH; ++j) step(r[j], rec_val(r[j], k + j), b[j], av[j], nf[j]);
        } else {
          for(uint32_t kk=k; kk < k1; ++kk) {
            uint4 const r = s_rec[wave][kk + kk / S];
            bool const nf = (r.z & RH_FST) != 0u || kk == k0;
            VT const b = loadb(r.x);
            VT const av = nf ? *(VT const *) (Ab + (r.y + colb)) : vzero<VT>();
            step(r, rec_val(r, kk), b, av, nf);
          }
        }
      }
      accs += accf * avc;
      // is the last slice complete at n1? It is if n1 starts a new slice
      // (or ends the tensor)
      bool complete = (n1 == a.nnz);
      if(!complete && n1 < W1) {
        complete = (s_rec[wave][k1 + k1 / S].z & ~RH_FST) != cur_s;
      } else if(!complete) {
        uint32_t const fs = a.fptr0[S0 + cur_s + 1];
        complete = (fs < a.nfibers) && ((uint64_t) a.fptr1[fs] == n1);
      }
      uint32_t const orow = s_srow[wave][cur_s];
      if(complete && !head_partial) {
        vstore<T, V>(a.out, orow, R, col, cok, accs);
      } else if(complete) {
        s_crow[wave][e0] = orow;
        if(cok) *(VT *) &s_cval[e0][c * V] = accs;
      } else {
        s_crow[wave][e0 + 1] = orow;
        if(cok) *(VT *) &s_cval[e0 + 1][c * V] = accs;
      }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- combine the chunk's partial slices with the carried run ----
    // Entries are in nonzero order, so a change of row closes the previous
    // run; the last run stays open and is carried into the next chunk.
    if(g == 0) {
      for(uint32_t e=0; e < (uint32_t) NE; ++e) {
        uint32_t const r = s_crow[wave][e];
        if(r == CARRY_NONE) continue;
        if(r != crow) {
          if(crow != CARRY_NONE) {
            if(chead) {
              vstore<T, V>(cv, 0u, R, col, cok, cval);
              hrow = crow;
            } else {
              vstore<T, V>(a.out, crow, R, col, cok, cval);
            }
          }
          crow = r;
          cval = vzero<VT>();
          chead = whead;
          whead = false;
        }
        if(cok) cval += *(VT const *) &s_cval[e][c * V];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // the wave's open runs go to the carry buffer (the fix-up adds them)
  if(g == 0) {
    if(crow != CARRY_NONE) vstore<T, V>(cv, 1u, R, col, cok, cval);
    if(c == 0 && blockIdx.y == 0) {
      a.carry_row[2 * gw] = hrow;
      a.carry_row[2 * gw + 1] = crow;
    }
  }
}

__global__ void k_encode_hot(uint32_t const * __restrict__ ids, uint64_t n,
    int32_t const * __restrict__ rank, uint32_t nhot, uint32_t * __restrict__ out)
{
  uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const stride = (uint64_t) gridDim.x * blockDim.x;
  for(; i < n; i += stride) {
    uint32_t const id = ids[i];
    int32_t const rk = rank[id];
    out[i] = (rk >= 0 && (uint32_t) rk < nhot) ? (RH_HOTF | (uint32_t) rk) : id;
  }
}
