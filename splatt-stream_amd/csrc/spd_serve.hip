// spd_serve.hip -- many independent CPD streams served by one GPU.
//
// No counterpart in the reference (it factors one stream per process). A
// server holds streams; each stream owns its device state (StreamState: the
// same per-slice step as the whole-stream driver), its own HIP stream and
// scratch, and a FIFO of submitted slices. Streams are admitted while their
// device footprint fits the HBM budget; a stream that does not fit waits
// until closed streams release memory. A pool of worker threads (as many as
// slices may be in flight at once) takes ready streams in arrival order and
// runs one slice each, so slices of different streams overlap on the device
// while the slices of one stream stay in order. A slice's host part (the
// parser's ordering and first-appearance relabelling, ParserSimple.cc:13-80,
// applied slice by slice) runs in the worker before its device part.
#include "spd_internal.hpp"
#include "spd_stream_state.hpp"
#include "spd_batch.hpp"
#include "spd_batch_host.hpp"
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace {

struct Slice {
  int64_t ticket = 0;
  int64_t stream = 0;
  uint64_t nnz = 0;
  std::vector<uint64_t> ind[SPD_MAX_MODES];   // original labels
  std::vector<double> vals;
  double t_submit = 0.;
  double t_done = 0.;
  double gpu_s = 0.;
  double local_err = 0.;
  uint64_t outer = 0;
  int status = 1;                             // 1 pending, 0 done, < 0 error
  char err[160] = "";
};

struct Stream {
  int64_t id = 0;
  int N = 0;
  spd_stream_cfg cfg = {};
  uint64_t cap[SPD_MAX_MODES] = {};
  size_t footprint = 0;
  bool admitted = false;
  bool busy = false;
  bool closing = false;
  bool allocated = false;
  std::deque<Slice *> queue;
  // parser state: original index -> label per mode
  std::vector<uint32_t> lab[SPD_MAX_MODES];
  uint64_t next_label[SPD_MAX_MODES] = {};
  // the stream's own rand() sequence (glibc TYPE_3 generator, as srand/rand)
  struct random_data rd = {};
  char rstate[128] = {};
  // device
  hipStream_t hs = nullptr;
  SpdScratchCtx scr = {};
  StreamState * state = nullptr;
  uint32_t * d_ind[SPD_MAX_MODES] = {};
  double * d_vals = nullptr;
  uint64_t slice_cap = 0;
  std::vector<double> time_rows;
  // factors kept on the host once the stream is closed and its memory freed
  std::vector<double> kept[SPD_MAX_MODES];
  uint64_t kept_rows[SPD_MAX_MODES] = {};
  bool has_kept = false;
  // batched mode: slices prepared on the host ahead of the device step
  bool batchable = false;
  bool prepping = false;
  std::deque<struct PSlice *> prepared;
  uint64_t prep_rows[SPD_MAX_MODES] = {};   // active rows after the prepared slices
  uint64_t prep_count = 0;
  uint8_t * d_slice = nullptr;              // the slice's device arrays
  size_t d_slice_cap = 0;
  double * d_part = nullptr;                // partial sums of the batched kernels
};

// A slice prepared for the batched step: ordered, relabelled, its copies
// ordered by each other mode, the new rows' values, all in one pinned buffer
// laid out as it is copied to the device.
struct PSlice {
  Slice * sl = nullptr;
  uint64_t nnz = 0;
  uint64_t rows[SPD_MAX_MODES] = {};
  uint64_t grow0[SPD_MAX_MODES] = {};
  int first = 0;
  int small = 0;
  double xnorm = 0.;
  uint8_t * buf = nullptr;                  // pinned
  size_t cap = 0;
  size_t bytes = 0;
  size_t off_ind[SPD_MAX_MODES][SPD_MAX_MODES] = {};
  size_t off_vals[SPD_MAX_MODES] = {};
  size_t off_grow[SPD_MAX_MODES] = {};
};

// n factor values from a glibc generator state (src/util.c:15-40: the
// magnitude from one draw, the sign from the next). The draws are those of
// random_r: for the additive feedback generator (TYPE_3, a 128-byte state)
// the step is inlined here, with the two state pointers kept in registers
// (random_r's own call costs several times more); other types call it.
void fill_rand_state(struct random_data * rd, double * vals, uint64_t n)
{
  if(rd->rand_type != 3) {
    for(uint64_t i=0; i < n; ++i) {
      int32_t a = 0, b = 0;
      random_r(rd, &a);
      random_r(rd, &b);
      double v = 3.0 * ((double) a / (double) RAND_MAX);
      if(b % 2 == 0) v = -v;
      vals[i] = v;
    }
    return;
  }
  int32_t * fptr = rd->fptr;
  int32_t * rptr = rd->rptr;
  int32_t * const state = rd->state;
  int32_t * const end = rd->end_ptr;
  auto draw = [&]() -> int32_t {
    uint32_t const val = (uint32_t) *fptr + (uint32_t) *rptr;
    *fptr = (int32_t) val;
    ++fptr;
    if(fptr >= end) {
      fptr = state;
      ++rptr;
    } else {
      ++rptr;
      if(rptr >= end) rptr = state;
    }
    return (int32_t) (val >> 1);
  };
  for(uint64_t i=0; i < n; ++i) {
    int32_t const a = draw();
    int32_t const b = draw();
    double v = 3.0 * ((double) a / (double) RAND_MAX);
    if(b % 2 == 0) v = -v;
    vals[i] = v;
  }
  rd->fptr = fptr;
  rd->rptr = rptr;
}

void fill_stream_rand(void * ctx, double * vals, uint64_t n)
{
  fill_rand_state(&((Stream *) ctx)->rd, vals, n);
}

}  // namespace

// A worker's HIP stream and pinned staging for slice uploads (reused by
// every slice the worker runs, so copies are asynchronous and no stream is
// created per slice or per tensor stream).
struct WorkerRes {
  hipStream_t hs = nullptr;
  uint32_t * hind[SPD_MAX_MODES] = {};
  double * hv = nullptr;
  uint64_t cap = 0;
};

struct spd_server {
  size_t budget = 0;
  size_t used = 0;
  int max_inflight = 1;
  std::vector<WorkerRes> wres;
  std::mutex mu;
  std::condition_variable cv_work;
  std::condition_variable cv_done;
  std::vector<std::thread> workers;
  std::map<int64_t, Stream *> streams;
  std::deque<int64_t> waiting;      // streams waiting for memory
  std::deque<int64_t> ready;        // admitted, not busy, with queued slices
  std::vector<Slice *> slices;      // by ticket
  int64_t next_stream = 0;
  uint64_t pending = 0;
  bool stop = false;
  // batched mode (max_inflight > 1): host preparation threads and engines
  bool batched = false;
  std::vector<std::thread> preppers;
  std::vector<std::thread> engines;
  std::vector<BtEngine *> eng;
  std::vector<PSlice *> pool;       // prepared-slice buffers to reuse
  std::condition_variable cv_prep;
};

namespace {

void admit_waiting(spd_server * sv)
{
  // caller holds sv->mu
  while(!sv->waiting.empty()) {
    Stream * s = sv->streams[sv->waiting.front()];
    if(sv->used + s->footprint > sv->budget) break;
    sv->waiting.pop_front();
    s->admitted = true;
    sv->used += s->footprint;
    if(sv->batched) {
      sv->cv_work.notify_all();
    } else if(!s->queue.empty() && !s->busy) {
      sv->ready.push_back(s->id);
    }
  }
}

// Device part of releasing an idle stream: its factors to the host, its
// device memory back to the pool (stream-ordered). Runs without the server
// lock: nothing else touches an idle, closing stream (it is marked busy).
void release_device(Stream * s)
{
  if(!s->allocated) return;
  if(s->state != nullptr) {
    uint32_t const R = s->state->R;
    for(int m=0; m < s->N; ++m) {
      if(m == s->cfg.stream_mode) continue;
      s->kept_rows[m] = s->state->rows[m];
      s->kept[m].resize(s->kept_rows[m] * R);
      if(hipMemcpyAsync(s->kept[m].data(), s->state->Unew[m],
            s->kept_rows[m] * R * sizeof(double), hipMemcpyDeviceToHost, s->hs) != hipSuccess) {
        s->kept_rows[m] = 0;
      }
    }
    (void) hipStreamSynchronize(s->hs);
    s->has_kept = true;
  }
  // stream-ordered frees on the worker stream the tensor stream last ran
  // on (its work is complete)
  spd_scratch_release(&s->scr);
  delete s->state;
  s->state = nullptr;
  for(int m=0; m < s->N; ++m) {
    (void) spd_free_on(s->d_ind[m], s->hs);
    s->d_ind[m] = nullptr;
  }
  (void) spd_free_on(s->d_vals, s->hs);
  s->d_vals = nullptr;
  s->slice_cap = 0;
  (void) spd_free_on(s->d_slice, s->hs);
  s->d_slice = nullptr;
  s->d_slice_cap = 0;
  (void) spd_free_on(s->d_part, s->hs);
  s->d_part = nullptr;
  s->allocated = false;
}

// Budget part (caller holds sv->mu): the footprint returns, waiting
// streams are admitted.
void release_budget(spd_server * sv, Stream * s)
{
  if(s->admitted) {
    sv->used -= s->footprint;
    s->admitted = false;
  }
  admit_waiting(sv);
}

// Stable order of a slice by the modes after the streamed one (cyclically),
// the order the reference's parser gives (ParserSimple.cc:13-80). When the
// capacities' bits fit 64, an LSD radix sort of packed keys (11-bit digits,
// stable); otherwise a comparison sort.
static void slice_order(Slice const * sl, int N, int t, uint64_t const * cap,
    std::vector<uint32_t> & order)
{
  uint64_t const n = sl->nnz;
  int keys[SPD_MAX_MODES];
  int kb[SPD_MAX_MODES];
  int nk = 0;
  int tot = 0;
  for(int k=1; k < N; ++k) {
    int const m = (t + k) % N;
    int b = 1;
    while(b < 63 && (1ull << b) < cap[m]) ++b;
    keys[nk] = m;
    kb[nk++] = b;
    tot += b;
  }
  order.resize(n);
  if(tot > 64) {
    for(uint64_t x=0; x < n; ++x) order[x] = (uint32_t) x;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      for(int q=0; q < nk; ++q) {
        uint64_t const ia = sl->ind[keys[q]][a];
        uint64_t const ib = sl->ind[keys[q]][b];
        if(ia != ib) return ia < ib;
      }
      return false;
    });
    return;
  }
  // packed key: the first sort mode most significant; indices beyond a
  // capacity are caught by the caller's range check
  std::vector<uint64_t> key(n), key2(n);
  std::vector<uint32_t> ord2(n);
  for(uint64_t x=0; x < n; ++x) {
    uint64_t kv = 0;
    for(int q=0; q < nk; ++q) {
      uint64_t const v = sl->ind[keys[q]][x] & ((kb[q] >= 64) ? ~0ull : ((1ull << kb[q]) - 1));
      kv = (kb[q] >= 64) ? v : ((kv << kb[q]) | v);
    }
    key[x] = kv;
    order[x] = (uint32_t) x;
  }
  constexpr int DB = 11;
  constexpr uint32_t NB = 1u << DB;
  std::vector<uint64_t> cnt(NB);
  for(int sh=0; sh < tot; sh += DB) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for(uint64_t x=0; x < n; ++x) ++cnt[(key[x] >> sh) & (NB - 1)];
    uint64_t run = 0;
    for(uint32_t d=0; d < NB; ++d) {
      uint64_t const c = cnt[d];
      cnt[d] = run;
      run += c;
    }
    for(uint64_t x=0; x < n; ++x) {
      uint64_t const pos = cnt[(key[x] >> sh) & (NB - 1)]++;
      key2[pos] = key[x];
      ord2[pos] = order[x];
    }
    key.swap(key2);
    order.swap(ord2);
  }
}

// Host part of a slice: order by the modes after the streamed one
// (cyclically), relabel by first appearance, upload. Device part: the step.
static int run_slice_body(Stream * s, Slice * sl, WorkerRes & w)
{
  int const N = s->N;
  int const t = s->cfg.stream_mode;
  uint64_t const n = sl->nnz;
  // this worker's stream runs the slice; the tensor stream's earlier work
  // (on any worker stream) is complete, and so are the worker's earlier
  // copies out of its staging buffers
  SPD_HIP(hipStreamSynchronize(w.hs));
  s->hs = w.hs;
  s->scr.st = w.hs;
  spd_scratch_bind(&s->scr);
  if(!s->allocated) {
    s->allocated = true;
    s->state = new StreamState();
    SPD_TRY(s->state->init(N, s->cfg, s->cap, w.hs, 1));
  } else {
    s->state->set_stream(w.hs);
  }
  if(n > w.cap) {
    uint64_t nc = w.cap ? w.cap : 4096;
    while(nc < n) nc *= 2;
    for(int m=0; m < SPD_MAX_MODES; ++m) {
      if(w.hind[m]) (void) hipHostFree(w.hind[m]);
      w.hind[m] = nullptr;
    }
    if(w.hv) (void) hipHostFree(w.hv);
    w.hv = nullptr;
    w.cap = 0;
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      SPD_HIP(hipHostMalloc((void **) &w.hind[m], nc * sizeof(uint32_t), hipHostMallocDefault));
    }
    SPD_HIP(hipHostMalloc((void **) &w.hv, nc * sizeof(double), hipHostMallocDefault));
    w.cap = nc;
  }
  std::vector<uint32_t> order;
  slice_order(sl, N, t, s->cap, order);
  uint32_t * const * hind = w.hind;
  double * const hv = w.hv;
  double xn = 0.;
  for(uint64_t x=0; x < n; ++x) {
    hv[x] = sl->vals[order[x]];
    xn += hv[x] * hv[x];
  }
  // every index is checked before any is labelled, so a rejected slice
  // leaves the stream's labels as they were
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    for(uint64_t x=0; x < n; ++x) {
      uint64_t const o = sl->ind[m][x];
      if(o >= s->cap[m]) {
        spd_set_error(SPD_EINPUT, "serve: index %llu of mode %d beyond its capacity %llu",
            (unsigned long long) o, m, (unsigned long long) s->cap[m]);
        return SPD_EINPUT;
      }
    }
  }
  for(int m=0; m < N; ++m) {
    s->state->batch_max[m] = 0;
    if(m == t) continue;
    for(uint64_t x=0; x < n; ++x) {
      uint64_t const o = sl->ind[m][order[x]];
      if(s->lab[m][o] == UINT32_MAX) s->lab[m][o] = (uint32_t) s->next_label[m]++;
      uint32_t const l = s->lab[m][o];
      hind[m][x] = l;
      if(l > s->state->batch_max[m]) s->state->batch_max[m] = l;
    }
  }
  if(n > s->slice_cap) {
    uint64_t nc = s->slice_cap ? s->slice_cap : 4096;
    while(nc < n) nc *= 2;
    for(int m=0; m < N; ++m) {
      (void) spd_free_on(s->d_ind[m], s->hs);
      s->d_ind[m] = nullptr;
      if(m != t) SPD_TRY(spd_malloc_on((void **) &s->d_ind[m], nc * sizeof(uint32_t), s->hs));
    }
    (void) spd_free_on(s->d_vals, s->hs);
    SPD_TRY(spd_malloc_on((void **) &s->d_vals, nc * sizeof(double), s->hs));
    s->slice_cap = nc;
  }
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    SPD_HIP(hipMemcpyAsync(s->d_ind[m], hind[m], n * sizeof(uint32_t),
        hipMemcpyHostToDevice, s->hs));
  }
  SPD_HIP(hipMemcpyAsync(s->d_vals, hv, n * sizeof(double), hipMemcpyHostToDevice, s->hs));
  uint32_t const * ind[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) ind[m] = s->d_ind[m];
  SPD_TRY(s->state->slice(ind, s->d_vals, n, xn, &sl->local_err, &sl->outer, &sl->gpu_s));
  uint32_t const R = s->state->R;
  size_t const off = s->time_rows.size();
  s->time_rows.resize(off + R);
  SPD_HIP(hipMemcpyAsync(s->time_rows.data() + off, s->state->time_row, R * sizeof(double),
      hipMemcpyDeviceToHost, s->hs));
  SPD_HIP(hipStreamSynchronize(s->hs));
  return SPD_OK;
}

// Every exit leaves the slice's work finished: the stream's next slice may
// run on another worker's stream, which waits only for its own work.
int run_slice(Stream * s, Slice * sl, WorkerRes & w)
{
  int const rc = run_slice_body(s, sl, w);
  if(rc != SPD_OK && s->hs != nullptr) (void) hipStreamSynchronize(s->hs);
  return rc;
}

void worker_main(spd_server * sv, int wi)
{
  (void) hipSetDevice(0);
  std::unique_lock<std::mutex> lk(sv->mu);
  while(true) {
    sv->cv_work.wait(lk, [&] { return sv->stop || !sv->ready.empty(); });
    if(sv->stop) break;
    int64_t const id = sv->ready.front();
    sv->ready.pop_front();
    Stream * s = sv->streams[id];
    s->busy = true;
    Slice * sl = s->queue.front();
    s->queue.pop_front();
    lk.unlock();
    int const rc = run_slice(s, sl, sv->wres[wi]);
    spd_scratch_bind(nullptr);
    double const now = sst_wall_s();
    lk.lock();
    sl->t_done = now;
    sl->status = rc;
    if(rc != SPD_OK) snprintf(sl->err, sizeof(sl->err), "%s", spd_last_error());
    if(!s->queue.empty()) {
      s->busy = false;
      --sv->pending;
      sv->ready.push_back(id);
      sv->cv_work.notify_one();
    } else if(s->closing) {
      // the copy-back and frees run outside the lock; the stream stays busy
      // and its slice pending until they are done
      lk.unlock();
      release_device(s);
      lk.lock();
      s->busy = false;
      --sv->pending;
      release_budget(sv, s);
      sv->cv_work.notify_all();
    } else {
      s->busy = false;
      --sv->pending;
    }
    sv->cv_done.notify_all();
  }
}

// ---- batched mode -----------------------------------------------------------

// Streams whose every mode the batched step handles: rank <= 64, ADMM with
// an entry-wise or column-norm prox, or closed forms.
bool stream_batchable(Stream const * s)
{
  if(s->cfg.rank > 64) return false;
  for(int m=0; m < s->N; ++m) {
    spd_con const & c = s->cfg.con[m];
    if(c.kind != SPD_CON_ADMM) continue;
    if(m == s->cfg.stream_mode) return false;
    switch(c.prox) {
    case SPD_PROX_MAXCOLNORM: case SPD_PROX_MAXCOLNORM_NONNEG: case SPD_PROX_NONNEG:
    case SPD_PROX_LASSO: case SPD_PROX_NTF_FROB: case SPD_PROX_NTF_LASSO:
      break;
    default:
      return false;
    }
  }
  return true;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t) 255; }

// Host part of a slice for the batched step (the per-slice path does the
// same in run_slice_body and StreamState::slice): the parser's order and
// first-appearance labels, the running dims, the new rows' rand() values
// (drawn in mode order, as grow_mats does), a copy of the nonzeros ordered
// by each mode other than the streamed one and the one after it (a stable
// counting sort by label, the order the device radix sort gives), and the
// slice's squared norm.
int prepare_slice(Stream * s, Slice * sl, PSlice * ps)
{
  int const N = s->N;
  int const t = s->cfg.stream_mode;
  int const lead = (t + 1) % N;
  uint32_t const R = (uint32_t) s->cfg.rank;
  uint64_t const n = sl->nnz;
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    for(uint64_t x=0; x < n; ++x) {
      uint64_t const o = sl->ind[m][x];
      if(o >= s->cap[m]) {
        spd_set_error(SPD_EINPUT, "serve: index %llu of mode %d beyond its capacity %llu",
            (unsigned long long) o, m, (unsigned long long) s->cap[m]);
        return SPD_EINPUT;
      }
    }
  }
  std::vector<uint32_t> order;
  slice_order(sl, N, t, s->cap, order);
  // layout: copy lead, the other copies, then the new rows
  size_t off = 0;
  int copies[SPD_MAX_MODES];
  int nc = 0;
  copies[nc++] = lead;
  for(int c=0; c < N; ++c) if(c != t && c != lead) copies[nc++] = c;
  for(int q=0; q < nc; ++q) {
    int const c = copies[q];
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      ps->off_ind[c][m] = off;
      off = align256(off + n * sizeof(uint32_t));
    }
    ps->off_vals[c] = off;
    off = align256(off + n * sizeof(double));
  }
  // labels and running dims
  uint64_t bmax[SPD_MAX_MODES] = {};
  std::vector<uint32_t> lab_tmp[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) {
    if(m == t) continue;
    lab_tmp[m].resize(n);
    uint32_t * const labs = s->lab[m].data();
    for(uint64_t x=0; x < n; ++x) {
      uint64_t const o = sl->ind[m][order[x]];
      if(labs[o] == UINT32_MAX) labs[o] = (uint32_t) s->next_label[m]++;
      uint32_t const l = labs[o];
      lab_tmp[m][x] = l;
      if(l > bmax[m]) bmax[m] = l;
    }
  }
  uint64_t bd[SPD_MAX_MODES];
  for(int m=0; m < N; ++m) {
    if(m == t) { bd[m] = 1; continue; }
    bd[m] = (bmax[m] + 1 > s->prep_rows[m]) ? bmax[m] + 1 : s->prep_rows[m];
    ps->grow0[m] = s->prep_rows[m];
    ps->rows[m] = bd[m];
    if(bd[m] > ps->grow0[m]) {
      ps->off_grow[m] = off;
      off = align256(off + (bd[m] - ps->grow0[m]) * R * sizeof(double));
    }
  }
  ps->rows[t] = 1;
  ps->bytes = off;
  if(off > ps->cap) {
    if(ps->buf) (void) hipHostFree(ps->buf);
    ps->buf = nullptr;
    ps->cap = 0;
    size_t nc2 = 1 << 20;
    while(nc2 < off) nc2 *= 2;
    SPD_HIP(hipHostMalloc((void **) &ps->buf, nc2, hipHostMallocDefault));
    ps->cap = nc2;
  }
  uint8_t * const B = ps->buf;
  // copy lead: the slice's own order
  double xn = 0.;
  {
    double * const v = (double *) (B + ps->off_vals[lead]);
    for(uint64_t x=0; x < n; ++x) {
      v[x] = sl->vals[order[x]];
      xn += v[x] * v[x];
    }
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      memcpy(B + ps->off_ind[lead][m], lab_tmp[m].data(), n * sizeof(uint32_t));
    }
  }
  // the other copies: stable counting sort by that mode's label
  std::vector<uint64_t> cnt;
  std::vector<uint32_t> pos(n);
  for(int q=1; q < nc; ++q) {
    int const c = copies[q];
    cnt.assign(bd[c] + 1, 0);
    for(uint64_t x=0; x < n; ++x) ++cnt[lab_tmp[c][x] + 1];
    for(uint64_t k=0; k < bd[c]; ++k) cnt[k + 1] += cnt[k];
    for(uint64_t x=0; x < n; ++x) pos[cnt[lab_tmp[c][x]]++] = (uint32_t) x;
    double const * const v0 = (double const *) (B + ps->off_vals[lead]);
    double * const v = (double *) (B + ps->off_vals[c]);
    for(uint64_t x=0; x < n; ++x) v[x] = v0[pos[x]];
    for(int m=0; m < N; ++m) {
      if(m == t) continue;
      uint32_t * const d = (uint32_t *) (B + ps->off_ind[c][m]);
      uint32_t const * const src = lab_tmp[m].data();
      for(uint64_t x=0; x < n; ++x) d[x] = src[pos[x]];
    }
  }
  // new rows, modes in order (grow_mats)
  for(int m=0; m < N; ++m) {
    if(m == t || bd[m] <= ps->grow0[m]) continue;
    fill_rand_state(&s->rd, (double *) (B + ps->off_grow[m]), (bd[m] - ps->grow0[m]) * R);
    s->prep_rows[m] = bd[m];
  }
  // the local error's mode: the smallest dim, first on ties (the streamed
  // mode's dim is 1)
  int small = 0;
  for(int m=1; m < N; ++m) if(bd[m] < bd[small]) small = m;
  ps->small = small;
  ps->first = (s->prep_count == 0) ? 1 : 0;
  ++s->prep_count;
  ps->xnorm = xn;
  ps->nnz = n;
  ps->sl = sl;
  return SPD_OK;
}

PSlice * pool_get(spd_server * sv)
{
  // caller holds sv->mu
  if(!sv->pool.empty()) {
    PSlice * p = sv->pool.back();
    sv->pool.pop_back();
    return p;
  }
  return new PSlice();
}

void pool_put(spd_server * sv, PSlice * p)
{
  // caller holds sv->mu
  p->sl = nullptr;
  sv->pool.push_back(p);
}

// slices prepared ahead of the device step, per stream
#define PREP_AHEAD 2

void finish_stream_locked(spd_server * sv, Stream * s, std::unique_lock<std::mutex> & lk);

void prep_main(spd_server * sv)
{
  std::unique_lock<std::mutex> lk(sv->mu);
  while(true) {
    Stream * pick = nullptr;
    sv->cv_prep.wait(lk, [&] {
      if(sv->stop) return true;
      double best = 0.;
      for(auto & kv : sv->streams) {
        Stream * s = kv.second;
        if(!s->batchable || s->prepping || s->queue.empty() ||
            s->prepared.size() >= PREP_AHEAD) continue;
        double const ts = s->queue.front()->t_submit;
        if(pick == nullptr || ts < best) { pick = s; best = ts; }
      }
      return pick != nullptr;
    });
    if(sv->stop) break;
    Stream * s = pick;
    s->prepping = true;
    Slice * sl = s->queue.front();
    s->queue.pop_front();
    PSlice * ps = pool_get(sv);
    lk.unlock();
    int const rc = prepare_slice(s, sl, ps);
    lk.lock();
    s->prepping = false;
    if(rc == SPD_OK) {
      s->prepared.push_back(ps);
    } else {
      pool_put(sv, ps);
      sl->t_done = sst_wall_s();
      sl->status = rc;
      snprintf(sl->err, sizeof(sl->err), "%s", spd_last_error());
      --sv->pending;
      if(s->closing && s->queue.empty() && s->prepared.empty() && !s->busy) {
        s->busy = true;
        finish_stream_locked(sv, s, lk);
      }
      sv->cv_done.notify_all();
    }
    sv->cv_work.notify_all();
    sv->cv_prep.notify_all();
  }
}

// the shape a batch shares: modes, streamed mode, padded rank
uint64_t stream_class(Stream const * s)
{
  return ((uint64_t) s->N << 16) | ((uint64_t) s->cfg.stream_mode << 8) |
      ((s->cfg.rank + 15) / 16);
}

bool stream_ready(Stream const * s)
{
  if(!s->admitted || s->busy) return false;
  return s->batchable ? !s->prepared.empty() : !s->queue.empty();
}

// Device part of a batch of prepared slices (the streams are busy: nobody
// else touches them).
int run_batch(spd_server * sv, BtEngine * e, WorkerRes & w, std::vector<Stream *> & ss,
    std::vector<PSlice *> & pp)
{
  hipStream_t const est = bt_engine_stream(e);
  uint32_t const K = (uint32_t) ss.size();
  std::vector<BtSlice> bs(K);
  std::vector<BtSlice *> ptrs(K);
  for(uint32_t q=0; q < K; ++q) {
    Stream * s = ss[q];
    PSlice * ps = pp[q];
    int const N = s->N;
    int const t = s->cfg.stream_mode;
    uint32_t const R = (uint32_t) s->cfg.rank;
    s->hs = est;
    if(!s->allocated) {
      s->allocated = true;
      s->state = new StreamState();
      SPD_TRY(s->state->init(N, s->cfg, s->cap, est, 1));
      uint64_t maxcap = 1;
      for(int m=0; m < N; ++m) if(m != t && s->cap[m] > maxcap) maxcap = s->cap[m];
      SPD_TRY(spd_malloc_on((void **) &s->d_part, bt_part_doubles(R, maxcap) * sizeof(double),
          est));
    } else {
      s->state->set_stream(est);
    }
    if(ps->bytes > s->d_slice_cap) {
      size_t nc = s->d_slice_cap ? s->d_slice_cap : (1 << 20);
      while(nc < ps->bytes) nc *= 2;
      (void) spd_free_on(s->d_slice, est);
      s->d_slice = nullptr;
      s->d_slice_cap = 0;
      SPD_TRY(spd_malloc_on((void **) &s->d_slice, nc, est));
      s->d_slice_cap = nc;
    }
    SPD_HIP(hipMemcpyAsync(s->d_slice, ps->buf, ps->bytes, hipMemcpyHostToDevice, est));
    BtSlice & b = bs[q];
    b.S = s->state;
    b.nnz = ps->nnz;
    int const lead = (t + 1) % N;
    for(int m=0; m < N; ++m) {
      b.rows[m] = ps->rows[m];
      b.grow0[m] = ps->grow0[m];
      b.d_grow[m] = (m != t && ps->rows[m] > ps->grow0[m]) ?
          (double const *) (s->d_slice + ps->off_grow[m]) : nullptr;
      for(int c=0; c < N; ++c) {
        if(c == t || m == t) continue;
        b.d_ind[c][m] = (uint32_t const *) (s->d_slice + ps->off_ind[c][m]);
      }
    }
    for(int c=0; c < N; ++c) {
      if(c != t) b.d_vals[c] = (double const *) (s->d_slice + ps->off_vals[c]);
    }
    (void) lead;
    b.d_part = s->d_part;
    b.first = ps->first;
    b.small = ps->small;
    b.xnorm = ps->xnorm;
    ptrs[q] = &b;
  }
  double const t0 = sst_wall_s();
  SPD_TRY(bt_engine_run(e, ptrs.data(), K));
  double const secs = sst_wall_s() - t0;
  for(uint32_t q=0; q < K; ++q) {
    Stream * s = ss[q];
    PSlice * ps = pp[q];
    BtSlice const & b = bs[q];
    for(int m=0; m < s->N; ++m) {
      if(m == s->cfg.stream_mode) continue;
      s->state->rows[m] = ps->rows[m];
      s->state->prev_dim[m] = ps->rows[m];
    }
    ++s->state->nslices;
    s->time_rows.insert(s->time_rows.end(), b.time_row.begin(), b.time_row.end());
    ps->sl->local_err = b.local_err;
    ps->sl->outer = b.outer;
    ps->sl->gpu_s = secs;
  }
  (void) w;
  (void) sv;
  return SPD_OK;
}

void finish_stream_locked(spd_server * sv, Stream * s, std::unique_lock<std::mutex> & lk)
{
  // caller holds lk; the stream is busy
  bool const more = s->batchable ? (!s->prepared.empty() || !s->queue.empty() || s->prepping)
                                 : !s->queue.empty();
  if(!more && s->closing) {
    lk.unlock();
    release_device(s);
    lk.lock();
    s->busy = false;
    release_budget(sv, s);
  } else {
    s->busy = false;
  }
}

void engine_main(spd_server * sv, int ei)
{
  (void) hipSetDevice(0);
  BtEngine * const e = sv->eng[ei];
  WorkerRes & w = sv->wres[ei];
  std::unique_lock<std::mutex> lk(sv->mu);
  while(true) {
    std::vector<Stream *> ss;
    sv->cv_work.wait(lk, [&] {
      if(sv->stop) return true;
      // the longest-waiting ready stream decides the batch's class
      Stream * first = nullptr;
      double best = 0.;
      for(auto & kv : sv->streams) {
        Stream * s = kv.second;
        if(!stream_ready(s)) continue;
        double const ts = s->batchable ? s->prepared.front()->sl->t_submit
                                       : s->queue.front()->t_submit;
        if(first == nullptr || ts < best) { first = s; best = ts; }
      }
      if(first == nullptr) return false;
      ss.push_back(first);
      if(first->batchable) {
        uint64_t const cls = stream_class(first);
        std::vector<std::pair<double, Stream *>> cand;
        for(auto & kv : sv->streams) {
          Stream * s = kv.second;
          if(s == first || !s->batchable || !stream_ready(s) || stream_class(s) != cls) continue;
          cand.emplace_back(s->prepared.front()->sl->t_submit, s);
        }
        std::sort(cand.begin(), cand.end(), [](auto const & a, auto const & b) {
          return a.first < b.first; });
        for(auto & c : cand) {
          if((int) ss.size() >= sv->max_inflight || ss.size() >= BT_MAXK) break;
          ss.push_back(c.second);
        }
      }
      return true;
    });
    if(sv->stop) break;
    for(Stream * s : ss) s->busy = true;
    if(!ss[0]->batchable) {
      // the per-slice path, one slice
      Stream * s = ss[0];
      Slice * sl = s->queue.front();
      s->queue.pop_front();
      lk.unlock();
      int const rc = run_slice(s, sl, w);
      spd_scratch_bind(nullptr);
      double const now = sst_wall_s();
      lk.lock();
      sl->t_done = now;
      sl->status = rc;
      if(rc != SPD_OK) snprintf(sl->err, sizeof(sl->err), "%s", spd_last_error());
      --sv->pending;
      finish_stream_locked(sv, s, lk);
      sv->cv_work.notify_all();
      sv->cv_done.notify_all();
      continue;
    }
    std::vector<PSlice *> pp;
    for(Stream * s : ss) {
      pp.push_back(s->prepared.front());
      s->prepared.pop_front();
    }
    sv->cv_prep.notify_all();
    lk.unlock();
    int const rc = run_batch(sv, e, w, ss, pp);
    char msg[160] = "";
    if(rc != SPD_OK) snprintf(msg, sizeof(msg), "%s", spd_last_error());
    if(rc != SPD_OK) (void) hipStreamSynchronize(bt_engine_stream(e));
    double const now = sst_wall_s();
    lk.lock();
    for(size_t q=0; q < ss.size(); ++q) {
      Slice * sl = pp[q]->sl;
      sl->t_done = now;
      sl->status = rc;
      if(rc != SPD_OK) snprintf(sl->err, sizeof(sl->err), "%s", msg);
      --sv->pending;
      pool_put(sv, pp[q]);
    }
    for(Stream * s : ss) finish_stream_locked(sv, s, lk);
    sv->cv_work.notify_all();
    sv->cv_prep.notify_all();
    sv->cv_done.notify_all();
  }
}

}  // namespace

// The batched step (spd_batch.hip) is switched off: its first run on the
// GPU (several streams with ranks 4 to 64 in flight) ended in an illegal
// memory access whose cause has not been found. With it off, slices of
// different streams run concurrently on per-worker HIP streams, one slice
// per launch sequence.
#define SPD_SERVE_BATCHED 0

extern "C" int spd_server_create(size_t budget, int max_inflight, spd_server ** out)
{
  *out = nullptr;
  SPD_TRY(spd_init());
  if(max_inflight < 1 || max_inflight > 64) {
    spd_set_error(SPD_EINPUT, "serve: 1 .. 64 slices in flight (got %d)", max_inflight);
    return SPD_EINPUT;
  }
  if(budget == 0) {
    size_t fr = 0, tot = 0;
    SPD_HIP(hipMemGetInfo(&fr, &tot));
    budget = fr / 10 * 8;
  }
  spd_server * sv = new spd_server();
  sv->budget = budget;
  sv->max_inflight = max_inflight;
  if(max_inflight > 1 && SPD_SERVE_BATCHED) {
    // batched: engines (each its own HIP stream) run batches of up to
    // max_inflight slices of one shape class; host threads prepare slices
    sv->batched = true;
    int const ne = BT_ENGINES;
    sv->wres.resize(ne);
    for(int k=0; k < ne; ++k) {
      BtEngine * e = nullptr;
      int const rc = bt_engine_create(&e);
      if(rc != SPD_OK) {
        for(BtEngine * x : sv->eng) bt_engine_free(x);
        delete sv;
        return rc;
      }
      sv->eng.push_back(e);
      sv->wres[k].hs = bt_engine_stream(e);
    }
    for(int k=0; k < ne; ++k) sv->engines.emplace_back(engine_main, sv, k);
    for(int k=0; k < BT_PREP_THREADS; ++k) sv->preppers.emplace_back(prep_main, sv);
    *out = sv;
    return SPD_OK;
  }
  sv->wres.resize(max_inflight);
  for(int w=0; w < max_inflight; ++w) {
    // default priority, non-blocking w.r.t. the legacy null stream
    if(hipStreamCreateWithFlags(&sv->wres[w].hs, hipStreamNonBlocking) != hipSuccess) {
      for(int k=0; k < w; ++k) (void) hipStreamDestroy(sv->wres[k].hs);
      delete sv;
      spd_set_error(SPD_EHIP, "serve: cannot create worker streams");
      return SPD_EHIP;
    }
  }
  for(int w=0; w < max_inflight; ++w) sv->workers.emplace_back(worker_main, sv, w);
  *out = sv;
  return SPD_OK;
}

extern "C" int spd_server_open(spd_server * sv, int N, uint64_t const * cap,
    spd_stream_cfg const * cfg, unsigned seed, int64_t * id_out)
{
  if(N < 2 || N > SPD_MAX_MODES || cfg->stream_mode < 0 || cfg->stream_mode >= N ||
      cfg->rank == 0 || cfg->rank > 256) {
    spd_set_error(SPD_EINPUT, "serve: bad stream shape");
    return SPD_EINPUT;
  }
  SPD_TRY(sst_check_cfg(N, *cfg));
  Stream * s = new Stream();
  s->N = N;
  s->cfg = *cfg;
  s->cfg.fill_rand = fill_stream_rand;
  s->cfg.rand_ctx = s;
  for(int m=0; m < N; ++m) {
    s->cap[m] = (m == cfg->stream_mode) ? 1 : cap[m];
    if(m != cfg->stream_mode) s->lab[m].assign(cap[m], UINT32_MAX);
  }
  initstate_r(seed, s->rstate, sizeof(s->rstate), &s->rd);
  s->footprint = StreamState::footprint(N, cfg->stream_mode, cfg->rank, s->cap);
  s->batchable = sv->batched && stream_batchable(s);
  if(s->batchable) {
    uint64_t maxcap = 1;
    for(int m=0; m < N; ++m) {
      if(m != cfg->stream_mode && s->cap[m] > maxcap) maxcap = s->cap[m];
    }
    s->footprint += bt_part_doubles((uint32_t) cfg->rank, maxcap) * sizeof(double);
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  if(s->footprint > sv->budget) {
    delete s;
    spd_set_error(SPD_ENOMEM, "serve: stream needs more than the whole HBM budget");
    return SPD_ENOMEM;
  }
  s->id = sv->next_stream++;
  sv->streams[s->id] = s;
  sv->waiting.push_back(s->id);
  admit_waiting(sv);
  *id_out = s->id;
  return SPD_OK;
}

extern "C" int spd_server_submit(spd_server * sv, int64_t id, uint64_t nnz,
    uint64_t const * const * inds, double const * vals, int64_t * ticket_out)
{
  double const now = sst_wall_s();
  Slice * sl = new Slice();
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->streams.find(id);
  if(it == sv->streams.end() || it->second->closing || nnz == 0) {
    delete sl;
    spd_set_error(SPD_EINPUT, "serve: bad stream %lld or empty slice", (long long) id);
    return SPD_EINPUT;
  }
  Stream * s = it->second;
  sl->ticket = (int64_t) sv->slices.size();
  sl->stream = id;
  sl->nnz = nnz;
  for(int m=0; m < s->N; ++m) {
    if(m == s->cfg.stream_mode) continue;
    sl->ind[m].assign(inds[m], inds[m] + nnz);
  }
  sl->vals.assign(vals, vals + nnz);
  sl->t_submit = now;
  sv->slices.push_back(sl);
  s->queue.push_back(sl);
  ++sv->pending;
  if(sv->batched) {
    if(s->batchable) sv->cv_prep.notify_one();
    else sv->cv_work.notify_all();
  } else if(s->admitted && !s->busy && s->queue.size() == 1) {
    sv->ready.push_back(id);
    sv->cv_work.notify_one();
  }
  *ticket_out = sl->ticket;
  return SPD_OK;
}

extern "C" int spd_server_drain(spd_server * sv)
{
  std::unique_lock<std::mutex> lk(sv->mu);
  bool stuck = false;
  sv->cv_done.wait(lk, [&] {
    if(sv->pending == 0) return true;
    // the remaining slices belong to streams that wait for memory while no
    // admitted stream has work left: only closing streams frees memory
    bool running = !sv->ready.empty();
    for(auto & kv : sv->streams) {
      Stream const * s = kv.second;
      if(s->busy || s->prepping || (s->admitted && (!s->queue.empty() || !s->prepared.empty())))
        running = true;
    }
    stuck = !running;
    return stuck;
  });
  if(stuck) {
    spd_set_error(SPD_ENOMEM, "serve: slices wait for HBM that only closing a stream frees");
    return SPD_ENOMEM;
  }
  return SPD_OK;
}

extern "C" int spd_server_slice(spd_server * sv, int64_t ticket, double * latency_s,
    double * gpu_s, double * local_err, uint64_t * outer)
{
  std::lock_guard<std::mutex> lk(sv->mu);
  if(ticket < 0 || ticket >= (int64_t) sv->slices.size()) {
    spd_set_error(SPD_EINPUT, "serve: bad ticket");
    return SPD_EINPUT;
  }
  Slice const * sl = sv->slices[ticket];
  if(sl->status > 0) return 1;
  if(sl->status < 0) {
    spd_set_error(sl->status, "%s", sl->err);
    return sl->status;
  }
  if(latency_s) *latency_s = sl->t_done - sl->t_submit;
  if(gpu_s) *gpu_s = sl->gpu_s;
  if(local_err) *local_err = sl->local_err;
  if(outer) *outer = sl->outer;
  return SPD_OK;
}

extern "C" int spd_server_factors(spd_server * sv, int64_t id, int mode, double * out,
    uint64_t * rows)
{
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->streams.find(id);
  if(it == sv->streams.end()) {
    spd_set_error(SPD_EINPUT, "serve: bad stream");
    return SPD_EINPUT;
  }
  Stream * s = it->second;
  sv->cv_done.wait(lk, [&] {
    return s->queue.empty() && s->prepared.empty() && !s->prepping && !s->busy;
  });
  if(mode < 0 || mode >= s->N) {
    spd_set_error(SPD_EINPUT, "serve: bad mode %d", mode);
    return SPD_EINPUT;
  }
  uint32_t const R = (uint32_t) s->cfg.rank;
  if(mode == s->cfg.stream_mode) {
    *rows = s->time_rows.size() / R;
    if(out) memcpy(out, s->time_rows.data(), s->time_rows.size() * sizeof(double));
    return SPD_OK;
  }
  if(s->state == nullptr) {
    *rows = s->has_kept ? s->kept_rows[mode] : 0;
    if(out && *rows) memcpy(out, s->kept[mode].data(), *rows * R * sizeof(double));
    return SPD_OK;
  }
  if(mode == s->cfg.stream_mode) {
    *rows = s->time_rows.size() / R;
    if(out) memcpy(out, s->time_rows.data(), s->time_rows.size() * sizeof(double));
    return SPD_OK;
  }
  *rows = s->state->rows[mode];
  if(out) {
    SPD_HIP(hipMemcpyAsync(out, s->state->Unew[mode], *rows * R * sizeof(double),
        hipMemcpyDeviceToHost, s->hs));
    SPD_HIP(hipStreamSynchronize(s->hs));
  }
  return SPD_OK;
}

extern "C" int spd_server_close(spd_server * sv, int64_t id)
{
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->streams.find(id);
  if(it == sv->streams.end()) {
    spd_set_error(SPD_EINPUT, "serve: bad stream");
    return SPD_EINPUT;
  }
  Stream * s = it->second;
  if(s->closing) return SPD_OK;
  s->closing = true;
  if(s->queue.empty() && s->prepared.empty() && !s->prepping && !s->busy) {
    for(auto w = sv->waiting.begin(); w != sv->waiting.end(); ++w) {
      if(*w == id) { sv->waiting.erase(w); break; }
    }
    s->busy = true;
    lk.unlock();
    release_device(s);
    lk.lock();
    s->busy = false;
    release_budget(sv, s);
    sv->cv_work.notify_all();
    sv->cv_done.notify_all();
  }
  return SPD_OK;
}

extern "C" void spd_server_stats(spd_server * sv, size_t * used, size_t * budget,
    int * admitted, int * waiting)
{
  std::lock_guard<std::mutex> lk(sv->mu);
  if(used) *used = sv->used;
  if(budget) *budget = sv->budget;
  int a = 0;
  for(auto & kv : sv->streams) a += kv.second->admitted ? 1 : 0;
  if(admitted) *admitted = a;
  if(waiting) *waiting = (int) sv->waiting.size();
}

extern "C" void spd_server_free(spd_server * sv)
{
  if(sv == nullptr) return;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] {
      for(auto & kv : sv->streams) if(kv.second->busy) return false;
      return true;
    });
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  sv->cv_prep.notify_all();
  for(auto & w : sv->workers) w.join();
  for(auto & w : sv->engines) w.join();
  for(auto & w : sv->preppers) w.join();
  for(auto & kv : sv->streams) {
    kv.second->closing = true;
    kv.second->busy = false;
    release_device(kv.second);
    release_budget(sv, kv.second);
    delete kv.second;
  }
  for(Slice * sl : sv->slices) delete sl;
  for(PSlice * p : sv->pool) {
    if(p->buf) (void) hipHostFree(p->buf);
    delete p;
  }
  for(auto & kv : sv->streams) (void) kv;
  for(WorkerRes & w : sv->wres) {
    (void) hipStreamSynchronize(w.hs);
    for(int m=0; m < SPD_MAX_MODES; ++m) {
      if(w.hind[m]) (void) hipHostFree(w.hind[m]);
    }
    if(w.hv) (void) hipHostFree(w.hv);
    if(!sv->batched) (void) hipStreamDestroy(w.hs);
  }
  for(BtEngine * e : sv->eng) bt_engine_free(e);
  delete sv;
}

// The per-stream generator alone (tests compare it with glibc's random_r).
extern "C" void spd_serve_rand_fill(unsigned seed, double * vals, uint64_t n)
{
  struct random_data rd = {};
  char st[128] = {};
  initstate_r(seed, st, sizeof(st), &rd);
  fill_rand_state(&rd, vals, n);
}
