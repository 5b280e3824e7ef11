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
    if(!s->queue.empty() && !s->busy) sv->ready.push_back(s->id);
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

}  // namespace

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
  if(s->admitted && !s->busy && s->queue.size() == 1) {
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
      if(s->busy || (s->admitted && !s->queue.empty())) running = true;
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
  sv->cv_done.wait(lk, [&] { return s->queue.empty() && !s->busy; });
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
  if(s->queue.empty() && !s->busy) {
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
  for(auto & w : sv->workers) w.join();
  for(auto & kv : sv->streams) {
    kv.second->closing = true;
    kv.second->busy = false;
    release_device(kv.second);
    release_budget(sv, kv.second);
    delete kv.second;
  }
  for(Slice * sl : sv->slices) delete sl;
  for(WorkerRes & w : sv->wres) {
    (void) hipStreamSynchronize(w.hs);
    for(int m=0; m < SPD_MAX_MODES; ++m) {
      if(w.hind[m]) (void) hipHostFree(w.hind[m]);
    }
    if(w.hv) (void) hipHostFree(w.hv);
    (void) hipStreamDestroy(w.hs);
  }
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
