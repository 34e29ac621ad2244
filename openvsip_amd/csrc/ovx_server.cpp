// ovx_server.cpp -- continuous-batching request server (C-ABI part).
//
// Requests (fast convolution against a resident replica, forward Fftm,
// FFT-based correlation against a prepared reference) arrive on a queue.  A
// scheduler thread takes a free slot -- a stream with its own device tile
// tables, four-step workspace and a hipGraph of tile-queue launches (the
// persistent launches, one per four-step pass and line length) captured once
// at start-up -- admits queued requests into it in
// arrival order while their device working sets fit the HBM budget, uploads
// the slot's tables and replays the graph.  A completion thread retires slots
// in launch order, stamps completion times and returns their budget.  With
// several slots, requests that arrive while others run are admitted as slots
// free up and their batches overlap on the device.  In plain mode every
// request is its own slot and runs the standalone kernels (the
// one-request-per-launch path the batched server is compared with).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "ovx_batch.hpp"
#include "ovx_batch_host.hpp"
#include "ovx_hier.hpp"
#include "ovx_ols.hpp"
#include "ovx_plan.hpp"

namespace ovx {

// Batches of at most this many requests run as standalone launches.
constexpr int kDirectMax = 4;

static double now_s()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- correlation references
// A correlation reference prepared once: its overlap-save segment length,
// the register-Stockham path of that length and H = FFT_L(reversed ref).
struct CorrRef {
  int support = 0;
  long m = 0, n = 0, p = 0, out_off = 0;
  int L = 0;
  LineFFT f;
  DevBuf H;
};
static std::mutex g_cref_mu;
static std::map<int, std::unique_ptr<CorrRef>> g_crefs;
static int g_next_cref = 1;

static int corr_ref_get(int id, const CorrRef **out)
{
  std::lock_guard<std::mutex> g(g_cref_mu);
  auto it = g_crefs.find(id);
  if (it == g_crefs.end()) return fail(OVX_ERR_INVALID, "unknown correlation reference %d", id);
  *out = it->second.get();
  return OVX_OK;
}

static OlsArgs corr_ols_args(const CorrRef &c, const float2 *in, float2 *out, int bias)
{
  OlsArgs a;
  a.in = in;
  a.N = c.n;
  a.M = c.m;
  a.H = (const float2 *)c.H.p;
  a.tw = (const float2 *)c.f.tw.p;
  a.scale = 1.0f / (float)c.L;
  a.conj_input = 1;
  a.unbiased = bias == OVX_UNBIASED;
  a.support = c.support;
  a.decimation = 1;
  a.out_off = c.out_off;
  a.c_first = c.out_off;
  a.c_step = c.L - (c.m - 1);
  a.P = c.p;
  a.out = out;
  return a;
}

static long corr_segments(const CorrRef &c) { return (c.p - 1 + (c.L - (c.m - 1))) / (c.L - (c.m - 1)); }

// ---------------------------------------------------------------- server
struct Req {
  unsigned long id = 0;
  ovx_request r{};
  double t_submit = 0, t_done = 0;
  size_t footprint = 0, ws_need = 0;
  int status = -1; // -1 pending
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  DevBuf d_jobs, d_tiles[kBatchQueues], d_meta, ws;
  BatchJob *h_jobs = nullptr;
  BatchTile *h_tiles[kBatchQueues] = {};
  int *h_meta = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<Req *> reqs;
  size_t bytes = 0;
};

struct Server {
  ovx_server_config cfg{};
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_wait;
  std::deque<Req *> queue;
  std::deque<Slot *> free_slots, inflight;
  std::map<unsigned long, std::unique_ptr<Req>> reqs;
  unsigned long next_id = 1;
  size_t budget = 0, in_use = 0;
  bool stop = false;
  bool sched_done = false;       // the scheduler has exited: nothing more enters `inflight`
  unsigned long completions = 0; // bumped by every retired slot
  int error = OVX_OK;
  std::thread sched, compl_;
  long max_tiles = 0;
  int device = 0;

  ~Server();
  int init(const ovx_server_config &c);
  int build_slot(Slot &s);
  int admit_and_launch(Slot &s, std::vector<Req *> &batch);
  int launch_plain(Slot &s, Req &r);
  void scheduler();
  void completer();
};

static void free_pinned(void *p)
{
  if (p) (void)hipHostFree(p);
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> g(mu);
    stop = true;
  }
  cv_sched.notify_all();
  if (sched.joinable()) sched.join();
  // the completer retires every slot the scheduler launched before it exits
  {
    std::lock_guard<std::mutex> g(mu);
    sched_done = true;
  }
  cv_done.notify_all();
  if (compl_.joinable()) compl_.join();
  for (auto &s : slots) {
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->done) (void)hipEventDestroy(s->done);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    free_pinned(s->h_jobs);
    for (int k = 0; k < kBatchQueues; ++k) free_pinned(s->h_tiles[k]);
    free_pinned(s->h_meta);
  }
}

int Server::build_slot(Slot &s)
{
  hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
  if (e != hipSuccess) return hip_fail(e, "slot stream");
  int st;
  size_t jb = (size_t)cfg.max_jobs * sizeof(BatchJob), tb = (size_t)max_tiles * sizeof(BatchTile);
  if ((st = s.d_jobs.alloc(jb)) || (st = s.d_meta.alloc(kMetaInts * sizeof(int)))) return st;
  for (int k = 0; k < kBatchQueues; ++k)
    if ((st = s.d_tiles[k].alloc(tb))) return st;
  if (cfg.workspace_mb > 0 && (st = s.ws.alloc((size_t)cfg.workspace_mb << 20))) return st;
  e = hipHostMalloc((void **)&s.h_jobs, jb, hipHostMallocDefault);
  for (int k = 0; k < kBatchQueues && e == hipSuccess; ++k)
    e = hipHostMalloc((void **)&s.h_tiles[k], tb, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&s.h_meta, kMetaInts * sizeof(int), hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(e, "pinned staging");
  if (cfg.plain) return OVX_OK;
  // capture the launches (the persistent one and one per four-step queue)
  // once; every batch reuses them with new table contents (tile counts are
  // read from device memory, an empty queue's launch exits at once)
  hipStream_t cap;
  e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "capture stream");
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    BatchTile *tiles[kBatchQueues];
    for (int k = 0; k < kBatchQueues; ++k) tiles[k] = (BatchTile *)s.d_tiles[k].p;
    e = capture_batch_launches((const BatchJob *)s.d_jobs.p, tiles, (int *)s.d_meta.p, nullptr, true, cap);
    hipError_t e2 = hipStreamEndCapture(cap, &s.graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess) return hip_fail(e, "slot graph capture");
  return OVX_OK;
}

int Server::init(const ovx_server_config &c)
{
  cfg = c;
  if (cfg.slots < 1) cfg.slots = 1;
  if (cfg.max_jobs < 1) cfg.max_jobs = 1;
  if (cfg.plain) cfg.max_jobs = 1; // one request per launch, `slots` requests in flight
  budget = (size_t)std::max(cfg.hbm_budget_mb, 1l) << 20;
  max_tiles = cfg.max_tiles > 0 ? cfg.max_tiles : 65536;
  (void)hipGetDevice(&device);
  for (int i = 0; i < cfg.slots; ++i) {
    slots.emplace_back(new (std::nothrow) Slot());
    if (!slots.back()) return fail(OVX_ERR_NOMEM, "host allocation");
    int st = build_slot(*slots.back());
    if (st) return st;
    free_slots.push_back(slots.back().get());
  }
  sched = std::thread([this] { scheduler(); });
  compl_ = std::thread([this] { completer(); });
  return OVX_OK;
}

// Tiles a request needs (to respect the slot's table capacity).
static long tiles_of(const Req &q, const LineFFT *f, const CorrRef *cr)
{
  const ovx_request &r = q.r;
  if (r.kind == OVX_REQ_CORR) return (corr_segments(*cr) + batch_ols_segments_per_tile(cr->L) - 1) /
                                       batch_ols_segments_per_tile(cr->L);
  int path = batch_path_of(*f);
  if (path == BP_FS) { // one line group per tile, per pass
    long gc = f->n2 / batch_fs_lines_per_group((int)f->n1);
    long gr = f->n1 / batch_fs_lines_per_group((int)f->n2);
    return r.rows * std::max(gc, gr);
  }
  long rpt = batch_rows_per_tile(path, (int)r.n);
  return (r.rows + rpt - 1) / rpt;
}

int Server::admit_and_launch(Slot &s, std::vector<Req *> &batch)
{
  std::vector<BatchJob> jobs;
  std::vector<BatchTile> q[kBatchQueues];
  size_t ws_off = 0;
  for (size_t i = 0; i < batch.size(); ++i) {
    const ovx_request &r = batch[i]->r;
    int st;
    BatchJob j{};
    if (r.kind == OVX_REQ_CORR) {
      const CorrRef *cr;
      if ((st = corr_ref_get(r.ref, &cr))) return st;
      j.path = 0;
      j.n = (int)cr->n;
      j.ols = corr_ols_args(*cr, (const float2 *)r.x, (float2 *)r.y, r.bias);
      j.ols_len = cr->L;
      jobs.push_back(j);
      int spt = batch_ols_segments_per_tile(cr->L);
      std::vector<BatchTile> &dst = q[batch_ols_queue(cr->L)];
      for (long sgm = 0; sgm < corr_segments(*cr); sgm += spt) dst.push_back(BatchTile{ (int)i, BT_OLS, (int)sgm, 0 });
      continue;
    }
    const LineFFT *f = shared_path(r.n, r.rows, &st);
    if (!f) return st;
    if (r.kind == OVX_REQ_FASTCONV) {
      const float2 *prep;
      long rn;
      if ((st = replica_prepared(r.ref, &prep, &rn, f))) return st;
      if ((st = make_batch_job(*f, r.rows, (const float2 *)r.x, r.ldx, (float2 *)r.y, r.ldy, prep, &j))) return st;
      jobs.push_back(j);
      make_batch_tiles(j, (int)i, q);
    } else { // forward Fftm
      if ((st = make_batch_job(*f, r.rows, (const float2 *)r.x, r.ldx, (float2 *)r.y, r.ldy, nullptr, &j))) return st;
      j.scale = 1.0f;
      if (j.path == BP_FS) {
        j.ws = (float2 *)((char *)s.ws.p + ws_off);
        ws_off += batch[i]->ws_need;
        make_batch_fft_tiles(j, (int)i, q);
      } else {
        int rpt = batch_rows_per_tile(j.path, j.n);
        int op = j.path == BP_HIER ? BT_HIER_FFT_FWD : BT_ROW_FFT_FWD;
        std::vector<BatchTile> &dst = q[batch_queue(j.path, j.n)];
        for (long rr = 0; rr < j.rows; rr += rpt) dst.push_back(BatchTile{ (int)i, op, (int)rr, 0 });
      }
      jobs.push_back(j);
    }
  }
  for (int k = 0; k < kBatchQueues; ++k)
    if ((long)q[k].size() > max_tiles) return fail(OVX_ERR_STATE, "slot tile table overflow");
  std::copy(jobs.begin(), jobs.end(), s.h_jobs);
  int meta[kMetaInts] = { 0 };
  for (int k = 0; k < kBatchQueues; ++k) {
    std::copy(q[k].begin(), q[k].end(), s.h_tiles[k]);
    meta[k] = (int)q[k].size();
  }
  std::copy(meta, meta + kMetaInts, s.h_meta);
  hipError_t e = hipMemcpyAsync(s.d_jobs.p, s.h_jobs, jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice, s.stream);
  for (int k = 0; k < kBatchQueues && e == hipSuccess; ++k)
    if (!q[k].empty())
      e = hipMemcpyAsync(s.d_tiles[k].p, s.h_tiles[k], q[k].size() * sizeof(BatchTile), hipMemcpyHostToDevice,
                         s.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s.d_meta.p, s.h_meta, kMetaInts * sizeof(int), hipMemcpyHostToDevice, s.stream);
  if (e == hipSuccess) e = hipGraphLaunch(s.exec, s.stream);
  if (e != hipSuccess) return hip_fail(e, "batch launch");
  return OVX_OK;
}

int Server::launch_plain(Slot &s, Req &q)
{
  const ovx_request &r = q.r;
  int st;
  if (r.kind == OVX_REQ_CORR) {
    const CorrRef *cr;
    if ((st = corr_ref_get(r.ref, &cr))) return st;
    OlsArgs a = corr_ols_args(*cr, (const float2 *)r.x, (float2 *)r.y, r.bias);
    hipError_t e = launch_ols(cr->L, a, corr_segments(*cr), s.stream);
    if (e != hipSuccess) return hip_fail(e, "correlation launch");
    return OVX_OK;
  }
  const LineFFT *f = shared_path(r.n, r.rows, &st);
  if (!f) return st;
  if (r.kind == OVX_REQ_FASTCONV) {
    const float2 *prep;
    long rn;
    if ((st = replica_prepared(r.ref, &prep, &rn, f))) return st;
    return f->fastconv_prepared((const float2 *)r.x, r.ldx, prep, (float2 *)r.y, r.ldy, r.rows, s.stream);
  }
  return f->transform_ws(-1, (const float2 *)r.x, r.ldx, (float2 *)r.y, r.ldy, r.rows, 1.0f, (float2 *)s.ws.p,
                         s.stream);
}

void Server::scheduler()
{
  (void)hipSetDevice(device);
  for (;;) {
    std::vector<Req *> batch;
    Slot *s = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_sched.wait(lk, [&] { return stop || (!queue.empty() && !free_slots.empty()); });
      if (stop) return;
      s = free_slots.front();
      free_slots.pop_front();
      size_t ws_used = 0;
      long tiles = 0;
      while (!queue.empty() && (long)batch.size() < cfg.max_jobs) {
        Req *r = queue.front();
        if (r->footprint > budget || r->ws_need > s->ws.bytes) { // can never run: refuse it
          queue.pop_front();
          r->status = r->footprint > budget ? OVX_ERR_BUDGET : OVX_ERR_UNSUPPORTED;
          r->t_done = now_s();
          cv_wait.notify_all();
          continue;
        }
        if (in_use + r->footprint > budget) break; // wait for budget to free up
        if (ws_used + r->ws_need > s->ws.bytes) break;
        const LineFFT *f = nullptr;
        const CorrRef *cr = nullptr;
        int st = OVX_OK;
        if (r->r.kind == OVX_REQ_CORR) st = corr_ref_get(r->r.ref, &cr);
        else f = shared_path(r->r.n, r->r.rows, &st);
        if (st) {
          queue.pop_front();
          r->status = st;
          r->t_done = now_s();
          cv_wait.notify_all();
          continue;
        }
        long t = cfg.plain ? 0 : tiles_of(*r, f, cr);
        if (!batch.empty() && tiles + t > max_tiles / 2) break;
        tiles += t;
        ws_used += r->ws_need;
        in_use += r->footprint;
        queue.pop_front();
        batch.push_back(r);
      }
      if (batch.empty()) { // nothing admissible now: give the slot back, wait for a completion
        free_slots.push_front(s);
        if (!queue.empty() && in_use > 0) {
          unsigned long seen = completions;
          cv_sched.wait(lk, [&] { return stop || completions != seen; });
        }
        continue;
      }
    }
    int st = OVX_OK;
    if (cfg.plain) {
      st = launch_plain(*s, *batch[0]);
    } else if ((int)batch.size() <= kDirectMax) {
      // a small batch (light load): the requests' standalone launches on the
      // slot's stream cost less than replaying the slot graph (a persistent
      // launch over every CU plus one launch per four-step queue); same
      // device bodies, same results
      for (Req *r : batch)
        if ((st = launch_plain(*s, *r))) break;
    } else {
      st = admit_and_launch(*s, batch);
    }
    if (st == OVX_OK) {
      hipError_t e = hipEventRecord(s->done, s->stream);
      if (e != hipSuccess) st = hip_fail(e, "event record");
    }
    std::lock_guard<std::mutex> g(mu);
    s->reqs = batch;
    if (st != OVX_OK) {
      error = st;
      for (Req *r : batch) {
        r->status = st;
        r->t_done = now_s();
        in_use -= r->footprint;
      }
      s->reqs.clear();
      free_slots.push_back(s);
      cv_wait.notify_all();
      continue;
    }
    inflight.push_back(s);
    cv_done.notify_all();
  }
}

void Server::completer()
{
  (void)hipSetDevice(device);
  for (;;) {
    Slot *s = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_done.wait(lk, [&] { return sched_done || !inflight.empty(); });
      if (inflight.empty()) return; // the scheduler has exited and nothing is in flight
      s = inflight.front();
    }
    hipError_t e = hipEventSynchronize(s->done);
    double t = now_s();
    std::lock_guard<std::mutex> g(mu);
    inflight.pop_front();
    for (Req *r : s->reqs) {
      r->status = e == hipSuccess ? OVX_OK : OVX_ERR_DEVICE;
      r->t_done = t;
      in_use -= r->footprint;
    }
    s->reqs.clear();
    free_slots.push_back(s);
    ++completions;
    cv_sched.notify_all();
    cv_wait.notify_all();
  }
}

static int prepare_request(const ovx_request &r, Req &q)
{
  if (!r.x || !r.y) return fail(OVX_ERR_INVALID, "null data pointer");
  q.r = r;
  if (r.kind == OVX_REQ_CORR) {
    const CorrRef *cr;
    int st = corr_ref_get(r.ref, &cr);
    if (st) return st;
    if (r.bias != OVX_BIASED && r.bias != OVX_UNBIASED) return fail(OVX_ERR_INVALID, "bad bias");
    q.footprint = (size_t)(cr->n + cr->p) * 8;
    return OVX_OK;
  }
  if (r.kind != OVX_REQ_FASTCONV && r.kind != OVX_REQ_FFTM) return fail(OVX_ERR_INVALID, "bad request kind");
  if (r.n < 1 || r.rows < 1) return fail(OVX_ERR_INVALID, "bad request size");
  if (r.rows > 1 && (r.ldx < r.n || r.ldy < r.n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  q.footprint = (size_t)r.rows * (size_t)(r.ldx + r.ldy) * 8;
  int st;
  const LineFFT *f = shared_path(r.n, r.rows, &st);
  if (!f) return st;
  if (r.kind == OVX_REQ_FFTM && f->kind == LineFFT::FOURSTEP) q.ws_need = (size_t)r.rows * r.n * 8;
  q.footprint += q.ws_need;
  if (r.kind == OVX_REQ_FASTCONV) {
    const float2 *prep;
    long rn;
    if ((st = replica_prepared(r.ref, &prep, &rn, f))) return st;
    if (rn != r.n) return fail(OVX_ERR_INVALID, "replica length %ld != request length %ld", rn, r.n);
  }
  return OVX_OK;
}

} // namespace ovx

using namespace ovx;

struct ovx_server_s {
  Server srv;
};

extern "C" int ovx_server_create(ovx_server *s, const ovx_server_config *cfg)
{
  if (!s || !cfg) return fail(OVX_ERR_INVALID, "null pointer");
  *s = nullptr;
  std::unique_ptr<ovx_server_s> p(new (std::nothrow) ovx_server_s());
  if (!p) return fail(OVX_ERR_NOMEM, "host allocation");
  int st = p->srv.init(*cfg);
  if (st) return st;
  *s = p.release();
  return OVX_OK;
}

extern "C" int ovx_server_destroy(ovx_server s)
{
  if (!s) return fail(OVX_ERR_INVALID, "null server");
  delete s;
  return OVX_OK;
}

extern "C" int ovx_server_submit(ovx_server s, const ovx_request *r, unsigned long *ticket)
{
  if (!s || !r) return fail(OVX_ERR_INVALID, "null pointer");
  std::unique_ptr<Req> q(new (std::nothrow) Req());
  if (!q) return fail(OVX_ERR_NOMEM, "host allocation");
  int st = prepare_request(*r, *q);
  if (st) return st;
  Server &v = s->srv;
  std::lock_guard<std::mutex> g(v.mu);
  q->id = v.next_id++;
  q->t_submit = now_s();
  if (ticket) *ticket = q->id;
  v.queue.push_back(q.get());
  v.reqs[q->id] = std::move(q);
  v.cv_sched.notify_all();
  return OVX_OK;
}

extern "C" int ovx_server_wait(ovx_server s, unsigned long ticket, double timeout_s, double *latency_s)
{
  if (!s) return fail(OVX_ERR_INVALID, "null server");
  Server &v = s->srv;
  std::unique_lock<std::mutex> lk(v.mu);
  auto it = v.reqs.find(ticket);
  if (it == v.reqs.end()) return fail(OVX_ERR_INVALID, "unknown ticket %lu", ticket);
  Req *q = it->second.get();
  bool ok = v.cv_wait.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return q->status != -1; });
  if (!ok) return fail(OVX_ERR_TIMEOUT, "request %lu not completed within %.3f s", ticket, timeout_s);
  int st = q->status;
  if (latency_s) *latency_s = q->t_done - q->t_submit;
  v.reqs.erase(it);
  if (st) return fail(st, "request %lu failed with status %d", ticket, st);
  return OVX_OK;
}

extern "C" int ovx_server_run_trace(ovx_server s, const ovx_request *reqs, long count, double *latency_s,
                                    double *makespan_s)
{
  if (!s || (!reqs && count > 0)) return fail(OVX_ERR_INVALID, "null pointer");
  std::vector<unsigned long> tickets((size_t)count);
  double t0 = now_s();
  for (long i = 0; i < count; ++i) {
    double target = t0 + reqs[i].t_arrival;
    while (now_s() < target) {
      double d = target - now_s();
      if (d > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(d - 100e-6));
    }
    int st = ovx_server_submit(s, &reqs[i], &tickets[(size_t)i]);
    if (st) return st;
  }
  int first_err = OVX_OK;
  double last = t0;
  for (long i = 0; i < count; ++i) {
    double lat = 0;
    int st = ovx_server_wait(s, tickets[(size_t)i], 600.0, &lat);
    if (st && !first_err) first_err = st;
    if (latency_s) latency_s[i] = st ? -1.0 : lat;
    last = std::max(last, t0 + reqs[i].t_arrival + lat);
  }
  if (makespan_s) *makespan_s = last - t0;
  return first_err;
}

// ---------------------------------------------------------------- correlation references
extern "C" int ovx_corr_reference_create(int support, long m, long n, const void *ref, int *id)
{
  if (!ref || !id) return fail(OVX_ERR_INVALID, "null pointer");
  if (support < OVX_SUPPORT_FULL || support > OVX_SUPPORT_MIN) return fail(OVX_ERR_INVALID, "bad support");
  if (m < 1 || n < 1 || m > n) return fail(OVX_ERR_INVALID, "bad sizes m=%ld n=%ld", m, n);
  std::unique_ptr<CorrRef> c(new (std::nothrow) CorrRef());
  if (!c) return fail(OVX_ERR_NOMEM, "host allocation");
  c->support = support;
  c->m = m;
  c->n = n;
  c->p = ovx_corr_output_size(support, m, n);
  long shift = support == OVX_SUPPORT_FULL ? m - 1 : support == OVX_SUPPORT_SAME ? m / 2 : 0;
  c->out_off = m - 1 - shift;
  long L = 1024;
  while (L < 4 * (m - 1) && L < kMaxRowN) L *= 2;
  if (m - 1 >= L) return fail(OVX_ERR_UNSUPPORTED, "reference length %ld too long", m);
  c->L = (int)L;
  int st = c->f.init(L, hier_supported(L) ? 1 : 0);
  if (st) return st;
  if (c->f.kind != LineFFT::LDS || !c->f.tw.p) return fail(OVX_ERR_UNSUPPORTED, "no segment path for L=%ld", L);
  DevBuf href, hpad;
  if ((st = href.alloc((size_t)m * 8)) || (st = hpad.alloc((size_t)L * 8)) || (st = c->H.alloc((size_t)L * 8)))
    return st;
  hipError_t e = hipMemcpy(href.p, ref, (size_t)m * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_ols_prep((const float2 *)href.p, m, 1, (float2 *)hpad.p, L, nullptr);
  if (e != hipSuccess) return hip_fail(e, "reference upload");
  if ((st = c->f.transform(-1, (const float2 *)hpad.p, L, (float2 *)c->H.p, L, 1, 1.0f, nullptr))) return st;
  e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return hip_fail(e, "reference spectrum");
  std::lock_guard<std::mutex> g(g_cref_mu);
  *id = g_next_cref++;
  g_crefs[*id] = std::move(c);
  return OVX_OK;
}

extern "C" int ovx_corr_reference_destroy(int id)
{
  std::lock_guard<std::mutex> g(g_cref_mu);
  if (!g_crefs.erase(id)) return fail(OVX_ERR_INVALID, "unknown correlation reference %d", id);
  return OVX_OK;
}
