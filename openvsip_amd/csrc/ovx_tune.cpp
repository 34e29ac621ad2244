// ovx_tune.cpp -- autotuner for the kernel variant of a (transform length,
// rows) pair, with a persistent cache.
//
// Every compiled variant of a length (hierarchical, register-Stockham with
// different values per thread) gives the same result to rounding; which is
// fastest depends on the length and on how many rows share the GPU.  The
// tuner times each variant's fused fastconv against a spectrum prepared in
// that variant's layout (the resident-replica path) on scratch data of the
// requested shape -- after ~50 ms of untimed launches, because the device
// clocks take about that long to settle under this load, then five
// interleaved rounds, best round per variant -- and records the winner under
// (n, rows rounded up to a power of two).
// Plans created with the default variant flag consult the cache.  Batches
// and the server have their own entries: which tile body (hierarchical, or
// the register-Stockham configuration the persistent kernels hold) runs the
// rows of an (n, rows) job inside a batch, timed as a one-job batch replayed
// from its graph; a batch tunes the keys it lacks when it is committed.  The
// cache is a text file, one "n rows_bucket variant microseconds" line per
// plan entry and one "b n rows_bucket variant microseconds" line per batch
// entry, at $OVX_TUNE_CACHE (default ~/.cache/ovxhip/tune.txt); it is the
// only state the library persists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <sys/stat.h>
#include <utility>
#include <vector>

#include "ovx_batch.hpp"
#include "ovx_batch_host.hpp"
#include "ovx_plan.hpp"

namespace ovx {

static std::mutex g_tune_mu;
// (n, rows bucket) -> (variant, microseconds): standalone fastconv plans
static std::map<std::pair<long, long>, std::pair<int, double>> g_tune;
// the same for batch and server tiles ("b" lines in the cache file)
static std::map<std::pair<long, long>, std::pair<int, double>> g_tune_batch;
static bool g_tune_loaded = false;

static long rows_bucket(long rows)
{
  long b = 1;
  while (b < rows) b <<= 1;
  return b;
}

static std::string cache_path()
{
  const char *p = getenv("OVX_TUNE_CACHE");
  if (p && *p) return p;
  const char *home = getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/ovxhip/tune.txt";
}

static void load_locked()
{
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  FILE *f = fopen(cache_path().c_str(), "r");
  if (!f) return;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    long n, rb;
    int v;
    double us;
    if (sscanf(line, "b %ld %ld %d %lf", &n, &rb, &v, &us) == 4) g_tune_batch[{ n, rb }] = { v, us };
    else if (sscanf(line, "%ld %ld %d %lf", &n, &rb, &v, &us) == 4) g_tune[{ n, rb }] = { v, us };
  }
  fclose(f);
}

static void save_locked()
{
  std::string path = cache_path();
  std::string dir = path.substr(0, path.find_last_of('/'));
  // create the directory chain (ignore failures: the cache is optional)
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
  std::string tmp = path + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return;
  for (auto &e : g_tune) fprintf(f, "%ld %ld %d %.3f\n", e.first.first, e.first.second, e.second.first, e.second.second);
  for (auto &e : g_tune_batch)
    fprintf(f, "b %ld %ld %d %.3f\n", e.first.first, e.first.second, e.second.first, e.second.second);
  fclose(f);
  rename(tmp.c_str(), path.c_str());
}

// Variant cached for (n, rows), or -1.
int tuned_variant(long n, long rows)
{
  std::lock_guard<std::mutex> g(g_tune_mu);
  load_locked();
  auto it = g_tune.find({ n, rows_bucket(rows) });
  return it == g_tune.end() ? -1 : it->second.first;
}

// One compiled variant with its own spectrum in the layout its kernel reads.
struct TuneCand {
  LineFFT f;
  DevBuf wp;
  bool ok = false;
};

static int run_prepared(const TuneCand &c, long n, long rows, const float2 *x, float2 *y, hipStream_t s)
{
  return c.f.fastconv_prepared(x, n, (const float2 *)c.wp.p, y, n, rows, s);
}

static int time_variant(const TuneCand &c, long n, long rows, const float2 *x, float2 *y, hipStream_t s, double *us)
{
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(OVX_ERR_DEVICE, "event create");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(OVX_ERR_DEVICE, "event create");
  }
  const int iters = 50;
  int st = OVX_OK;
  hipError_t e = hipEventRecord(e0, s);
  for (int i = 0; i < iters && !st; ++i) st = run_prepared(c, n, rows, x, y, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  if (e != hipSuccess) return hip_fail(e, "tuning timing");
  *us = ms * 1e3 / iters;
  return OVX_OK;
}

} // namespace ovx

using namespace ovx;

extern "C" int ovx_autotune_fastconv(long n, long rows, int *best_variant, double *best_us)
{
  if (n < 1 || rows < 1) return fail(OVX_ERR_INVALID, "bad shape");
  int nv = LineFFT::variants(n);
  if (nv < 1) return fail(OVX_ERR_UNSUPPORTED, "length %ld not supported", n);
  DevBuf x, w, y;
  size_t bytes = (size_t)rows * n * 8;
  int st;
  if ((st = x.alloc(bytes)) || (st = y.alloc(bytes)) || (st = w.alloc((size_t)n * 8))) return st;
  hipError_t e = hipMemset(x.p, 0, bytes);
  if (e == hipSuccess) e = hipMemset(w.p, 0, (size_t)n * 8);
  if (e == hipSuccess) e = hipDeviceSynchronize(); // the tuning stream does not wait for the null stream
  if (e != hipSuccess) return hip_fail(e, "tuning buffers");
  hipStream_t s;
  if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "tuning stream");
  // every variant's plan and prepared spectrum, built once
  std::vector<std::unique_ptr<TuneCand>> cand;
  for (int v = 0; v < nv && !st; ++v) {
    cand.emplace_back(new (std::nothrow) TuneCand());
    TuneCand *c = cand.back().get();
    if (!c) {
      st = fail(OVX_ERR_NOMEM, "host allocation");
      break;
    }
    int r = c->f.init(n, v);
    if (r == OVX_ERR_UNSUPPORTED) continue; // shape outside this variant's limits
    if (r || (r = c->wp.alloc((size_t)n * 8)) ||
        (r = c->f.prepare_spectrum((const float2 *)w.p, (float2 *)c->wp.p, 1.0f / (float)n, s))) {
      st = r;
      break;
    }
    c->ok = true;
  }
  // untimed launches until the clocks have settled (~50 ms, bounded count)
  if (!st) {
    auto t0 = std::chrono::steady_clock::now();
    for (int loop = 0; loop < 2000 && !st; ++loop) {
      for (auto &c : cand)
        if (c && c->ok && !st) st = run_prepared(*c, n, rows, (const float2 *)x.p, (float2 *)y.p, s);
      if ((loop & 7) == 7) {
        if ((e = hipStreamSynchronize(s)) != hipSuccess) st = hip_fail(e, "tuning warm-up");
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
      }
    }
  }
  // five interleaved rounds, the fastest round of each variant counts
  // (robust to a busy device during one round)
  std::vector<double> t_min(nv, 1e30);
  for (int round = 0; round < 5 && !st; ++round) {
    for (int v = 0; v < nv && !st; ++v) {
      if (!cand[v] || !cand[v]->ok) continue;
      double us;
      st = time_variant(*cand[v], n, rows, (const float2 *)x.p, (float2 *)y.p, s, &us);
      if (!st) t_min[v] = std::min(t_min[v], us);
    }
  }
  int best = 0;
  double best_t = 1e30;
  for (int v = 0; v < nv; ++v)
    if (t_min[v] < best_t) {
      best_t = t_min[v];
      best = v;
    }
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  if (st) return st;
  if (best_t >= 1e30) return fail(OVX_ERR_UNSUPPORTED, "no variant ran for n=%ld rows=%ld", n, rows);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    load_locked();
    g_tune[{ n, rows_bucket(rows) }] = { best, best_t };
    save_locked();
  }
  if (best_variant) *best_variant = best;
  if (best_us) *best_us = best_t;
  return OVX_OK;
}

extern "C" int ovx_tuned_variant(long n, long rows) { return tuned_variant(n, rows); }

// ---------------------------------------------------------------- batch tiles
// The tile bodies a batch (and the server) can run for length n: the
// hierarchical body and the register-Stockham body with batch_row_p(n)
// values per thread, each the device body of one standalone variant.
namespace ovx {

std::vector<int> batch_candidates(long n)
{
  std::vector<int> out;
  if (n < 1 || n > kMaxRowN) return out;
  const bool hier = hier_supported(n);
  if (hier && batch_supported(BP_HIER, (int)n)) out.push_back(0);
  if (batch_supported(BP_ROW, (int)n)) {
    for (int v = 0; v < row_variants((int)n); ++v) {
      RowConfig rc;
      if (row_config((int)n, v, &rc) && rc.p == batch_row_p((int)n)) out.push_back(hier ? v + 1 : v);
    }
  }
  return out;
}

int tuned_batch_variant(long n, long rows)
{
  std::lock_guard<std::mutex> g(g_tune_mu);
  load_locked();
  auto it = g_tune_batch.find({ n, rows_bucket(rows) });
  return it == g_tune_batch.end() ? -1 : it->second.first;
}

// Time one batch of a single job (n x rows, zero data, a spectrum prepared
// for the candidate's path) per candidate tile body, replayed from its
// captured graph after ~20 ms of untimed replays; the fastest of five
// rounds of 20 replays counts.
int autotune_batch(long n, long rows, int *best_variant, double *best_us)
{
  std::vector<int> cands = batch_candidates(n);
  if (cands.empty()) return fail(OVX_ERR_UNSUPPORTED, "length %ld has no batch tile to tune", n);
  DevBuf x, y, w, wp;
  const size_t bytes = (size_t)rows * n * 8;
  int st;
  if ((st = x.alloc(bytes)) || (st = y.alloc(bytes)) || (st = w.alloc((size_t)n * 8)) || (st = wp.alloc((size_t)n * 8)))
    return st;
  hipError_t e = hipMemset(x.p, 0, bytes);
  if (e == hipSuccess) e = hipMemset(w.p, 0, (size_t)n * 8);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "batch tuning buffers");
  hipStream_t s;
  if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "tuning stream");
  int best = -1;
  double best_t = 1e30;
  for (int v : cands) {
    LineFFT f;
    if ((st = f.init(n, v))) break;
    if ((st = f.prepare_spectrum((const float2 *)w.p, (float2 *)wp.p, 1.0f / (float)n, s))) break;
    BatchJob j;
    if ((st = make_batch_job(f, rows, (const float2 *)x.p, n, (float2 *)y.p, n, (const float2 *)wp.p, &j))) break;
    std::vector<BatchJob> jobs(1, j);
    std::vector<BatchTile> q[kBatchQueues];
    make_batch_tiles(j, 0, q);
    std::vector<TileQueue> queues;
    if (!q[0].empty()) queues.push_back(TileQueue{ 0, 0, std::move(q[0]) });
    if (!q[kMidQueue].empty()) queues.push_back(TileQueue{ kMidLine, 0, std::move(q[kMidQueue]) });
    GraphBatch gb;
    if ((st = gb.build(jobs, queues, s))) break;
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < 400 && !st; ++k) {
      st = gb.launch(s);
      if ((k & 7) == 7) {
        if ((e = hipStreamSynchronize(s)) != hipSuccess) st = hip_fail(e, "batch tuning");
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
      }
    }
    double t_min = 1e30;
    for (int round = 0; round < 5 && !st; ++round) {
      hipEvent_t e0, e1;
      if (hipEventCreate(&e0) != hipSuccess) { st = fail(OVX_ERR_DEVICE, "event create"); break; }
      if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); st = fail(OVX_ERR_DEVICE, "event create"); break; }
      e = hipEventRecord(e0, s);
      for (int k = 0; k < 20 && !st; ++k) st = gb.launch(s);
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (!st && e != hipSuccess) st = hip_fail(e, "batch tuning timing");
      if (!st) t_min = std::min(t_min, ms * 1e3 / 20);
    }
    if (st) break;
    if (t_min < best_t) {
      best_t = t_min;
      best = v;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  if (st) return st;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    load_locked();
    g_tune_batch[{ n, rows_bucket(rows) }] = { best, best_t };
    save_locked();
  }
  if (best_variant) *best_variant = best;
  if (best_us) *best_us = best_t;
  return OVX_OK;
}

} // namespace ovx

extern "C" int ovx_autotune_batch(long n, long rows, int *best_variant, double *best_us)
{
  if (n < 1 || rows < 1) return fail(OVX_ERR_INVALID, "bad shape");
  return autotune_batch(n, rows, best_variant, best_us);
}

extern "C" int ovx_tuned_batch_variant(long n, long rows) { return tuned_batch_variant(n, rows); }
