// ovx_batch_host.cpp -- batches of uneven fastconv jobs (C-ABI part):
// tile-queue construction, device tables, the persistent launches over the
// LDS-resident tiles and the four-step pass launches captured once into a
// hipGraph and replayed, and the plain one-launch-per-job path it is
// compared with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "ovx_batch.hpp"
#include "ovx_batch_host.hpp"
#include "ovx_plan.hpp"

namespace ovx {

// ---------------------------------------------------------------- shared paths
// Kernel paths (twiddle tables) per (transform length, variant), shared by
// batches, replicas and the server.
static std::mutex g_path_mu;
static std::map<std::pair<long, int>, std::unique_ptr<LineFFT>> g_paths;

// Variant the batch tiles of an (n, rows) job run: the autotuner's batch
// entry for (n, rows) when it has one, else the default -- for the 3 * 2^k
// lengths the 12-values-per-thread register-Stockham configuration (variant
// 1), which fits the 1024-thread persistent kernel's register budget, for
// the others the standalone default.
int batch_variant(long n, long rows)
{
  int v = tuned_batch_variant(n, rows);
  if (v >= 0) return v;
  return (n == 3072 || n == 6144 || n == 12288) ? 1 : LineFFT::default_variant(n);
}

const LineFFT *shared_path(long n, int variant, int *status)
{
  std::lock_guard<std::mutex> g(g_path_mu);
  auto it = g_paths.find({ n, variant });
  if (it != g_paths.end()) {
    *status = OVX_OK;
    return it->second.get();
  }
  std::unique_ptr<LineFFT> f(new (std::nothrow) LineFFT());
  if (!f) {
    *status = fail(OVX_ERR_NOMEM, "host allocation");
    return nullptr;
  }
  *status = f->init(n, variant);
  if (*status) return nullptr;
  const LineFFT *p = f.get();
  g_paths[{ n, variant }] = std::move(f);
  return p;
}

const LineFFT *shared_path(long n, long rows, int *status) { return shared_path(n, batch_variant(n, rows), status); }

int batch_path_of(const LineFFT &f)
{
  if (f.kind == LineFFT::HIER) return BP_HIER;
  if (f.kind == LineFFT::LDS) return BP_ROW;
  if (f.kind == LineFFT::FOURSTEP) return BP_FS;
  return 0;
}

int device_cu_count()
{
  static int cus = 0;
  if (!cus) {
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
    cus = p.multiProcessorCount;
  }
  return cus;
}

// Fill `job` for a fastconv of n x rows against a prepared spectrum.
int make_batch_job(const LineFFT &f, long rows, const float2 *x, long ldx, float2 *y, long ldy, const float2 *wprep,
                   BatchJob *job)
{
  BatchJob j{};
  j.path = batch_path_of(f);
  j.n = (int)f.n;
  if (!batch_supported(j.path, j.n))
    return fail(OVX_ERR_UNSUPPORTED, "length %ld has no batch tile (path %d)", f.n, j.path);
  if (!f.tables_ready()) return fail(OVX_ERR_STATE, "length %ld: twiddle tables missing", f.n);
  if (rows > 0x7fffffff) return fail(OVX_ERR_UNSUPPORTED, "too many rows in one job");
  j.rows = rows;
  j.x = x;
  j.ldx = ldx;
  j.y = y;
  j.ldy = ldy;
  j.w = wprep;
  j.scale = 1.0f;
  if (j.path == BP_HIER) {
    j.tw0 = (const float2 *)f.tw_outer.p;
    j.tw1 = (const float2 *)f.tw_wave.p;
  } else if (j.path == BP_ROW) {
    // the tile body is Engine<n, batch_row_p(n)>: the path must be exactly
    // that configuration, or its twiddle table would not match
    if (f.rc.p != batch_row_p(j.n))
      return fail(OVX_ERR_STATE, "length %ld: path has %d values per thread, batch tile needs %d", f.n, f.rc.p,
                  batch_row_p(j.n));
    j.tw0 = (const float2 *)f.tw.p;
  } else {
    j.tw0 = (const float2 *)f.tw1.p;
    j.tw1 = (const float2 *)f.tw2.p;
    j.tlo = (const float2 *)f.tlo.p;
    j.thi = (const float2 *)f.thi.p;
    j.sbits = f.sbits;
    j.tfull = (const float2 *)f.tfull.p;
    j.n1 = (int)f.n1;
    j.n2 = (int)f.n2;
  }
  *job = j;
  return OVX_OK;
}

// Tiles of job `ji`: LDS-resident rows go to the queue of their persistent
// launch (batch_queue), the three four-step passes to their (pass, line
// length) queues, one line group per tile.
void make_batch_tiles(const BatchJob &j, int ji, std::vector<BatchTile> q[kBatchQueues])
{
  if (j.path == BP_HIER || j.path == BP_ROW) {
    int rpt = batch_rows_per_tile(j.path, j.n);
    int op = j.path == BP_HIER ? BT_HIER_FASTCONV : BT_ROW_FASTCONV;
    std::vector<BatchTile> &dst = q[batch_queue(j.path, j.n)];
    for (long r = 0; r < j.rows; r += rpt) dst.push_back(BatchTile{ ji, op, (int)r, 0 });
    return;
  }
  const int gcol = j.n2 / batch_fs_lines_per_group(j.n1), grow = j.n1 / batch_fs_lines_per_group(j.n2);
  std::vector<BatchTile> &a = q[fs_queue(0, j.n1)], &b = q[fs_queue(1, j.n2)], &c = q[fs_queue(2, j.n1)];
  for (long r = 0; r < j.rows; ++r) {
    for (int g = 0; g < gcol; ++g) a.push_back(BatchTile{ ji, BT_FS_A, (int)r, g });
    for (int g = 0; g < grow; ++g) b.push_back(BatchTile{ ji, BT_FS_B, (int)r, g });
    for (int g = 0; g < gcol; ++g) c.push_back(BatchTile{ ji, BT_FS_C, (int)r, g });
  }
}

std::vector<TileQueue> chunked_fs_queues(const std::vector<BatchJob> &jobs, const std::vector<int> &order,
                                         size_t chunk_bytes)
{
  std::vector<TileQueue> out;
  std::vector<BatchTile> cq[kBatchQueues];
  size_t used = 0;
  auto flush = [&] {
    for (int k = 1; k < kFsQueueEnd; ++k)
      if (!cq[k].empty()) {
        out.push_back(TileQueue{ fs_queue_line(k), fs_queue_pass(k), std::move(cq[k]) });
        cq[k].clear();
      }
    used = 0;
  };
  for (int ji : order) {
    const BatchJob &j = jobs[(size_t)ji];
    if (j.path != BP_FS) continue;
    const size_t row_bytes = (size_t)j.n * 8;
    const int gcol = j.n2 / batch_fs_lines_per_group(j.n1), grow = j.n1 / batch_fs_lines_per_group(j.n2);
    const int qa = fs_queue(0, j.n1), qb = fs_queue(1, j.n2), qc = fs_queue(2, j.n1);
    for (long r = 0; r < j.rows; ++r) {
      if (used > 0 && used + row_bytes > chunk_bytes) flush();
      for (int g = 0; g < gcol; ++g) cq[qa].push_back(BatchTile{ ji, BT_FS_A, (int)r, g });
      for (int g = 0; g < grow; ++g) cq[qb].push_back(BatchTile{ ji, BT_FS_B, (int)r, g });
      for (int g = 0; g < gcol; ++g) cq[qc].push_back(BatchTile{ ji, BT_FS_C, (int)r, g });
      used += row_bytes;
    }
  }
  flush();
  return out;
}

void make_batch_fft_tiles(const BatchJob &j, int ji, std::vector<BatchTile> q[kBatchQueues])
{
  const int gcol = j.n2 / batch_fs_lines_per_group(j.n1), grow = j.n1 / batch_fs_lines_per_group(j.n2);
  std::vector<BatchTile> &a = q[fs_queue(0, j.n1)], &b = q[fs_queue(1, j.n2)];
  for (long r = 0; r < j.rows; ++r) {
    for (int g = 0; g < gcol; ++g) a.push_back(BatchTile{ ji, BT_FS_FFT_A, (int)r, g });
    for (int g = 0; g < grow; ++g) b.push_back(BatchTile{ ji, BT_FS_FFT_B, (int)r, g });
  }
}

hipError_t capture_batch_launches(const BatchJob *jobs, BatchTile *const tiles[kBatchQueues], int *meta,
                                  const int *counts, bool all, hipStream_t cap)
{
  hipError_t e = hipSuccess;
  if (all || counts[0] > 0) e = launch_batch(jobs, tiles[0], meta, device_cu_count(), cap);
  if (e == hipSuccess && (all || counts[kMidQueue] > 0))
    e = launch_batch_mid(jobs, tiles[kMidQueue], meta + kMidQueue, kBatchMidBlocksPerCu * device_cu_count(), cap);
  for (int k = 1; k < kFsQueueEnd && e == hipSuccess; ++k) {
    if (!all && counts[k] == 0) continue;
    const int line = fs_queue_line(k);
    // a batch with known counts launches one workgroup per tile (as the
    // standalone passes do); the server's graph, whose counts change per
    // batch, strides a resident grid over the queue
    const int grid = all ? batch_fs_blocks_per_cu(line) * device_cu_count() : counts[k];
    e = launch_batch_fs(line, fs_queue_pass(k), jobs, tiles[k], meta + k, grid, cap);
  }
  return e;
}

GraphBatch::~GraphBatch() { release_graph(); }

void GraphBatch::release_graph()
{
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
}

// Upload jobs and the tile queues; capture the launches once, in queue order.
int GraphBatch::build(const std::vector<BatchJob> &jobs, const std::vector<TileQueue> &queues, hipStream_t s)
{
  release_graph();
  const int nq = (int)queues.size();
  size_t nt = 0;
  for (const auto &q : queues) nt += q.tiles.size();
  int st;
  if ((st = d_jobs.alloc(std::max<size_t>(jobs.size(), 1) * sizeof(BatchJob))) ||
      (st = d_tiles.alloc(std::max<size_t>(nt, 1) * sizeof(BatchTile))) || (st = d_meta.alloc(std::max(nq, 1) * sizeof(int))))
    return st;
  std::vector<BatchTile> all;
  all.reserve(nt);
  std::vector<int> counts((size_t)std::max(nq, 1), 0); // tile counts
  std::vector<size_t> offset(nq);
  nlaunch = 0;
  for (int k = 0; k < nq; ++k) {
    offset[k] = all.size();
    counts[k] = (int)queues[k].tiles.size();
    all.insert(all.end(), queues[k].tiles.begin(), queues[k].tiles.end());
    if (counts[k]) ++nlaunch;
  }
  hipError_t e = hipSuccess;
  if (!jobs.empty()) e = hipMemcpyAsync(d_jobs.p, jobs.data(), jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !all.empty())
    e = hipMemcpyAsync(d_tiles.p, all.data(), all.size() * sizeof(BatchTile), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_meta.p, counts.data(), counts.size() * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "batch table upload");
  ntiles = nt;
  if (nlaunch == 0) return OVX_OK;

  hipStream_t cap;
  e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "capture stream");
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const BatchJob *dj = (const BatchJob *)d_jobs.p;
    int *meta = (int *)d_meta.p;
    for (int k = 0; k < nq && e == hipSuccess; ++k) {
      if (!counts[k]) continue;
      const BatchTile *t = (const BatchTile *)d_tiles.p + offset[k];
      if (queues[k].line == 0) e = launch_batch(dj, t, meta + k, device_cu_count(), cap);
      else if (queues[k].line == kMidLine)
        e = launch_batch_mid(dj, t, meta + k, kBatchMidBlocksPerCu * device_cu_count(), cap);
      else e = launch_batch_fs(queues[k].line, queues[k].pass, dj, t, meta + k, counts[k], cap); // one workgroup per tile
    }
    hipError_t e2 = hipStreamEndCapture(cap, &graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cap);
  if (e != hipSuccess) {
    release_graph();
    return hip_fail(e, "batch graph capture");
  }
  return OVX_OK;
}

int GraphBatch::launch(hipStream_t s) const
{
  if (nlaunch == 0) return OVX_OK;
  hipError_t e = hipGraphLaunch(exec, s);
  if (e != hipSuccess) return hip_fail(e, "batch graph launch");
  return OVX_OK;
}

} // namespace ovx

using namespace ovx;

// ---------------------------------------------------------------- C-ABI
struct ovx_batch_s {
  struct HostJob {
    long n, rows;
    const void *x;
    long ldx;
    int replica;
    void *y;
    long ldy;
  };
  std::vector<HostJob> jobs;
  std::vector<BatchJob> djobs;
  std::vector<const LineFFT *> paths;
  std::vector<const float2 *> preps;
  GraphBatch gb;
  bool committed = false;
};

extern "C" int ovx_batch_create(ovx_batch *b)
{
  if (!b) return fail(OVX_ERR_INVALID, "null batch pointer");
  *b = new (std::nothrow) ovx_batch_s();
  if (!*b) return fail(OVX_ERR_NOMEM, "host allocation");
  return OVX_OK;
}

extern "C" int ovx_batch_destroy(ovx_batch b)
{
  if (!b) return fail(OVX_ERR_INVALID, "null batch");
  delete b;
  return OVX_OK;
}

extern "C" int ovx_batch_add_fastconv(ovx_batch b, long n, long rows, const void *x, long ldx, int replica, void *y,
                                      long ldy)
{
  if (!b) return fail(OVX_ERR_INVALID, "null batch");
  if (b->committed) return fail(OVX_ERR_STATE, "batch already committed");
  if (!x || !y) return fail(OVX_ERR_INVALID, "null data pointer");
  if (n < 1 || rows < 0) return fail(OVX_ERR_INVALID, "bad job size");
  if (rows > 1 && (ldx < n || ldy < n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  b->jobs.push_back(ovx_batch_s::HostJob{ n, rows, x, ldx, replica, y, ldy });
  return OVX_OK;
}

extern "C" int ovx_batch_commit(ovx_batch b, void *stream)
{
  if (!b) return fail(OVX_ERR_INVALID, "null batch");
  if (b->committed) return fail(OVX_ERR_STATE, "batch already committed");
  hipStream_t s = (hipStream_t)stream;
  b->djobs.clear();
  b->paths.clear();
  b->preps.clear();
  // tile bodies for the (n, rows) keys the autotuner has not seen yet
  if (!getenv("OVX_NO_BATCH_TUNE")) {
    for (size_t i = 0; i < b->jobs.size(); ++i) {
      const auto &h = b->jobs[i];
      if (batch_candidates(h.n).size() > 1 && tuned_batch_variant(h.n, h.rows) < 0) {
        int st = autotune_batch(h.n, h.rows, nullptr, nullptr);
        if (st) return st;
      }
    }
  }
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    const auto &h = b->jobs[i];
    int st;
    const LineFFT *f = shared_path(h.n, h.rows, &st);
    if (!f) return st;
    const float2 *prep = nullptr;
    long rn = 0;
    if ((st = replica_prepared(h.replica, &prep, &rn, f))) return st;
    if (rn != h.n) return fail(OVX_ERR_INVALID, "replica %d has length %ld, job %zu has %ld", h.replica, rn, i, h.n);
    BatchJob j;
    if ((st = make_batch_job(*f, h.rows, (const float2 *)h.x, h.ldx, (float2 *)h.y, h.ldy, prep, &j))) return st;
    b->djobs.push_back(j);
    b->paths.push_back(f);
    b->preps.push_back(prep);
  }
  // largest jobs first inside each queue
  std::vector<int> order(b->djobs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->djobs[a].n > b->djobs[c].n; });
  // Launch order: the two persistent launches over the LDS-resident tiles
  // (1024-thread rows, then the shorter ones), then the four-step jobs' rows
  // in Infinity-Cache-sized chunks, one launch per non-empty (pass, line
  // length) queue of a chunk, in pass order.
  std::vector<BatchTile> q[kBatchQueues];
  for (int i : order)
    if (b->djobs[i].path != BP_FS) make_batch_tiles(b->djobs[i], i, q);
  std::vector<TileQueue> queues;
  if (!q[0].empty()) queues.push_back(TileQueue{ 0, 0, std::move(q[0]) });
  if (!q[kMidQueue].empty()) queues.push_back(TileQueue{ kMidLine, 0, std::move(q[kMidQueue]) });
  for (auto &fq : chunked_fs_queues(b->djobs, order, kFsChunkBytes)) queues.push_back(std::move(fq));
  int st = b->gb.build(b->djobs, queues, s);
  if (st) return st;
  b->committed = true;
  return OVX_OK;
}

extern "C" int ovx_batch_exec(ovx_batch b, void *stream)
{
  if (!b) return fail(OVX_ERR_INVALID, "null batch");
  if (!b->committed) return fail(OVX_ERR_STATE, "batch not committed");
  return b->gb.launch((hipStream_t)stream);
}

extern "C" int ovx_batch_exec_plain(ovx_batch b, void *stream)
{
  if (!b) return fail(OVX_ERR_INVALID, "null batch");
  if (!b->committed) return fail(OVX_ERR_STATE, "batch not committed");
  hipStream_t s = (hipStream_t)stream;
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    const auto &h = b->jobs[i];
    int st = b->paths[i]->fastconv_prepared((const float2 *)h.x, h.ldx, b->preps[i], (float2 *)h.y, h.ldy, h.rows, s);
    if (st) return st;
  }
  return OVX_OK;
}

extern "C" long ovx_batch_tiles(ovx_batch b)
{
  if (!b) return -1;
  return (long)b->gb.ntiles;
}

extern "C" int ovx_batch_launches(ovx_batch b)
{
  if (!b) return -1;
  return b->gb.nlaunch;
}

extern "C" int ovx_batch_job_variant(ovx_batch b, long job)
{
  if (!b || !b->committed || job < 0 || job >= (long)b->paths.size()) return -1;
  return b->paths[(size_t)job]->variant;
}
