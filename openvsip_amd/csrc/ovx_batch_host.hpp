// ovx_batch_host.hpp -- host-side pieces shared by batches and the server.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "ovx_batch.hpp"
#include "ovx_plan.hpp"

namespace ovx {

// Kernel path (and its twiddles) of length n in `variant`, created on first
// use and shared.  The (n, rows) form takes batch_variant(n, rows).
const LineFFT *shared_path(long n, int variant, int *status);
const LineFFT *shared_path(long n, long rows, int *status);
// Variant of the tile bodies an (n, rows) job runs in a batch or the server:
// the autotuner's batch entry, else the default.
int batch_variant(long n, long rows);
// Batch tile autotuner (ovx_tune.cpp): the variants whose device body a
// batch tile can run for length n, the cached choice for (n, rows) or -1,
// and the tuning itself (records the choice).
std::vector<int> batch_candidates(long n);
int tuned_batch_variant(long n, long rows);
int autotune_batch(long n, long rows, int *best_variant, double *best_us);
int batch_path_of(const LineFFT &f);
int device_cu_count();
int make_batch_job(const LineFFT &f, long rows, const float2 *x, long ldx, float2 *y, long ldy, const float2 *wprep,
                   BatchJob *job);
// Tiles of job `ji` appended to the batch's queues (kBatchQueues of them).
void make_batch_tiles(const BatchJob &j, int ji, std::vector<BatchTile> q[kBatchQueues]);
// Four-step forward Fftm tiles (x -> workspace -> y) of job `ji`.
void make_batch_fft_tiles(const BatchJob &j, int ji, std::vector<BatchTile> q[kBatchQueues]);
// Capture the launches of a batch into `graph` on stream `cap`: the
// persistent launches over queue 0 and kMidQueue, then one launch per
// four-step queue in pass order; `all` launches every queue (tile counts read from device memory at
// run time), otherwise only the non-empty ones in `counts`.
hipError_t capture_batch_launches(const BatchJob *jobs, BatchTile *const tiles[kBatchQueues], int *meta,
                                  const int *counts, bool all, hipStream_t cap);
// Prepared spectrum of a resident replica in the layout of path `f`.
int replica_prepared(int id, const float2 **prep, long *n, const LineFFT *f);

// Server slot tables: tile counts [0 .. kBatchQueues).
constexpr int kMetaInts = 16;

// One launch of a batch: a persistent kernel over LDS-resident tiles
// (line == 0: the 1024-thread one, line == kMidLine: the 512-thread one) or
// four-step pass `pass` over lines of length `line`.
constexpr int kMidLine = -1;
struct TileQueue {
  int line = 0;
  int pass = 0;
  std::vector<BatchTile> tiles;
};

// Four-step passes of a batch in chunks of about kFsChunkBytes of rows
// (rows of one job may fall into different chunks), each chunk's three passes
// launched back to back: a chunk's matrix then stays in the 256 MB Infinity
// Cache from one pass to the next.  Standalone 65536-point fastconvs run at
// 0.56 us per row with 96-128 MB matrices and at 0.70 us per row from 512 MB
// up, which is what one pass over all of a batch's four-step rows gives.
constexpr size_t kFsChunkBytes = (size_t)128 << 20;
// Four-step queues (in launch order) of jobs order[0], order[1], ... that
// take the four-step path.
std::vector<TileQueue> chunked_fs_queues(const std::vector<BatchJob> &jobs, const std::vector<int> &order,
                                         size_t chunk_bytes);

// Device job/tile tables plus the batch's launches (queues in launch order)
// captured in one hipGraph.
struct GraphBatch {
  DevBuf d_jobs, d_tiles, d_meta;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nlaunch = 0;
  size_t ntiles = 0;
  ~GraphBatch();
  void release_graph();
  int build(const std::vector<BatchJob> &jobs, const std::vector<TileQueue> &queues, hipStream_t s);
  int launch(hipStream_t s) const;
};

} // namespace ovx
