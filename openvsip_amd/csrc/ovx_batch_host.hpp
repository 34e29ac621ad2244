// ovx_batch_host.hpp -- host-side pieces shared by batches and the server.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "ovx_batch.hpp"
#include "ovx_plan.hpp"

namespace ovx {

// Default kernel path (and its twiddles) of length n, created on first use.
const LineFFT *shared_path(long n, int *status);
int batch_path_of(const LineFFT &f);
int device_cu_count();
int make_batch_job(const LineFFT &f, long rows, const float2 *x, long ldx, float2 *y, long ldy, const float2 *wprep,
                   BatchJob *job);
void make_batch_tiles(const BatchJob &j, int ji, std::vector<BatchTile> &lds, std::vector<BatchTile> pass[3]);
void balance_queues(std::vector<BatchTile> &lds, std::vector<BatchTile> pass[3], std::vector<BatchTile> q[3]);
// Prepared spectrum of a resident replica in the layout of path `f`.
int replica_prepared(int id, const float2 **prep, long *n, const LineFFT *f);

// Device job/tile tables plus up to three persistent launches captured in
// one hipGraph.
struct GraphBatch {
  DevBuf d_jobs, d_tiles, d_meta; // meta: ntiles[0..3], counters[4..7]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nq = 0;
  int offset[3] = { 0, 0, 0 };
  size_t ntiles = 0;
  ~GraphBatch();
  void release_graph();
  int build(const std::vector<BatchJob> &jobs, std::vector<BatchTile> q[3], hipStream_t s);
  int launch(hipStream_t s) const;
};

} // namespace ovx
