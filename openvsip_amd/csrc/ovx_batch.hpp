// ovx_batch.hpp -- persistent tile-queue kernel for batches of uneven jobs.
#pragma once
#include <hip/hip_runtime.h>

#include "ovx_ols.hpp"

namespace ovx {

// Kernel path of one job (which device body its tiles run).
enum BatchPath { BP_HIER = 1, BP_ROW = 2, BP_FS = 3 };
// Tile operations.
enum BatchOp {
  BT_HIER_FASTCONV = 1, // rows [a, a + rows_per_tile) of a hierarchical-path job
  BT_ROW_FASTCONV = 2,  // rows [a, ...) of a register-Stockham-path job
  BT_FS_A = 3,          // four-step fastconv pass 1: row a, column group b
  BT_FS_B = 4,          // pass 2: row a, line group b
  BT_FS_C = 5,          // pass 3: row a, column group b
  BT_HIER_FFT_FWD = 6,  // forward Fftm rows of a hierarchical-path job
  BT_ROW_FFT_FWD = 7,   // forward Fftm rows of a register-Stockham-path job
  BT_FS_FFT_A = 8,      // four-step forward Fftm pass 1: x -> workspace
  BT_FS_FFT_B = 9,      // pass 2: workspace -> y (natural order)
  BT_OLS = 10           // one overlap-save segment (a = segment index)
};

struct BatchJob {
  int path;
  int n;
  long rows;
  const float2 *x;
  long ldx;
  float2 *y;
  long ldy;
  const float2 *w;   // prepared spectrum (path layout, 1/n folded in)
  const float2 *tw0; // HIER: outer table; ROW: engine table; FS: n1 line table
  const float2 *tw1; // HIER: 1024-point wave table; FS: n2 line table
  const float2 *tlo, *thi;
  int sbits;
  const float2 *tfull; // FS: full twiddle table or null
  int n1, n2;
  float scale;       // forward Fftm scale
  float2 *ws;        // four-step Fftm workspace (rows x n)
  OlsArgs ols;       // overlap-save segment parameters
  int ols_len;       // segment length L
};

struct BatchTile {
  int job;
  int op;
  int a; // first row
  int b; // line group (four-step)
};

// Tile queues of a batch: queue 0 holds the LDS-resident tiles (rows,
// overlap-save segments) of the 1024-thread persistent launch, queue
// kMidQueue those of the 512-thread one; the four-step passes get one queue
// per (pass, line length 256 or 512), queues 1 .. kFsQueueEnd - 1, run in
// pass order, one line group per workgroup.
constexpr int kBatchQueues = 8;
constexpr int kFsQueueEnd = 7;
constexpr int kMidQueue = 7;
constexpr int kFsLines[2] = { 256, 512 };
inline int fs_queue(int pass, int n_line) { return 1 + 2 * pass + (n_line == 512 ? 1 : 0); }
inline int fs_queue_line(int q) { return kFsLines[(q - 1) % 2]; }

// Values per thread of a register-Stockham batch tile: 12 for the 3 * 2^k
// lengths (so that 1024 threads fit the 128-VGPR budget of a 1024-thread
// workgroup), 16 for the powers of two.
constexpr int batch_row_p(int n) { return (n == 256 || n == 512 || n == 2048 || n == 4096) ? 16 : 12; }

// Threads per persistent workgroup.  Rows whose line takes at most 512
// threads (every LDS-resident length up to 8192; overlap-save segments up to
// 8192) run in the 512-thread launch, two workgroups per CU, so that one
// workgroup's loads and stores overlap the other's arithmetic; 12288 and
// 16384 need the 1024-thread workgroup, one per CU.
constexpr int kBatchThreads = 1024;
constexpr int kBatchMidThreads = 512;
constexpr int batch_line_threads(int path, int n) // threads of one row's transform
{
  return path == BP_HIER ? 64 * (n / 1024) : n / batch_row_p(n);
}
constexpr bool batch_mid_routed(int path, int n) { return batch_line_threads(path, n) <= kBatchMidThreads; }
constexpr bool batch_ols_mid_routed(int L) { return L / 16 <= kBatchMidThreads; }
inline int batch_queue(int path, int n) { return batch_mid_routed(path, n) ? kMidQueue : 0; }
inline int batch_ols_queue(int L) { return batch_ols_mid_routed(L) ? kMidQueue : 0; }
int batch_rows_per_tile(int path, int n);
int batch_ols_segments_per_tile(int L);
int batch_fs_lines_per_group(int n_line); // C of a four-step group of lines of length n_line
int batch_fs_blocks_per_cu(int n_line);   // co-resident workgroups of a four-step pass launch
bool batch_supported(int path, int n);

// One persistent launch over tiles[0 .. *ntiles): `grid` workgroups, workgroup
// w taking tiles w, w + grid, ...
hipError_t launch_batch(const BatchJob *jobs, const BatchTile *tiles, const int *ntiles, int grid, hipStream_t s);
// The same over the 512-thread queue (two workgroups per CU).
hipError_t launch_batch_mid(const BatchJob *jobs, const BatchTile *tiles, const int *ntiles, int grid, hipStream_t s);
constexpr int kBatchMidBlocksPerCu = 2;
// Four-step pass `pass` (0, 1, 2) over tiles[0 .. *ntiles) of line length
// n_line (256 or 512): `grid` workgroups of one line group each take tiles
// blockIdx.x, blockIdx.x + grid, ... (no shared counter).
hipError_t launch_batch_fs(int n_line, int pass, const BatchJob *jobs, const BatchTile *tiles, const int *ntiles,
                           int grid, hipStream_t s);
inline int fs_queue_pass(int q) { return (q - 1) / 2; }

} // namespace ovx
