// ovx_batch.hip -- persistent launches over (instance, rows) tile queues.
//
// A batch of uneven fast-convolution jobs (different N and row counts) is cut
// into tiles: a few whole rows of an LDS-resident job (one workgroup's worth
// of threads), or one column/line group of one row of a four-step job.  The
// LDS-resident tiles run in two persistent launches, one of 1024-thread
// workgroups (one per CU; rows of 12288 and 16384 points) and one of
// 512-thread workgroups (two per CU; every shorter row), whose fixed grids
// stride over their queues, workgroup w taking tiles w, w + grid, ...  Every
// tile runs the same device body as the standalone kernels, so a job's
// result is bitwise the same alone and in a batch.  Four-step jobs need their
// three passes in order; each pass is a launch of its own over its queue, and
// the host captures all of a batch's launches in one hipGraph.
#include "ovx_batch.hpp"
#include "ovx_fs_dev.hpp"
#include "ovx_hier_dev.hpp"
#include "ovx_ols_dev.hpp"
#include "ovx_row_dev.hpp"

namespace ovx {

// LDS of the largest tile of each persistent launch.  1024 threads: 64
// register-Stockham rows of 256 (17408 float2), four-step groups
// 4 x 16 x 273.  512 threads: every tile routed there (rows of up to 8192
// points, overlap-save segments) needs 16 lines' worth of 512 threads,
// 16 * 512 * 33 / 32 = 8448 float2 (67.6 KB: two workgroups per CU).  Every
// tile function checks its need against its launch's figure.
constexpr int kBatchLds = 4 * 16 * 273 > 17408 ? 4 * 16 * 273 : 17408;
constexpr int kBatchMidLds = 8448;
template <int NT>
struct BatchLds {
  static constexpr int value = NT == kBatchThreads ? kBatchLds : kBatchMidLds;
};

// The job descriptor of the current tile lives in LDS (copied once per tile
// by the persistent loop); the tile functions read it from there.
typedef __attribute__((address_space(3))) const BatchJob lds_job;

// The tile functions are called (not inlined), so the compiler cannot see
// that the descriptor is the same for every lane: pointers read from LDS
// would sit in vector registers and every buffer access built on them would
// need a loop over the distinct values.  Each field a tile function uses is
// read into scalar registers through uniform().
template <typename T>
__device__ __forceinline__ T uniform(const __attribute__((address_space(3))) T &f)
{
  static_assert(sizeof(T) % 4 == 0, "dword copy");
  T u;
  const __attribute__((address_space(3))) int *src = (const __attribute__((address_space(3))) int *)&f;
  int *dst = (int *)&u;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return u;
}

// Tile functions of a workgroup of NT threads: NT / (threads of a row) rows
// (or overlap-save segments) per tile.
template <int NT, int N>
__device__ __noinline__ void tile_hier_fastconv(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using H = Hier<N>;
  static_assert(NT % H::TR == 0 && (NT / H::TR) * H::LROW <= BatchLds<NT>::value, "tile LDS");
  HierArgs a{};
  a.x = uniform(j.x);
  a.ldx = uniform(j.ldx);
  a.y = uniform(j.y);
  a.ldy = uniform(j.ldy);
  a.w = uniform(j.w);
  a.ldw = 0;
  a.rows = uniform(j.rows);
  a.tw_outer = uniform(j.tw0);
  a.tw_wave = uniform(j.tw1);
  a.scale = 1.0f;
  a.natural = 0;
  const int lrow = (int)threadIdx.x / H::TR;
  const int t = (int)threadIdx.x % H::TR;
  // a tile of one row is only made for a row that exists
  hier_fastconv_body<N, NT / H::TR == 1>(a, row0, lrow, t, lds + lrow * H::LROW);
}

template <int NT, int N>
__device__ __noinline__ void tile_hier_fwd(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using H = Hier<N>;
  static_assert(NT % H::TR == 0 && (NT / H::TR) * H::LROW <= BatchLds<NT>::value, "tile LDS");
  HierArgs a{};
  a.x = uniform(j.x);
  a.ldx = uniform(j.ldx);
  a.y = uniform(j.y);
  a.ldy = uniform(j.ldy);
  a.rows = uniform(j.rows);
  a.tw_outer = uniform(j.tw0);
  a.tw_wave = uniform(j.tw1);
  a.scale = uniform(j.scale);
  const int lrow = (int)threadIdx.x / H::TR;
  const int t = (int)threadIdx.x % H::TR;
  hier_fwd_body<N>(a, row0, lrow, t, lds + lrow * H::LROW);
}

template <int NT, int N, int P>
__device__ __noinline__ void tile_row_fastconv(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using E = Engine<N, P, (N / P <= 64)>;
  static_assert(NT % E::T == 0 && (NT / E::T) * E::LROW <= BatchLds<NT>::value, "tile LDS");
  const int lrow = (int)threadIdx.x / E::T;
  const int t = (int)threadIdx.x % E::T;
  row_fastconv_body<N, P>(uniform(j.x), uniform(j.ldx), uniform(j.w), 0, uniform(j.y), uniform(j.ldy),
                          uniform(j.rows), uniform(j.tw0), 1.0f, row0, lrow, t, lds + lrow * E::LROW);
}

template <int NT, int N, int P>
__device__ __noinline__ void tile_row_fwd(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using E = Engine<N, P, (N / P <= 64)>;
  static_assert(NT % E::T == 0 && (NT / E::T) * E::LROW <= BatchLds<NT>::value, "tile LDS");
  const int lrow = (int)threadIdx.x / E::T;
  const int t = (int)threadIdx.x % E::T;
  row_fft_body<N, P, -1>(uniform(j.x), uniform(j.ldx), uniform(j.y), uniform(j.ldy), uniform(j.rows),
                         uniform(j.tw0), uniform(j.scale), row0, lrow, t, lds + lrow * E::LROW);
}

// Overlap-save segments seg0, seg0 + 1, ... (NT / (L/16) of them per tile); a
// copy past the last segment runs with P = 0, which stores nothing, because
// ols_body's barriers need every thread.
template <int NT, int L>
__device__ __noinline__ void tile_ols(lds_job &j, int seg0_arg, lds_ptr lds)
{
  const int seg0 = __builtin_amdgcn_readfirstlane(seg0_arg); // uniform, see uniform()
  OlsArgs a = uniform(j.ols);
  constexpr int T = L / 16;
  static_assert(NT % T == 0 && (NT / T) * Engine<L, 16>::LROW <= BatchLds<NT>::value, "tile LDS");
  const int t = (int)threadIdx.x % T;
  const int copy = (int)threadIdx.x / T;
  const long seg = (long)seg0 + copy;
  const long nseg = ((a.P - 1) * a.decimation + a.c_step) / a.c_step; // segments of this job
  if (seg >= nseg) a.P = 0;
  ols_body<L, 16>(a, seg, t, lds + copy * Engine<L, 16>::LROW);
}

int batch_ols_segments_per_tile(int L)
{
  return (batch_ols_mid_routed(L) ? kBatchMidThreads : kBatchThreads) / (L / 16);
}

template <typename J>
__device__ __forceinline__ FsArgs fs_args_of(const J &j)
{
  FsArgs a{};
  a.n1 = j.n1;
  a.n2 = j.n2;
  a.nbig = j.n;
  a.tw1 = j.tw0;
  a.tw2 = j.tw1;
  a.tlo = j.tlo;
  a.thi = j.thi;
  a.sbits = j.sbits;
  a.tfull = j.tfull;
  a.wp = j.w;
  a.ldwp = 0;
  a.scale = 1.0f;
  return a;
}

// Whether a tile of (path, n) -- or an overlap-save segment of length L --
// runs in the launch of NT threads: the host routes each to exactly one of
// the two (batch_queue, batch_ols_queue), and the other launch's switch has
// no code for it.
template <int NT, int PATH, int N>
constexpr bool runs_in() { return batch_mid_routed(PATH, N) == (NT == kBatchMidThreads); }
template <int NT, int L>
constexpr bool ols_runs_in() { return batch_ols_mid_routed(L) == (NT == kBatchMidThreads); }
static_assert(batch_mid_routed(BP_HIER, 8192) && !batch_mid_routed(BP_HIER, 16384) &&
                  batch_mid_routed(BP_ROW, 6144) && !batch_mid_routed(BP_ROW, 12288) &&
                  batch_mid_routed(BP_ROW, 256) && batch_ols_mid_routed(8192) && !batch_ols_mid_routed(16384),
              "rows of 12288 and 16384 points (and segments of 16384) need the 1024-thread launch, all others "
              "run in the 512-thread one");

template <int NT, int N>
__device__ __forceinline__ void run_hier(bool fwd_only, lds_job &j, int a, lds_ptr lds)
{
  if constexpr (runs_in<NT, BP_HIER, N>()) {
    if (fwd_only) tile_hier_fwd<NT, N>(j, a, lds);
    else tile_hier_fastconv<NT, N>(j, a, lds);
  }
}
template <int NT, int N, int P>
__device__ __forceinline__ void run_row(bool fwd_only, lds_job &j, int a, lds_ptr lds)
{
  if constexpr (runs_in<NT, BP_ROW, N>()) {
    if (fwd_only) tile_row_fwd<NT, N, P>(j, a, lds);
    else tile_row_fastconv<NT, N, P>(j, a, lds);
  }
}
template <int NT, int L>
__device__ __forceinline__ void run_ols(lds_job &j, int a, lds_ptr lds)
{
  if constexpr (ols_runs_in<NT, L>()) tile_ols<NT, L>(j, a, lds);
}

// The persistent loop of a launch of NT threads: workgroup w takes tiles w,
// w + grid, ...: consecutive tiles (rows of one job, jobs of like size next to
// each other) go to different workgroups, so each gets a like share of the
// queue (a shared atomic counter cost each workgroup a cross-XCD round trip
// per tile).
template <int NT>
__device__ __forceinline__ void batch_loop(const BatchJob *__restrict__ jobs, const BatchTile *__restrict__ tiles,
                                           const int *__restrict__ ntiles, float2 *lds_mem, BatchJob *s_job)
{
  const lds_ptr lds = as_lds(lds_mem);
  lds_job &j = *(lds_job *)s_job;
  static_assert(sizeof(BatchJob) % 4 == 0 && sizeof(BatchJob) / 4 <= NT, "job copy");
  const int total = *ntiles;
  for (int ti = blockIdx.x; ti < total; ti += gridDim.x) {
    const BatchTile tile = tiles[ti];
    if (threadIdx.x < sizeof(BatchJob) / 4) ((int *)s_job)[threadIdx.x] = ((const int *)&jobs[tile.job])[threadIdx.x];
    __syncthreads();
    switch (tile.op) {
    case BT_HIER_FASTCONV:
    case BT_HIER_FFT_FWD: {
      const bool fwd = tile.op == BT_HIER_FFT_FWD;
      switch (j.n) {
      case 1024: run_hier<NT, 1024>(fwd, j, tile.a, lds); break;
      case 2048: run_hier<NT, 2048>(fwd, j, tile.a, lds); break;
      case 4096: run_hier<NT, 4096>(fwd, j, tile.a, lds); break;
      case 8192: run_hier<NT, 8192>(fwd, j, tile.a, lds); break;
      case 16384: run_hier<NT, 16384>(fwd, j, tile.a, lds); break;
      default: break;
      }
      break;
    }
    case BT_ROW_FASTCONV:
    case BT_ROW_FFT_FWD: {
      const bool fwd = tile.op == BT_ROW_FFT_FWD;
      switch (j.n) {
      case 256: run_row<NT, 256, 16>(fwd, j, tile.a, lds); break;
      case 512: run_row<NT, 512, 16>(fwd, j, tile.a, lds); break;
      case 2048: run_row<NT, 2048, 16>(fwd, j, tile.a, lds); break;
      case 4096: run_row<NT, 4096, 16>(fwd, j, tile.a, lds); break;
      case 3072: run_row<NT, 3072, 12>(fwd, j, tile.a, lds); break;
      case 6144: run_row<NT, 6144, 12>(fwd, j, tile.a, lds); break;
      case 12288: run_row<NT, 12288, 12>(fwd, j, tile.a, lds); break;
      default: break;
      }
      break;
    }
    case BT_OLS:
      switch (j.ols_len) {
      case 1024: run_ols<NT, 1024>(j, tile.a, lds); break;
      case 2048: run_ols<NT, 2048>(j, tile.a, lds); break;
      case 4096: run_ols<NT, 4096>(j, tile.a, lds); break;
      case 8192: run_ols<NT, 8192>(j, tile.a, lds); break;
      case 16384: run_ols<NT, 16384>(j, tile.a, lds); break;
      default: break;
      }
      break;
    default: break; // four-step tiles run in k_batch_fs
    }
    __syncthreads(); // LDS reuse by the next tile
  }
}

__global__ __launch_bounds__(kBatchThreads) void k_batch(const BatchJob *__restrict__ jobs,
                                                          const BatchTile *__restrict__ tiles,
                                                          const int *__restrict__ ntiles)
{
  __shared__ float2 lds_mem[kBatchLds];
  __shared__ BatchJob s_job;
  batch_loop<kBatchThreads>(jobs, tiles, ntiles, lds_mem, &s_job);
}

// Two workgroups per CU: four waves per SIMD (the second launch bound is
// waves per SIMD; the tile functions inherit it, so they fit 128 VGPRs) and
// 2 x (67.6 KB + the descriptor) of LDS.
__global__ __launch_bounds__(kBatchMidThreads, kBatchMidBlocksPerCu * kBatchMidThreads / 256) void k_batch_mid(
    const BatchJob *__restrict__ jobs, const BatchTile *__restrict__ tiles, const int *__restrict__ ntiles)
{
  __shared__ float2 lds_mem[kBatchMidLds];
  __shared__ BatchJob s_job;
  batch_loop<kBatchMidThreads>(jobs, tiles, ntiles, lds_mem, &s_job);
}

hipError_t launch_batch(const BatchJob *jobs, const BatchTile *tiles, const int *ntiles, int grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(kBatchThreads), 0, s, jobs, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_batch_mid(const BatchJob *jobs, const BatchTile *tiles, const int *ntiles, int grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_mid, dim3(grid), dim3(kBatchMidThreads), 0, s, jobs, tiles, ntiles);
  return hipGetLastError();
}

// ---------------------------------------------------------------- four-step passes
// One line group of one row per workgroup (the standalone kernels' shape;
// four waves per SIMD, the occupancy the LDS of L = 256 groups allows), so
// co-resident workgroups run independently and overlap their loads with
// other groups' arithmetic; workgroup w takes tiles w, w + grid, ...  The job
// descriptor is read from global memory (uniform per tile, scalar loads).
template <int L, int PASS>
__global__ __launch_bounds__(FsCfg<L>::NT, 4) void k_batch_fs(const BatchJob *__restrict__ jobs,
                                                               const BatchTile *__restrict__ tiles,
                                                               const int *__restrict__ ntiles)
{
  using F = FsCfg<L>;
  __shared__ float2 lds[F::LDS];
  const int total = *ntiles;
  const int tid = (int)threadIdx.x;
  for (int ti = blockIdx.x; ti < total; ti += gridDim.x) {
    const BatchTile tile = tiles[ti];
    const BatchJob &j = jobs[tile.job];
    FsArgs a = fs_args_of(j);
    if constexpr (PASS == 0) { // column FFT * twiddle: fastconv x -> y, Fftm x -> workspace
      const bool fft = tile.op == BT_FS_FFT_A;
      if (fft) a.scale = j.scale;
      a.in = j.x;
      a.ld_in = j.ldx;
      a.out = fft ? j.ws : j.y;
      a.ld_out = fft ? (long)j.n : j.ldy;
      fs_col_body<L, F::P, F::C, 0>(a, tile.a, tile.b, tid, lds);
    } else if constexpr (PASS == 1) {
      if (tile.op == BT_FS_B) { // fastconv: row FFT * W * IFFT * twiddle in place on y
        a.in = j.y;
        a.ld_in = j.ldy;
        a.out = j.y;
        a.ld_out = j.ldy;
        fs_row_body<L, F::P, F::C, 0>(a, tile.a, tile.b, tid, lds);
      } else { // Fftm: workspace -> y in natural order, scaled
        a.scale = j.scale;
        a.in = j.ws;
        a.ld_in = j.n;
        a.out = j.y;
        a.ld_out = j.ldy;
        fs_row_body<L, F::P, F::C, 1>(a, tile.a, tile.b, tid, lds);
      }
    } else { // fastconv pass 3: column IFFT in place on y
      a.in = j.y;
      a.ld_in = j.ldy;
      a.out = j.y;
      a.ld_out = j.ldy;
      fs_col_body<L, F::P, F::C, 1>(a, tile.a, tile.b, tid, lds);
    }
    __syncthreads(); // LDS reuse by the next tile
  }
}

template <int L>
static hipError_t launch_fs_pass(int pass, const BatchJob *jobs, const BatchTile *tiles, const int *ntiles, int grid,
                                 hipStream_t s)
{
  constexpr int NT = FsCfg<L>::NT;
  if (pass == 0) hipLaunchKernelGGL((k_batch_fs<L, 0>), dim3(grid), dim3(NT), 0, s, jobs, tiles, ntiles);
  else if (pass == 1) hipLaunchKernelGGL((k_batch_fs<L, 1>), dim3(grid), dim3(NT), 0, s, jobs, tiles, ntiles);
  else hipLaunchKernelGGL((k_batch_fs<L, 2>), dim3(grid), dim3(NT), 0, s, jobs, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_batch_fs(int n_line, int pass, const BatchJob *jobs, const BatchTile *tiles, const int *ntiles,
                           int grid, hipStream_t s)
{
  if (grid <= 0) return hipSuccess;
  if (n_line == 256) return launch_fs_pass<256>(pass, jobs, tiles, ntiles, grid, s);
  if (n_line == 512) return launch_fs_pass<512>(pass, jobs, tiles, ntiles, grid, s);
  return hipErrorInvalidValue;
}

// Workgroups of a four-step pass that fit one CU together (LDS-bound).
int batch_fs_blocks_per_cu(int n_line)
{
  if (n_line == 256) return (160 * 1024) / (FsCfg<256>::LDS * 8);
  if (n_line == 512) return (160 * 1024) / (FsCfg<512>::LDS * 8);
  return 1;
}

bool batch_supported(int path, int n)
{
  if (path == BP_HIER) return hier_supported(n);
  if (path == BP_ROW) return n == 256 || n == 512 || n == 2048 || n == 4096 || n == 3072 || n == 6144 || n == 12288;
  if (path == BP_FS) return n == 65536 || n == 131072 || n == 262144;
  return false;
}

int batch_rows_per_tile(int path, int n)
{
  if (path != BP_HIER && path != BP_ROW) return 1;
  return (batch_mid_routed(path, n) ? kBatchMidThreads : kBatchThreads) / batch_line_threads(path, n);
}

int batch_fs_lines_per_group(int n_line)
{
  if (n_line == 256) return FsCfg<256>::C;
  if (n_line == 512) return FsCfg<512>::C;
  return 0;
}

} // namespace ovx
