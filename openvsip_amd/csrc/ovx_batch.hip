// ovx_batch.hip -- one persistent launch over an (instance, rows) tile queue.
//
// A batch of uneven fast-convolution jobs (different N and row counts) is cut
// into tiles of kBatchThreads threads' worth of work: a few whole rows of an
// LDS-resident job, or a few column/line groups of one row of a four-step
// job.  A fixed grid of workgroups (one per CU: the LDS footprint is sized for
// the largest tile) strides over the queue, workgroup w taking tiles w,
// w + grid, ..., and runs the same device bodies as the standalone kernels, so a
// job's result is bitwise the same alone and in a batch.  Four-step jobs need
// their three passes in order; the host splits a batch into up to three
// queues (pass 1 + LDS-resident work, pass 2, pass 3) launched back to back
// in one captured hipGraph.
#include "ovx_batch.hpp"
#include "ovx_fs_dev.hpp"
#include "ovx_hier_dev.hpp"
#include "ovx_ols_dev.hpp"
#include "ovx_row_dev.hpp"

namespace ovx {

// LDS of the largest tile (1024 threads): 64 register-Stockham rows of 256
// (17408 float2), four-step groups 4 x 16 x 273; every tile function checks
// its need against it.
constexpr int kBatchLds = 4 * 16 * 273 > 17408 ? 4 * 16 * 273 : 17408;

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

template <int N>
__device__ __noinline__ void tile_hier_fastconv(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using H = Hier<N>;
  static_assert((kBatchThreads / H::TR) * H::LROW <= kBatchLds, "tile LDS");
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
  hier_fastconv_body<N, kBatchThreads / H::TR == 1>(a, row0, lrow, t, lds + lrow * H::LROW);
}

template <int N>
__device__ __noinline__ void tile_hier_fwd(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using H = Hier<N>;
  static_assert((kBatchThreads / H::TR) * H::LROW <= kBatchLds, "tile LDS");
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

template <int N, int P>
__device__ __noinline__ void tile_row_fastconv(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using E = Engine<N, P, (N / P <= 64)>;
  static_assert((kBatchThreads / E::T) * E::LROW <= kBatchLds, "tile LDS");
  const int lrow = (int)threadIdx.x / E::T;
  const int t = (int)threadIdx.x % E::T;
  row_fastconv_body<N, P>(uniform(j.x), uniform(j.ldx), uniform(j.w), 0, uniform(j.y), uniform(j.ldy),
                          uniform(j.rows), uniform(j.tw0), 1.0f, row0, lrow, t, lds + lrow * E::LROW);
}

template <int N, int P>
__device__ __noinline__ void tile_row_fwd(lds_job &j, int row0_arg, lds_ptr lds)
{
  const int row0 = __builtin_amdgcn_readfirstlane(row0_arg); // uniform, see uniform()
  using E = Engine<N, P, (N / P <= 64)>;
  static_assert((kBatchThreads / E::T) * E::LROW <= kBatchLds, "tile LDS");
  const int lrow = (int)threadIdx.x / E::T;
  const int t = (int)threadIdx.x % E::T;
  row_fft_body<N, P, -1>(uniform(j.x), uniform(j.ldx), uniform(j.y), uniform(j.ldy), uniform(j.rows),
                         uniform(j.tw0), uniform(j.scale), row0, lrow, t, lds + lrow * E::LROW);
}

// Overlap-save segments seg0, seg0 + 1, ... (kBatchThreads / (L/16) of them
// per tile); a copy past the last segment runs with P = 0, which stores
// nothing, because ols_body's barriers need every thread.
template <int L>
__device__ __noinline__ void tile_ols(lds_job &j, int seg0_arg, lds_ptr lds)
{
  const int seg0 = __builtin_amdgcn_readfirstlane(seg0_arg); // uniform, see uniform()
  OlsArgs a = uniform(j.ols);
  constexpr int T = L / 16;
  static_assert(T <= kBatchThreads, "segment too long");
  static_assert((kBatchThreads / T) * Engine<L, 16>::LROW <= kBatchLds, "tile LDS");
  const int t = (int)threadIdx.x % T;
  const int copy = (int)threadIdx.x / T;
  const long seg = (long)seg0 + copy;
  const long nseg = ((a.P - 1) * a.decimation + a.c_step) / a.c_step; // segments of this job
  if (seg >= nseg) a.P = 0;
  ols_body<L, 16>(a, seg, t, lds + copy * Engine<L, 16>::LROW);
}

int batch_ols_segments_per_tile(int L) { return kBatchThreads / (L / 16); }

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

__global__ __launch_bounds__(kBatchThreads) void k_batch(const BatchJob *__restrict__ jobs,
                                                          const BatchTile *__restrict__ tiles,
                                                          const int *__restrict__ ntiles)
{
  __shared__ float2 lds_mem[kBatchLds];
  __shared__ BatchJob s_job;
  const lds_ptr lds = as_lds(lds_mem);
  lds_job &j = *(lds_job *)&s_job;
  static_assert(sizeof(BatchJob) % 4 == 0 && sizeof(BatchJob) / 4 <= kBatchThreads, "job copy");
  const int total = *ntiles;
  // workgroup w takes tiles w, w + grid, ...: consecutive tiles (rows of one
  // job, jobs of like size next to each other) go to different workgroups,
  // so each gets a like share of the queue (a shared atomic counter cost
  // each workgroup a cross-XCD round trip per tile)
  for (int ti = blockIdx.x; ti < total; ti += gridDim.x) {
    const BatchTile tile = tiles[ti];
    if (threadIdx.x < sizeof(BatchJob) / 4)
      ((int *)&s_job)[threadIdx.x] = ((const int *)&jobs[tile.job])[threadIdx.x];
    __syncthreads();
    switch (tile.op) {
    case BT_HIER_FASTCONV:
      switch (j.n) {
      case 1024: tile_hier_fastconv<1024>(j, tile.a, lds); break;
      case 2048: tile_hier_fastconv<2048>(j, tile.a, lds); break;
      case 4096: tile_hier_fastconv<4096>(j, tile.a, lds); break;
      case 8192: tile_hier_fastconv<8192>(j, tile.a, lds); break;
      case 16384: tile_hier_fastconv<16384>(j, tile.a, lds); break;
      default: break;
      }
      break;
    case BT_HIER_FFT_FWD:
      switch (j.n) {
      case 1024: tile_hier_fwd<1024>(j, tile.a, lds); break;
      case 2048: tile_hier_fwd<2048>(j, tile.a, lds); break;
      case 4096: tile_hier_fwd<4096>(j, tile.a, lds); break;
      case 8192: tile_hier_fwd<8192>(j, tile.a, lds); break;
      case 16384: tile_hier_fwd<16384>(j, tile.a, lds); break;
      default: break;
      }
      break;
    case BT_ROW_FASTCONV:
      switch (j.n) {
      case 256: tile_row_fastconv<256, 16>(j, tile.a, lds); break;
      case 512: tile_row_fastconv<512, 16>(j, tile.a, lds); break;
      case 2048: tile_row_fastconv<2048, 16>(j, tile.a, lds); break;
      case 4096: tile_row_fastconv<4096, 16>(j, tile.a, lds); break;
      case 3072: tile_row_fastconv<3072, 12>(j, tile.a, lds); break;
      case 6144: tile_row_fastconv<6144, 12>(j, tile.a, lds); break;
      case 12288: tile_row_fastconv<12288, 12>(j, tile.a, lds); break;
      default: break;
      }
      break;
    case BT_ROW_FFT_FWD:
      switch (j.n) {
      case 256: tile_row_fwd<256, 16>(j, tile.a, lds); break;
      case 512: tile_row_fwd<512, 16>(j, tile.a, lds); break;
      case 2048: tile_row_fwd<2048, 16>(j, tile.a, lds); break;
      case 4096: tile_row_fwd<4096, 16>(j, tile.a, lds); break;
      case 3072: tile_row_fwd<3072, 12>(j, tile.a, lds); break;
      case 6144: tile_row_fwd<6144, 12>(j, tile.a, lds); break;
      case 12288: tile_row_fwd<12288, 12>(j, tile.a, lds); break;
      default: break;
      }
      break;
    case BT_OLS:
      switch (j.ols_len) {
      case 1024: tile_ols<1024>(j, tile.a, lds); break;
      case 2048: tile_ols<2048>(j, tile.a, lds); break;
      case 4096: tile_ols<4096>(j, tile.a, lds); break;
      case 8192: tile_ols<8192>(j, tile.a, lds); break;
      case 16384: tile_ols<16384>(j, tile.a, lds); break;
      default: break;
      }
      break;
    default: break; // four-step tiles run in k_batch_fs
    }
    __syncthreads(); // LDS reuse by the next tile
  }
}

hipError_t launch_batch(const BatchJob *jobs, const BatchTile *tiles, const int *ntiles, int grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(kBatchThreads), 0, s, jobs, tiles, ntiles);
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

// Batch tiles of the 3 * 2^k lengths use 12 values per thread, so that 1024
// threads fit the 128-VGPR budget of a 1024-thread workgroup; the powers of
// two use 16.
int batch_row_p(int n) { return (n == 256 || n == 512 || n == 2048 || n == 4096) ? 16 : 12; }

int batch_rows_per_tile(int path, int n)
{
  if (path == BP_HIER) return kBatchThreads / (64 * (n / 1024));
  if (path == BP_ROW) return kBatchThreads / (n / batch_row_p(n));
  return 1;
}

int batch_fs_lines_per_group(int n_line)
{
  if (n_line == 256) return FsCfg<256>::C;
  if (n_line == 512) return FsCfg<512>::C;
  return 0;
}

} // namespace ovx
