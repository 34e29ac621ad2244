// ovx_hier.hip -- hierarchical LDS-resident transforms for N = R0 * 1024
// (N = 2048 ... 16384), the hot shapes of the fast-convolution path.
//
// A row is split as n = n2 + 1024*n1 (n2 < 1024, n1 < R0).  With
// k = k1 + R0*k2:
//   X[k1 + R0 k2] = sum_n2 W_1024^{n2 k2} [ W_N^{n2 k1} sum_n1 x[n2 + 1024 n1] W_R0^{n1 k1} ]
// 1. outer stage: every thread loads R0-point columns x[j + 1024 n1] straight
//    from HBM (coalesced), runs DFT_R0 and the twiddle W_N^{j k1} and writes
//    slice k1 of the LDS row;                                 block barrier
// 2. wave stage: wave k1 owns slice k1 and runs a 1024-point Stockham FFT on
//    it (16 values per lane, exchanges synchronised within the wave only);
//    for fast convolution it then multiplies by the spectrum in [k1][k2]
//    order, runs the inverse 1024-point FFT and the inverse twiddle
//    W_N^{-n2 k1}, all without leaving the wave;             block barrier
// 3. outer inverse stage: DFT_R0 over k1, stores straight to HBM.
// Waves run unsynchronised through the whole middle of the computation, so
// LDS traffic of one wave overlaps arithmetic of the others; the fused
// fastconv needs two workgroup barriers in total.
#include "ovx_hier_dev.hpp"

namespace ovx {

// ---------------------------------------------------------------- kernels
template <int N, bool SPLIT>
__global__ __launch_bounds__(Hier<N>::NT) void k_hier_fastconv(HierArgs a)
{
  using H = Hier<N>;
  __shared__ float2 lds_all[H::B * H::LROW];
  const int lrow = (H::B == 1) ? 0 : (int)threadIdx.x / H::TR;
  const int t = (H::B == 1) ? (int)threadIdx.x : (int)threadIdx.x % H::TR;
  // one row per block and a grid of exactly `rows` blocks: every row exists
  hier_fastconv_body<N, H::B == 1, SPLIT>(a, (long)blockIdx.x * H::B, lrow, t, lds_all + lrow * H::LROW);
}

template <int N, bool SPLIT>
__global__ __launch_bounds__(Hier<N>::NT) void k_hier_fft_fwd(HierArgs a)
{
  using H = Hier<N>;
  __shared__ float2 lds_all[H::B * H::LROW];
  const int lrow = (H::B == 1) ? 0 : (int)threadIdx.x / H::TR;
  const int t = (H::B == 1) ? (int)threadIdx.x : (int)threadIdx.x % H::TR;
  hier_fwd_body<N, SPLIT>(a, (long)blockIdx.x * H::B, lrow, t, lds_all + lrow * H::LROW);
}

template <int N, bool SPLIT>
__global__ __launch_bounds__(Hier<N>::NT) void k_hier_fft_inv(HierArgs a)
{
  using H = Hier<N>;
  __shared__ float2 lds_all[H::B * H::LROW];
  const int lrow = (H::B == 1) ? 0 : (int)threadIdx.x / H::TR;
  const int t = (H::B == 1) ? (int)threadIdx.x : (int)threadIdx.x % H::TR;
  hier_inv_body<N, SPLIT>(a, (long)blockIdx.x * H::B, lrow, t, lds_all + lrow * H::LROW);
}

// Spectrum permutation to [k1][k2] order with scale folded in:
// wp[row][k1*1024 + k2] = scale * w[row*ldw + k1 + R0*k2].
__global__ void k_hier_permute(const float2 *__restrict__ w, long ldw, float2 *__restrict__ wp, int n, int r0,
                               float scale)
{
  const long row = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k1 = i >> 10, k2 = i & 1023;
  wp[row * n + i] = cscale(w[row * ldw + k1 + r0 * k2], scale);
}

// The same permutation read from a split-storage spectrum vector:
// wp[k1*1024 + k2] = scale * (wr + i wi)[k1 + R0*k2].
__global__ void k_hier_permute_split(const float *__restrict__ wr, const float *__restrict__ wi,
                                     float2 *__restrict__ wp, int n, int r0, float scale)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int src = (i >> 10) + r0 * (i & 1023);
  wp[i] = make_float2(scale * wr[src], scale * wi[src]);
}

// Outer twiddles: two[(k1-1)*1024 + j] = exp(-2 pi i j k1 / N), FP64-evaluated.
__global__ void k_hier_outer_twiddles(float2 *two, int n, int count)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int k1 = i / 1024 + 1, j = i % 1024;
  long m = (long)j * k1;
  double sn, cs;
  sincospi(-2.0 * (double)m / (double)n, &sn, &cs);
  two[i] = make_float2((float)cs, (float)sn);
}

// ---------------------------------------------------------------- dispatch
template <int N>
static hipError_t launch_kind(int kind, const HierArgs &a, hipStream_t s)
{
  using H = Hier<N>;
  dim3 grid((unsigned)((a.rows + H::B - 1) / H::B)), block(H::NT);
  if (kind == HIER_FASTCONV) hipLaunchKernelGGL((k_hier_fastconv<N, false>), grid, block, 0, s, a);
  else if (kind == HIER_FASTCONV_SPLIT) hipLaunchKernelGGL((k_hier_fastconv<N, true>), grid, block, 0, s, a);
  else if (kind == HIER_FWD) hipLaunchKernelGGL((k_hier_fft_fwd<N, false>), grid, block, 0, s, a);
  else if (kind == HIER_INV) hipLaunchKernelGGL((k_hier_fft_inv<N, false>), grid, block, 0, s, a);
  else if (kind == HIER_FWD_SPLIT) hipLaunchKernelGGL((k_hier_fft_fwd<N, true>), grid, block, 0, s, a);
  else if (kind == HIER_INV_SPLIT) hipLaunchKernelGGL((k_hier_fft_inv<N, true>), grid, block, 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

bool hier_supported(long n) { return n == 1024 || n == 2048 || n == 4096 || n == 8192 || n == 16384; }

int hier_rows_per_block(long n)
{
  switch (n) {
  case 1024: return Hier<1024>::B;
  case 2048: return Hier<2048>::B;
  case 4096: return Hier<4096>::B;
  case 8192: return Hier<8192>::B;
  case 16384: return Hier<16384>::B;
  default: return 0;
  }
}

hipError_t launch_hier(int kind, long n, const HierArgs &a, hipStream_t s)
{
  if (a.rows <= 0) return hipSuccess;
  switch (n) {
  case 1024: return launch_kind<1024>(kind, a, s);
  case 2048: return launch_kind<2048>(kind, a, s);
  case 4096: return launch_kind<4096>(kind, a, s);
  case 8192: return launch_kind<8192>(kind, a, s);
  case 16384: return launch_kind<16384>(kind, a, s);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_hier_permute(const float2 *w, long ldw, float2 *wp, long n, long rows, float scale, hipStream_t s)
{
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)rows);
  hipLaunchKernelGGL(k_hier_permute, grid, dim3(256), 0, s, w, ldw, wp, (int)n, (int)(n / 1024), scale);
  return hipGetLastError();
}

hipError_t launch_hier_permute_split(const float *wr, const float *wi, float2 *wp, long n, float scale,
                                     hipStream_t s)
{
  hipLaunchKernelGGL(k_hier_permute_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, wr, wi, wp, (int)n,
                     (int)(n / 1024), scale);
  return hipGetLastError();
}

hipError_t launch_hier_twiddles(long n, float2 *two, float2 *tw_wave, hipStream_t s)
{
  int count = (int)(n / 1024 - 1) * 1024;
  if (count > 0) {
    hipLaunchKernelGGL(k_hier_outer_twiddles, dim3((count + 255) / 256), dim3(256), 0, s, two, (int)n, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return launch_twiddles(Engine<1024, 16, true>::plan, tw_wave, s);
}

int hier_wave_twiddle_count() { return Engine<1024, 16, true>::plan.twtotal; }

} // namespace ovx
