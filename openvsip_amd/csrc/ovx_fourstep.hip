// ovx_fourstep.hip -- transforms and fast convolution for lines longer than
// one workgroup's LDS (N > 16384) on gfx950.
//
// A line of N = N1*N2 is viewed as an N1 x N2 row-major matrix,
// n = N2*n1 + n2.  Forward:
//   column pass:  B[k1][n2] = W_N^{n2 k1} * sum_n1 A[n1][n2] W_N1^{n1 k1}
//   row pass:     X[k1 + N1 k2] = sum_n2 B[k1][n2] W_N2^{n2 k2}
// Inverse (k = k1 + N1 k2, n = N2 n1 + n2):
//   row pass:     U[k1][n2] = W_N^{-n2 k1} * sum_k2 X[k1 + N1 k2] W_N2^{-n2 k2}
//   column pass:  x[N2 n1 + n2] = sum_k1 U[k1][n2] W_N1^{-n1 k1}
// Fftm uses column pass + row pass with an LDS-staged transposed store
// (output in natural order).  Fast convolution needs no transpose at all:
// column pass (forward), one row pass doing forward row FFT, multiply by the
// spectrum permuted to [k1][k2] order, inverse row FFT and the inverse
// twiddle, then the inverse column pass -- three passes over HBM.
//
// The column pass moves C adjacent columns per block with the column index
// fastest across lanes, so every global access is a C*8-byte contiguous run.
// W_N^m is rebuilt from two small FP64-generated tables:
// W_N^m = Thi[m >> S] * Tlo[m & (2^S - 1)].
#include "ovx_fs_dev.hpp"

namespace ovx {

template <int L, int P, int C, int MODE, bool SIN, bool SOUT>
__global__ __launch_bounds__(C *(L / P)) void k_fs_col(FsArgs a)
{
  __shared__ float2 lds_all[C * FsLds<L>::LROW];
  fs_col_body<L, P, C, MODE, SIN, SOUT>(a, blockIdx.y, blockIdx.x, threadIdx.x, lds_all);
}

template <int L, int P, int C, int MODE, bool SOUT>
__global__ __launch_bounds__(C *(L / P)) void k_fs_row(FsArgs a)
{
  __shared__ float2 lds_all[C * FsLds<L>::LROW];
  fs_row_body<L, P, C, MODE, SOUT>(a, blockIdx.y, blockIdx.x, threadIdx.x, lds_all);
}

// Wp[k1][k2] = scale * W[k1 + N1 k2] for `rows` coefficient rows (ldw = 0:
// a single vector).
__global__ void k_fs_permute(const float2 *__restrict__ w, long ldw, float2 *__restrict__ wp, long n1, long n2,
                             float scale)
{
  const long n = n1 * n2;
  const long row = blockIdx.y;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x; // index into wp (k1-major)
  if (i >= n) return;
  long k1 = i / n2, k2 = i % n2;
  float2 v = w[row * ldw + k1 + n1 * k2];
  wp[row * n + i] = cscale(v, scale);
}

// ---------------------------------------------------------------- dispatch

// Passes that read or write split storage move 4-byte values, so their
// groups take 32 lines where the thread count allows (128-byte runs per
// array instead of 64).
template <int L>
struct FsSplitCfg {
  static constexpr int P = FsCfg<L>::P;
  static constexpr int C = (L / P) * 32 <= 1024 ? 32 : FsCfg<L>::C;
  static constexpr int NT = C * (L / P);
};

template <int L, int MODE>
static hipError_t col_launch(const FsArgs &a, long rows, hipStream_t s)
{
  using F = FsCfg<L>;
  using G = FsSplitCfg<L>;
  // split storage is read by a call's first pass or written by its last
  if (a.in_re)
    hipLaunchKernelGGL((k_fs_col<L, G::P, G::C, MODE, true, false>), dim3((unsigned)(a.n2 / G::C), (unsigned)rows),
                       dim3(G::NT), 0, s, a);
  else if (a.out_re)
    hipLaunchKernelGGL((k_fs_col<L, G::P, G::C, MODE, false, true>), dim3((unsigned)(a.n2 / G::C), (unsigned)rows),
                       dim3(G::NT), 0, s, a);
  else
    hipLaunchKernelGGL((k_fs_col<L, F::P, F::C, MODE, false, false>), dim3((unsigned)(a.n2 / F::C), (unsigned)rows),
                       dim3(F::NT), 0, s, a);
  return hipGetLastError();
}

template <int L, int MODE>
static hipError_t row_launch(const FsArgs &a, long rows, hipStream_t s)
{
  using F = FsCfg<L>;
  using G = FsSplitCfg<L>;
  if (a.out_re)
    hipLaunchKernelGGL((k_fs_row<L, G::P, G::C, MODE, true>), dim3((unsigned)(a.n1 / G::C), (unsigned)rows),
                       dim3(G::NT), 0, s, a);
  else
    hipLaunchKernelGGL((k_fs_row<L, F::P, F::C, MODE, false>), dim3((unsigned)(a.n1 / F::C), (unsigned)rows),
                       dim3(F::NT), 0, s, a);
  return hipGetLastError();
}

#define OVX_FS_LENGTHS(X) X(128) X(256) X(512) X(1024) X(2048)

hipError_t launch_fs_col(int mode, const FsArgs &a, long rows, hipStream_t s)
{
  switch (a.n1) {
#define X(L)                                                  \
  case L:                                                     \
    if (mode == 0) return col_launch<L, 0>(a, rows, s);       \
    if (mode == 1) return col_launch<L, 1>(a, rows, s);       \
    return col_launch<L, 2>(a, rows, s);
    OVX_FS_LENGTHS(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fs_row(int mode, const FsArgs &a, long rows, hipStream_t s)
{
  switch (a.n2) {
#define X(L)                                                  \
  case L:                                                     \
    if (mode == 0) return row_launch<L, 0>(a, rows, s);       \
    if (mode == 1) return row_launch<L, 1>(a, rows, s);       \
    return row_launch<L, 2>(a, rows, s);
    OVX_FS_LENGTHS(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

// tfull[k1 * n2 + c] = W_N^{k1 c}, N = n1 n2, from FP64 values rounded once.
__global__ void k_fs_full_twiddles(long n1, long n2, float2 *__restrict__ out)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = n1 * n2;
  if (i >= n) return;
  const long m = (i / n2) * (i % n2); // < n1 n2
  double sn, cs;
  sincospi(-2.0 * (double)m / (double)n, &sn, &cs);
  out[i] = make_float2((float)cs, (float)sn);
}

hipError_t launch_fs_full_twiddles(long n1, long n2, float2 *out, hipStream_t s)
{
  long n = n1 * n2;
  hipLaunchKernelGGL(k_fs_full_twiddles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n1, n2, out);
  return hipGetLastError();
}

hipError_t launch_fs_permute(const float2 *w, long ldw, float2 *wp, long n1, long n2, long rows, float scale,
                             hipStream_t s)
{
  long n = n1 * n2;
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)rows);
  hipLaunchKernelGGL(k_fs_permute, grid, dim3(256), 0, s, w, ldw, wp, n1, n2, scale);
  return hipGetLastError();
}

bool fs_length_supported(long l)
{
  switch (l) {
#define X(L) case L:
    OVX_FS_LENGTHS(X)
#undef X
    return true;
  default: return false;
  }
}

} // namespace ovx
