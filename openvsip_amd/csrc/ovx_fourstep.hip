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
#include "ovx_fourstep.hpp"
#include "ovx_internal.hpp"

namespace ovx {

// LDS stride per line: odd number of float2 so that lanes that differ in the
// line index land on different banks.
template <int L>
struct FsLds {
  static constexpr int LROW = lds_row_elems(L) | 1;
};

__device__ __forceinline__ float2 big_twiddle(const float2 *__restrict__ tlo, const float2 *__restrict__ thi,
                                              int sbits, uint32_t m)
{
  float2 lo = tlo[m & ((1u << sbits) - 1u)];
  float2 hi = thi[m >> sbits];
  return cmul(hi, lo);
}

// ---------------------------------------------------------------- column pass
// MODE 0: forward column FFT, then * W_N^{n2 k1}      (fastconv / Fftm pass 1)
// MODE 1: inverse column FFT, then * scale             (fastconv pass 3)
// MODE 2: inverse column FFT, then * W_N^{-n2 k1}      (inverse Fftm pass 1)
template <int L, int P, int C, int MODE>
__global__ __launch_bounds__(C *(L / P)) void k_fs_col(FsArgs a)
{
  using E = Engine<L, P>;
  constexpr int T = E::T;
  constexpr int LROW = FsLds<L>::LROW;
  __shared__ float2 lds_all[C * LROW];
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const long row = blockIdx.y;
  const long n2 = (long)blockIdx.x * C + c;
  const long N2 = a.n2;
  const float2 *src = a.in + row * a.ld_in + n2;
  float2 *dst = a.out + row * a.ld_out + n2;
  const rsrc_t rtw = make_rsrc(a.tw1, (uint32_t)(E::plan.twtotal + 1) * 8u);
  float2 *lds = lds_all + c * LROW;

  float2 v[P];
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = src[(long)(t + T * q) * N2];
  if constexpr (MODE == 0) E::template fft<-1>(v, lds, t, rtw);
  else E::template fft<+1>(v, lds, t, rtw);
  if constexpr (MODE == 0 || MODE == 2) {
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const uint32_t k1 = (uint32_t)(t + T * q);
      const uint32_t m = (uint32_t)n2 * k1; // < N1*N2: no reduction needed
      float2 w = big_twiddle(a.tlo, a.thi, a.sbits, m);
      v[q] = (MODE == 0) ? cmul(v[q], w) : cmulc(v[q], w);
    }
  } else {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = cscale(v[q], a.scale);
  }
#pragma unroll
  for (int q = 0; q < P; ++q) dst[(long)(t + T * q) * N2] = v[q];
}

// ---------------------------------------------------------------- row pass
// MODE 0: fastconv middle pass (in place allowed):
//         forward row FFT, * Wp[k1][k2], inverse row FFT, * W_N^{-n2 k1}
// MODE 1: forward row FFT, transposed store X[k1 + N1 k2], * scale
// MODE 2: inverse row FFT, transposed store,               * scale
template <int L, int P, int C, int MODE>
__global__ __launch_bounds__(C *(L / P)) void k_fs_row(FsArgs a)
{
  using E = Engine<L, P>;
  constexpr int T = E::T;
  constexpr int LROW = FsLds<L>::LROW;
  __shared__ float2 lds_all[C * LROW];
  const int tid = threadIdx.x;
  const int c = tid / T; // line fastest-varying within a line: t
  const int t = tid % T;
  const long row = blockIdx.y;
  const long k1 = (long)blockIdx.x * C + c;
  const long N1 = a.n1;
  const rsrc_t rtw = make_rsrc(a.tw2, (uint32_t)(E::plan.twtotal + 1) * 8u);
  float2 *lds = lds_all + c * LROW;
  const float2 *src = a.in + row * a.ld_in + k1 * L;

  float2 v[P];
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = src[t + T * q];
  if constexpr (MODE == 0) {
    E::template fft<-1>(v, lds, t, rtw);
    const float2 *wp = a.wp + row * a.ldwp + k1 * L;
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = cmul(v[q], wp[t + T * q]);
    E::template fft<+1>(v, lds, t, rtw);
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const uint32_t n2 = (uint32_t)(t + T * q);
      const uint32_t m = n2 * (uint32_t)k1;
      v[q] = cmulc(v[q], big_twiddle(a.tlo, a.thi, a.sbits, m));
    }
    float2 *dst = a.out + row * a.ld_out + k1 * L;
#pragma unroll
    for (int q = 0; q < P; ++q) dst[t + T * q] = v[q];
  } else {
    if constexpr (MODE == 1) E::template fft<-1>(v, lds, t, rtw);
    else E::template fft<+1>(v, lds, t, rtw);
    if (a.scale != 1.0f) {
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = cscale(v[q], a.scale);
    }
    // transpose through LDS: element k2 of line c -> out[k1_0 + c + N1 k2]
#pragma unroll
    for (int q = 0; q < P; ++q) lds[t + T * q] = v[q];
    __syncthreads();
    const int c2 = tid % C;
    const int t2 = tid / C;
    const float2 *lsrc = lds_all + c2 * LROW;
    float2 *dst = a.out + row * a.ld_out + (long)blockIdx.x * C + c2;
#pragma unroll
    for (int q = 0; q < P; ++q) dst[(long)(t2 + T * q) * N1] = lsrc[t2 + T * q];
  }
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
template <int L>
struct FsCfg {
  static constexpr int P = 16;
  static constexpr int C = (L / P) * 16 <= 1024 ? 16 : 1024 / (L / P);
  static constexpr int NT = C * (L / P);
};

template <int L, int MODE>
static hipError_t col_launch(const FsArgs &a, long rows, hipStream_t s)
{
  using F = FsCfg<L>;
  dim3 grid((unsigned)(a.n2 / F::C), (unsigned)rows);
  hipLaunchKernelGGL((k_fs_col<L, F::P, F::C, MODE>), grid, dim3(F::NT), 0, s, a);
  return hipGetLastError();
}

template <int L, int MODE>
static hipError_t row_launch(const FsArgs &a, long rows, hipStream_t s)
{
  using F = FsCfg<L>;
  dim3 grid((unsigned)(a.n1 / F::C), (unsigned)rows);
  hipLaunchKernelGGL((k_fs_row<L, F::P, F::C, MODE>), grid, dim3(F::NT), 0, s, a);
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
