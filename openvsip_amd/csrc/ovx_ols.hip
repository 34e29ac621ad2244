// ovx_ols.hip -- FFT-based 1-D correlation and convolution by overlap-save
// on the LDS-resident engine.
//
// Both operations reduce to one linear convolution c = h * g:
//   correlation (src/ovxx/signal/corr.hpp:91-215)
//       out[n] = sum_k ref[k] conj(in[n + k - shift])
//              = c[n + M-1-shift],  h[j] = ref[M-1-j],  g = conj(in)
//       shift = M-1 (full), M/2 (same), 0 (min);
//   convolution (src/ovxx/signal/conv.hpp:232-330)
//       out[n] = sum_k kernel[k] in[n D + off - k] = c[n D + off],
//       h = kernel, g = in, off = 0 (full), M/2 (same), M-1 (min).
// Segment s of length L covers c[i0, i0 + S), S = L - (M-1): the kernel
// gathers g[i0 - (M-1) + u] (zero outside [0, N)), runs forward FFT,
// multiplies by H = FFT_L(h) / L, inverse FFT, and scatters the valid outputs
// u >= M-1, applying the unbiased divisors (correlation) or the decimation
// (convolution).  One block per segment; every segment is one "row".
#include "ovx_ols_dev.hpp"

namespace ovx {

template <int L, int P>
__global__ __launch_bounds__(L / P) void k_ols(OlsArgs a)
{
  __shared__ float2 lds[Engine<L, P>::LROW];
  ols_body<L, P>(a, blockIdx.x, threadIdx.x, lds);
}

// h_pad[u] = reversed (corr) or plain (conv) kernel, zero-padded to L
__global__ void k_ols_prep(const float2 *__restrict__ h, long M, int reverse, float2 *__restrict__ out, long L)
{
  long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= L) return;
  float2 v = make_float2(0.f, 0.f);
  if (u < M) v = h[reverse ? (M - 1 - u) : u];
  out[u] = v;
}

template <int L>
static hipError_t do_ols(const OlsArgs &a, long segments, hipStream_t s)
{
  constexpr int P = L >= 16 ? 16 : L;
  hipLaunchKernelGGL((k_ols<L, P>), dim3((unsigned)segments), dim3(L / P), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ols(int L, const OlsArgs &a, long segments, hipStream_t s)
{
  if (segments <= 0) return hipSuccess;
  switch (L) {
  case 64: return do_ols<64>(a, segments, s);
  case 128: return do_ols<128>(a, segments, s);
  case 256: return do_ols<256>(a, segments, s);
  case 512: return do_ols<512>(a, segments, s);
  case 1024: return do_ols<1024>(a, segments, s);
  case 2048: return do_ols<2048>(a, segments, s);
  case 4096: return do_ols<4096>(a, segments, s);
  case 8192: return do_ols<8192>(a, segments, s);
  case 16384: return do_ols<16384>(a, segments, s);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_ols_prep(const float2 *h, long M, int reverse, float2 *out, long L, hipStream_t s)
{
  long blocks = (L + 255) / 256;
  hipLaunchKernelGGL(k_ols_prep, dim3((unsigned)blocks), dim3(256), 0, s, h, M, reverse, out, L);
  return hipGetLastError();
}

} // namespace ovx
