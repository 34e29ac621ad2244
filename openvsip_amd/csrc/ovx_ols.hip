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
#include "ovx_internal.hpp"
#include "ovx_ols.hpp"

namespace ovx {

// unbiased divisor of output n (corr.hpp:115-124, 163-178, 204-206)
__device__ __forceinline__ float corr_divisor(int support, long n, long M, long N)
{
  if (support == 0) {
    if (n < M - 1) return (float)(n + 1);
    if (n >= N) return (float)(N + M - 1 - n);
    return (float)M;
  } else if (support == 1) {
    if (n < M / 2) return (float)(n + (M + 1) / 2);
    if (n >= N - M / 2) return (float)(N - 1 + (M + 1) / 2 - n);
    return (float)M;
  }
  return (float)M;
}

template <int L, int P>
__global__ __launch_bounds__(L / P) void k_ols(OlsArgs a)
{
  using E = Engine<L, P>;
  constexpr int T = E::T;
  __shared__ float2 lds[E::LROW];
  const int t = threadIdx.x;
  const long seg = blockIdx.x;
  const long S = L - (a.M - 1);
  const long i0 = a.c_first + seg * a.c_step;
  const rsrc_t rtw = make_rsrc(a.tw, (uint32_t)(E::plan.twtotal + 1) * 8u);
  const rsrc_t rh = make_rsrc(a.H, (uint32_t)L * 8u);

  float2 v[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const long idx = i0 - (a.M - 1) + t + (long)T * q;
    float2 g = make_float2(0.f, 0.f);
    if (idx >= 0 && idx < a.N) g = a.in[idx];
    if (a.conj_input) g.y = -g.y;
    v[q] = g;
  }
  E::template fft<-1>(v, lds, t, rtw);
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = cmul(v[q], cscale(bload(rh, (uint32_t)t * 8u, (uint32_t)(T * q) * 8u), a.scale));
  E::template fft<+1>(v, lds, t, rtw);
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const long u = t + (long)T * q;
    if (u < a.M - 1 || u - (a.M - 1) >= S) continue;
    const long i = i0 + u - (a.M - 1);     // index into c
    long rel = i - a.out_off;               // n*D
    if (rel < 0) continue;
    if (a.decimation > 1 && (rel % a.decimation) != 0) continue;
    const long n = a.decimation > 1 ? rel / a.decimation : rel;
    if (n >= a.P) continue;
    float2 r = v[q];
    if (a.unbiased) {
      float d = corr_divisor(a.support, n, a.M, a.N);
      r.x /= d;
      r.y /= d;
    }
    a.out[n] = r;
  }
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
