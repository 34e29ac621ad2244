// ovx_ols_dev.hpp -- device body of one overlap-save segment (see
// ovx_ols.hip), shared by the standalone kernel and the persistent kernel.
#pragma once
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

// Segment `seg` by thread t of L/P threads, LDS region `lds` (E::LROW).
template <int L, int P, typename LP>
__device__ __forceinline__ void ols_body(const OlsArgs &a, long seg, int t, LP lds)
{
  using E = Engine<L, P>;
  constexpr int T = E::T;
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

} // namespace ovx
