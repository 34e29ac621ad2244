// ovx_ols.hpp -- overlap-save launch interface (correlation / convolution).
#pragma once
#include <hip/hip_runtime.h>

namespace ovx {

struct OlsArgs {
  const float2 *in;  // input signal, length N
  long N;            // input length
  long M;            // kernel length
  const float2 *H;   // FFT_L of the padded kernel
  const float2 *tw;  // twiddles of the length-L plan
  float scale;       // 1/L
  int conj_input;    // correlation: conj(in)
  int unbiased;      // correlation: per-position divisors
  int support;       // 0 full, 1 same, 2 min
  long decimation;   // convolution: D (1 for correlation)
  long out_off;      // out[n] = c[n*D + out_off]
  long c_first;      // first c index of segment 0
  long c_step;       // c indices per segment (S = L - (M-1))
  long P;            // output length
  float2 *out;
};

hipError_t launch_ols(int L, const OlsArgs &a, long segments, hipStream_t s);
hipError_t launch_ols_prep(const float2 *h, long M, int reverse, float2 *out, long L, hipStream_t s);

} // namespace ovx
