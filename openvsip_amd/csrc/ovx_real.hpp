// ovx_real.hpp -- launch interface of the real <-> complex Fftm passes.
#pragma once
#include "ovx_internal.hpp"

namespace ovx {

// In place on `lines` lines of h + 1 complex values holding FFT_h of the
// packed real line in [0, h): produce the n = 2h point real-input spectrum
// X[0 .. h], times `scale`.  wt[k] = exp(-2 pi i k / n), k <= h/2.
hipError_t launch_r2c_post(float2 *y, long ldy, long lines, long h, const float2 *wt, float scale, hipStream_t s);
// From `lines` half spectra X[0 .. h] build the h-point complex lines whose
// inverse transform is the packed real signal (out of place).
hipError_t launch_c2r_pre(const float2 *x, long ldx, float2 *z, long ldz, long lines, long h, const float2 *wt,
                          hipStream_t s);
// The same two with the half spectra in split storage (separate real and
// imaginary arrays, line stride in floats): the post-pass reads FFT_h from z
// and writes X there; the pre-pass reads X from there.
hipError_t launch_r2c_post_split(const float2 *z, long ldz, float *yr, float *yi, long ldy, long lines, long h,
                                 const float2 *wt, float scale, hipStream_t s);
hipError_t launch_c2r_pre_split(const float *xr, const float *xi, long ldx, float2 *z, long ldz, long lines, long h,
                                const float2 *wt, hipStream_t s);

} // namespace ovx
