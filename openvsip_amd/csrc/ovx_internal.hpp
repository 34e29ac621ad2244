// ovx_internal.hpp -- internal (C++) interface between the kernel translation
// units and the C-ABI layer.  Not installed; the public surface is
// include/ovxhip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ovx_engine.hpp"

namespace ovx {

// One compiled LDS-resident row-transform configuration.
struct RowConfig {
  int n = 0;            // transform length
  int p = 0;            // values per thread
  int rows_per_block = 0;
  int threads = 0;      // block size
  size_t lds_bytes = 0; // static LDS per block
  RadixPlan plan{};
  int variant = 0;      // index among the configurations compiled for n
};

// Number of compiled configurations for length n (0 = not LDS-resident).
int row_variants(int n);
bool row_config(int n, int variant, RowConfig *out);
// Largest length with an LDS-resident configuration.
constexpr int kMaxRowN = 16384;

// Transform of `rows` lines of length cfg.n.  Element e of line r sits at
// in + r*ld_in + e (contiguous lines, complex<float>); scale multiplies the
// result (skipped when 1).
hipError_t launch_row_fft(const RowConfig &cfg, int dir, const float2 *in, long ld_in,
                          float2 *out, long ld_out, long rows, const float2 *tw,
                          float scale, hipStream_t s);

// Fused fast convolution of `rows` lines:
//   y_r = scale * IFFT( w_r (.) FFT(x_r) ),  w_r = w + r*ldw (ldw = 0: one vector).
hipError_t launch_row_fastconv(const RowConfig &cfg, const float2 *x, long ldx,
                               const float2 *w, long ldw, float2 *y, long ldy, long rows,
                               const float2 *tw, float scale, hipStream_t s);

// The same two on split storage: real and imaginary parts of the lines in
// two float arrays with the same strides (ld in floats); w stays interleaved.
hipError_t launch_row_fft_split(const RowConfig &cfg, int dir, const float *in_re, const float *in_im, long ld_in,
                                float *out_re, float *out_im, long ld_out, long rows, const float2 *tw, float scale,
                                hipStream_t s);
hipError_t launch_row_fastconv_split(const RowConfig &cfg, const float *xr, const float *xi, long ldx,
                                     const float2 *w, long ldw, float *yr, float *yi, long ldy, long rows,
                                     const float2 *tw, float scale, hipStream_t s);

// Forward twiddle table of a radix plan: tw[twoff_s + (r-1)*Ns + k] =
// exp(-2 pi i r k / (Ns R_s)) evaluated in FP64 on the device, rounded to FP32.
hipError_t launch_twiddles(const RadixPlan &plan, float2 *tw, hipStream_t s);

// out[i] = scale * in[i], i < n.
hipError_t launch_scale_copy(const float2 *in, float2 *out, long n, float scale, hipStream_t s);

// Table of exp(-2 pi i j / n), j in [0, count), FP64-evaluated, FP32-rounded.
hipError_t launch_unit_roots(long n, long count, float2 *out, hipStream_t s);

// Split-complex <-> interleaved, `lines` lines of n values: line l of the
// float arrays re/im starts at l*ld_s, of the complex array at l*ld_c.
hipError_t launch_interleave(const float *re, const float *im, long ld_s, float2 *out, long ld_c, long lines,
                             long n, hipStream_t s);
hipError_t launch_deinterleave(const float2 *in, long ld_c, float *re, float *im, long ld_s, long lines, long n,
                               hipStream_t s);

} // namespace ovx
