// ovx_bf16.hpp -- launch interface of the BF16x3 MFMA precision variant.
#pragma once
#include <hip/hip_runtime.h>

namespace ovx {

// Fused fastconv of `rows` rows of length bf16x3_length() with the radix-16
// stages on the MFMA cores (bf16 hi/lo split products, FP32 accumulation).
// w: natural-order spectrum (one vector), multiplied by `scale`; tw: the
// twiddle table of the register-Stockham configuration (16384, 16 per thread);
// u256: exp(-2 pi i m / 256), m < 256.
hipError_t launch_bf16x3_fastconv(const float2 *x, long ldx, const float2 *w, float2 *y, long ldy, long rows,
                                  const float2 *tw, const float2 *u256, float scale, hipStream_t s);
int bf16x3_length();
int bf16x3_twiddle_count();

} // namespace ovx
