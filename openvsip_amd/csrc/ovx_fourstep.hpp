// ovx_fourstep.hpp -- launch interface of the two-pass four-step kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace ovx {

struct FsArgs {
  const float2 *in;
  long ld_in;
  float2 *out;
  long ld_out;
  long n1, n2, nbig;       // N = n1 * n2
  const float2 *tw1, *tw2; // line twiddles of the n1 and n2 plans
  const float2 *tlo, *thi; // W_N^m = thi[m >> sbits] * tlo[m & (2^sbits - 1)]
  int sbits;
  const float2 *tfull;     // W_N^{k1 n2} at [k1 * n2count + n2], or null (two-level)
  const float2 *wp;        // fastconv: permuted, scaled spectrum [k1][k2]
  long ldwp;               // 0: one spectrum for every row
  float scale;
  // split storage for a call's first read / last write (separate real and
  // imaginary arrays at the strides ld_in / ld_out); null: in / out are used
  const float *in_re = nullptr, *in_im = nullptr;
  float *out_re = nullptr, *out_im = nullptr;
};

hipError_t launch_fs_col(int mode, const FsArgs &a, long rows, hipStream_t s);
hipError_t launch_fs_row(int mode, const FsArgs &a, long rows, hipStream_t s);
hipError_t launch_fs_permute(const float2 *w, long ldw, float2 *wp, long n1, long n2, long rows, float scale,
                             hipStream_t s);
bool fs_length_supported(long l);
// Longest transform whose full twiddle table (n entries) the plans keep.
constexpr long kFsFullTwiddleMax = 1l << 18;
hipError_t launch_fs_full_twiddles(long n1, long n2, float2 *out, hipStream_t s);

} // namespace ovx
