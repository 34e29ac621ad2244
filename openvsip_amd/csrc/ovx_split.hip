// ovx_split.hip -- moves between split-complex storage (real and imaginary
// parts in two float arrays, OpenVSIP's split_complex format) and the
// interleaved layout the transform kernels read.  A pure move of floats:
// exact, so a split-storage transform equals the interleaved one bit for bit.
#include <hip/hip_runtime.h>

#include "ovx_internal.hpp"

namespace ovx {

namespace {

constexpr int kThreads = 256;
constexpr unsigned kMaxGridY = 65535;

// lines x n values; line l of the split arrays starts at l * ld_s floats, of
// the interleaved array at l * ld_c complex values.  Block (bx, by) covers
// elements [bx * 256, bx * 256 + 256) of lines by, by + gridDim.y, ...
__global__ __launch_bounds__(kThreads) void k_interleave(const float *__restrict__ re,
                                                        const float *__restrict__ im, long ld_s,
                                                        float2 *__restrict__ out, long ld_c, long lines, long n)
{
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  for (long l = blockIdx.y; l < lines; l += gridDim.y)
    out[l * ld_c + i] = make_float2(re[l * ld_s + i], im[l * ld_s + i]);
}

__global__ __launch_bounds__(kThreads) void k_deinterleave(const float2 *__restrict__ in, long ld_c,
                                                          float *__restrict__ re, float *__restrict__ im,
                                                          long ld_s, long lines, long n)
{
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  for (long l = blockIdx.y; l < lines; l += gridDim.y) {
    const float2 v = in[l * ld_c + i];
    re[l * ld_s + i] = v.x;
    im[l * ld_s + i] = v.y;
  }
}

dim3 grid(long lines, long n)
{
  return dim3((unsigned)((n + kThreads - 1) / kThreads), (unsigned)(lines < (long)kMaxGridY ? lines : kMaxGridY));
}

} // namespace

hipError_t launch_interleave(const float *re, const float *im, long ld_s, float2 *out, long ld_c, long lines,
                             long n, hipStream_t s)
{
  if (lines <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_interleave, grid(lines, n), dim3(kThreads), 0, s, re, im, ld_s, out, ld_c, lines, n);
  return hipGetLastError();
}

hipError_t launch_deinterleave(const float2 *in, long ld_c, float *re, float *im, long ld_s, long lines, long n,
                               hipStream_t s)
{
  if (lines <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_deinterleave, grid(lines, n), dim3(kThreads), 0, s, in, ld_c, re, im, ld_s, lines, n);
  return hipGetLastError();
}

} // namespace ovx
