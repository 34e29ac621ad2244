// ovx_vmmul.hip -- vector-matrix element-wise multiply on the device:
//   vmmul<0>(v, m)(r, c) = v(c) * m(r, c)      (v runs along each row)
//   vmmul<1>(v, m)(r, c) = v(r) * m(r, c)      (v runs along each column)
// src/vsip/impl/vmmul.hpp:50-128 and src/ovxx/expr/vmmul.hpp:53-59 (the
// block's get()); the reference reduces it to a loop of vector multiplies
// over rows or columns (vmmul.hpp:79-101).  Here it is one pass: every
// element is read once and written once (HBM-bound, 8-24 bytes per element).
//
// Operands are complex<float> (interleaved) or float; the result is
// complex<float> when either operand is complex, else float (the
// reference's Promotion<T1, T2>).  Threads run along whichever matrix axis
// has unit stride, so that the loads and stores of a wavefront are
// contiguous for row-major and column-major operands alike.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ovxhip.h"
#include "ovx_plan.hpp"

namespace ovx {

namespace {

template <bool C>
struct elem;
template <>
struct elem<true> {
  typedef float2 type;
};
template <>
struct elem<false> {
  typedef float type;
};

__device__ inline float2 as_c(float2 a) { return a; }
__device__ inline float2 as_c(float a) { return make_float2(a, 0.0f); }

// complex product as the reference's complex<T> operator* computes it
// (re = ar*br - ai*bi, im = ar*bi + ai*br), then the real factor alpha
__device__ inline float2 cmul(float2 a, float2 b, float alpha)
{
  float re = a.x * b.x - a.y * b.y;
  float im = a.x * b.y + a.y * b.x;
  return make_float2(alpha * re, alpha * im);
}
__device__ inline float cmul(float a, float b, float alpha) { return alpha * (a * b); }

// (outer, inner) is (row, column) or (column, row): inner has the unit
// stride.  v is indexed by the inner index when v_inner, else the outer.
struct VmmulArgs {
  const void *v;
  long vs;
  const void *m;
  long m_outer, m_inner;
  void *out;
  long o_outer, o_inner;
  long outer, inner;
  int v_inner;
  float alpha;
};

template <bool VC, bool MC>
__global__ __launch_bounds__(256) void k_vmmul(VmmulArgs a)
{
  typedef typename elem<VC>::type vt;
  typedef typename elem<MC>::type mt;
  typedef typename elem<VC || MC>::type ot;
  const vt *v = static_cast<const vt *>(a.v);
  const mt *m = static_cast<const mt *>(a.m);
  ot *out = static_cast<ot *>(a.out);
  for (long o = blockIdx.y; o < a.outer; o += gridDim.y) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.inner; i += (long)gridDim.x * blockDim.x) {
      vt vv = v[(a.v_inner ? i : o) * a.vs];
      mt mm = m[o * a.m_outer + i * a.m_inner];
      if constexpr (VC || MC) out[o * a.o_outer + i * a.o_inner] = cmul(as_c(vv), as_c(mm), a.alpha);
      else out[o * a.o_outer + i * a.o_inner] = cmul(vv, mm, a.alpha);
    }
  }
}

} // namespace

} // namespace ovx

using namespace ovx;

extern "C" int ovx_vmmul(int dim, int v_complex, const void *v, long v_stride, int m_complex, const void *m,
                         long m_s0, long m_s1, void *out, long out_s0, long out_s1, long rows, long cols,
                         float alpha, void *stream)
{
  if (dim != 0 && dim != 1) return fail(OVX_ERR_INVALID, "vmmul: dimension %d is not 0 or 1", dim);
  if (rows < 0 || cols < 0) return fail(OVX_ERR_INVALID, "vmmul: negative size %ld x %ld", rows, cols);
  if (rows == 0 || cols == 0) return OVX_OK;
  if (!v || !m || !out) return fail(OVX_ERR_INVALID, "vmmul: null pointer");
  VmmulArgs a{};
  a.v = v;
  a.vs = v_stride;
  a.m = m;
  a.out = out;
  a.alpha = alpha;
  // threads along the axis on which the output (else the matrix) is
  // contiguous
  bool col_major = (out_s0 == 1 && out_s1 != 1) || (out_s0 != 1 && out_s1 != 1 && m_s0 == 1 && m_s1 != 1);
  if (!col_major) {
    a.outer = rows;
    a.inner = cols;
    a.m_outer = m_s0;
    a.m_inner = m_s1;
    a.o_outer = out_s0;
    a.o_inner = out_s1;
    a.v_inner = dim == 0;
  } else {
    a.outer = cols;
    a.inner = rows;
    a.m_outer = m_s1;
    a.m_inner = m_s0;
    a.o_outer = out_s1;
    a.o_inner = out_s0;
    a.v_inner = dim == 1;
  }
  const int threads = 256;
  long bx = std::min<long>((a.inner + threads - 1) / threads, 1024);
  long by = std::min<long>(a.outer, 65535);
  // enough workgroups to fill the chip when the lines are short
  dim3 grid((unsigned)bx, (unsigned)by);
  hipStream_t s = static_cast<hipStream_t>(stream);
  bool vc = v_complex != 0, mc = m_complex != 0;
  if (vc && mc) k_vmmul<true, true><<<grid, threads, 0, s>>>(a);
  else if (vc) k_vmmul<true, false><<<grid, threads, 0, s>>>(a);
  else if (mc) k_vmmul<false, true><<<grid, threads, 0, s>>>(a);
  else k_vmmul<false, false><<<grid, threads, 0, s>>>(a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? OVX_OK : hip_fail(e, "vmmul launch");
}
