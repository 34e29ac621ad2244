// ovx_device.hpp -- device-side building blocks for the gfx950 FFT kernels.
//
// Complex values are float2 (re, im), the interleaved complex<float> layout
// OpenVSIP uses by default (src/ovxx/storage/block.hpp:27-42).  DIR is the
// transform exponent sign: -1 forward, +1 inverse
// (src/ovxx/signal/fft/util.hpp:43-51).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ovx {

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
  return make_float2(__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -a.x * b.y));
}
// multiply by the table twiddle w (stored for the forward sign); the inverse
// direction uses its conjugate.
template <int DIR>
__device__ __forceinline__ float2 twmul(float2 a, float2 w)
{
  if constexpr (DIR < 0) return cmul(a, w);
  else return cmulc(a, w);
}
// a * (DIR * i): forward multiplies by -i, inverse by +i.
template <int DIR>
__device__ __forceinline__ float2 rot(float2 a)
{
  if constexpr (DIR < 0) return make_float2(a.y, -a.x);
  else return make_float2(-a.y, a.x);
}

// ---------------------------------------------------------------- small DFTs
// In-place DFT of R values in natural order: y_k = sum_n x_n W_R^{nk},
// W_R = exp(DIR*2*pi*i/R).
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> {
  __device__ __forceinline__ static void run(float2 *) {}
};

template <int DIR> struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2 *v)
  {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR> struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2 *v)
  {
    const float c = -0.5f, s = 0.86602540378443864676f; // cos, sin(2pi/3)
    float2 a = v[0], t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    float2 m = make_float2(__builtin_fmaf(c, t.x, a.x), __builtin_fmaf(c, t.y, a.y));
    float2 r = rot<DIR>(cscale(d, s));
    v[0] = cadd(a, t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <int DIR> struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2 *v)
  {
    float2 s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    float2 s1 = cadd(v[1], v[3]), d1 = rot<DIR>(csub(v[1], v[3]));
    v[0] = cadd(s0, s1);
    v[1] = cadd(d0, d1);
    v[2] = csub(s0, s1);
    v[3] = csub(d0, d1);
  }
};

template <int DIR> struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2 *v)
  {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 x0 = v[0];
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    float2 a1 = make_float2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
    float2 a2 = make_float2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
    float2 b1 = rot<DIR>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    float2 b2 = rot<DIR>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = cadd(x0, cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

// multiply by W8^1 = sqrt(1/2)(1 + DIR i)
template <int DIR> __device__ __forceinline__ float2 mul_w8_1(float2 a)
{
  const float h = 0.70710678118654752440f;
  if constexpr (DIR < 0) return make_float2((a.x + a.y) * h, (a.y - a.x) * h);
  else return make_float2((a.x - a.y) * h, (a.y + a.x) * h);
}
// multiply by W8^3 = sqrt(1/2)(-1 + DIR i)
template <int DIR> __device__ __forceinline__ float2 mul_w8_3(float2 a)
{
  const float h = 0.70710678118654752440f;
  if constexpr (DIR < 0) return make_float2((a.y - a.x) * h, -(a.x + a.y) * h);
  else return make_float2(-(a.x + a.y) * h, (a.x - a.y) * h);
}

template <int DIR> struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2 *v)
  {
    float2 u[4], w[4];
    u[0] = cadd(v[0], v[4]); w[0] = csub(v[0], v[4]);
    u[1] = cadd(v[1], v[5]); w[1] = mul_w8_1<DIR>(csub(v[1], v[5]));
    u[2] = cadd(v[2], v[6]); w[2] = rot<DIR>(csub(v[2], v[6]));
    u[3] = cadd(v[3], v[7]); w[3] = mul_w8_3<DIR>(csub(v[3], v[7]));
    Dft<4, DIR>::run(u);
    Dft<4, DIR>::run(w);
#pragma unroll
    for (int k = 0; k < 4; ++k) { v[2 * k] = u[k]; v[2 * k + 1] = w[k]; }
  }
};

// 16 = 4 x 4: y[k1 + 4 k2] = DFT4_n( W16^{n k1} * DFT4_{n2}(x[n + 4 n2])[k1] )[k2]
template <int DIR> struct Dft<16, DIR> {
  __device__ __forceinline__ static void run(float2 *v)
  {
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float d = (float)DIR;
    float2 z[4][4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      float2 a[4] = { v[n], v[n + 4], v[n + 8], v[n + 12] };
      Dft<4, DIR>::run(a);
#pragma unroll
      for (int k = 0; k < 4; ++k) z[n][k] = a[k];
    }
    // twiddles W16^{n k1}, n,k1 in 1..3
    z[1][1] = cmul(z[1][1], make_float2(c1, d * s1));
    z[1][2] = mul_w8_1<DIR>(z[1][2]);
    z[1][3] = cmul(z[1][3], make_float2(s1, d * c1));
    z[2][1] = mul_w8_1<DIR>(z[2][1]);
    z[2][2] = rot<DIR>(z[2][2]);
    z[2][3] = mul_w8_3<DIR>(z[2][3]);
    z[3][1] = cmul(z[3][1], make_float2(s1, d * c1));
    z[3][2] = mul_w8_3<DIR>(z[3][2]);
    z[3][3] = cmul(z[3][3], make_float2(-c1, -d * s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 b[4] = { z[0][k1], z[1][k1], z[2][k1], z[3][k1] };
      Dft<4, DIR>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

// ---------------------------------------------------------------- buffer IO
// Raw buffer loads/stores: one 128-bit descriptor per base pointer, a
// per-lane 32-bit byte offset and an SGPR/immediate constant offset, so the
// unrolled accesses of a row need no 64-bit address arithmetic.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
static constexpr int kRsrcFlags = 0x00020000; // raw buffer, 32-bit data format

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, uint32_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kRsrcFlags);
}
__device__ __forceinline__ float2 bload(rsrc_t r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// non-temporal (streaming) variant for data read exactly once
__device__ __forceinline__ float2 bload_nt(rsrc_t r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}
__device__ __forceinline__ void bstore(rsrc_t r, float2 v, uint32_t voff, uint32_t soff)
{
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void bstore_nt(rsrc_t r, float2 v, uint32_t voff, uint32_t soff)
{
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 2);
}
__device__ __forceinline__ float bload1_nt(rsrc_t r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}
__device__ __forceinline__ void bstore1_nt(rsrc_t r, float v, uint32_t voff, uint32_t soff)
{
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 2);
}

// Streaming access to one line of complex values, element e at a per-lane
// base plus a constant offset.  IoC: interleaved complex<float> (one 64-bit
// access per value).  IoS: split storage, real and imaginary parts in two
// float arrays with the same layout (two 32-bit accesses, each coalesced).
struct IoC {
  rsrc_t r;
  uint32_t v; // lane base, bytes
  __device__ __forceinline__ float2 ld(uint32_t e) const { return bload_nt(r, v, e * 8u); }
  __device__ __forceinline__ void st(float2 x, uint32_t e) const { bstore_nt(r, x, v, e * 8u); }
};
struct IoS {
  rsrc_t re, im;
  uint32_t v; // lane base, bytes
  __device__ __forceinline__ float2 ld(uint32_t e) const
  {
    return make_float2(bload1_nt(re, v, e * 4u), bload1_nt(im, v, e * 4u));
  }
  __device__ __forceinline__ void st(float2 x, uint32_t e) const
  {
    bstore1_nt(re, x.x, v, e * 4u);
    bstore1_nt(im, x.y, v, e * 4u);
  }
};

} // namespace ovx
