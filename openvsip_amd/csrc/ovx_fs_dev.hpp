// ovx_fs_dev.hpp -- device bodies of the two-pass four-step kernels (see
// ovx_fourstep.hip for the algorithm), shared by the standalone kernels and
// the persistent batch kernel.
#pragma once
#include "ovx_fourstep.hpp"
#include "ovx_internal.hpp"

namespace ovx {

// LDS stride per line: odd number of float2 so that lanes that differ in the
// line index land on different banks.
template <int L>
struct FsLds {
  static constexpr int LROW = LdsLayout<L, 16>::elems | 1;
};

__device__ __forceinline__ float2 big_twiddle(const float2 *__restrict__ tlo, const float2 *__restrict__ thi,
                                              int sbits, uint32_t m)
{
  float2 lo = tlo[m & ((1u << sbits) - 1u)];
  float2 hi = thi[m >> sbits];
  return cmul(hi, lo);
}

// ---------------------------------------------------------------- column pass
// MODE 0: forward column FFT, then * W_N^{n2 k1}      (fastconv / Fftm pass 1)
// MODE 1: inverse column FFT, then * scale             (fastconv pass 3)
// MODE 2: inverse column FFT, then * W_N^{-n2 k1}      (inverse Fftm pass 1)
// One group of C columns of one row: `group` selects columns
// [group*C, group*C + C), `tid` is the thread within the group's C*(L/P)
// threads, `lds_all` the group's C*LROW LDS elements.
// The passes' global reads and writes at element offset o from a call's
// base: interleaved lines (a.in, a.out) or, when SPLIT, separate real and
// imaginary arrays (a.in_re / a.in_im, a.out_re / a.out_im).
template <bool SPLIT>
__device__ __forceinline__ float2 fs_load(const FsArgs &a, long o)
{
  if constexpr (SPLIT) return make_float2(a.in_re[o], a.in_im[o]);
  else return a.in[o];
}
template <bool SPLIT>
__device__ __forceinline__ void fs_store(const FsArgs &a, long o, float2 v)
{
  if constexpr (SPLIT) {
    a.out_re[o] = v.x;
    a.out_im[o] = v.y;
  } else {
    a.out[o] = v;
  }
}

template <int L, int P, int C, int MODE, bool SIN = false, bool SOUT = false, typename LP>
__device__ __forceinline__ void fs_col_body(const FsArgs &a, long row, long group, int tid, LP lds_all)
{
  using E = Engine<L, P>;
  constexpr int T = E::T;
  constexpr int LROW = FsLds<L>::LROW;
  const int c = tid % C;
  const int t = tid / C;
  const long n2 = group * C + c;
  const long N2 = a.n2;
  const long ib = row * a.ld_in + n2, ob = row * a.ld_out + n2;
  const rsrc_t rtw = make_rsrc(a.tw1, (uint32_t)(E::plan.twtotal + 1) * 8u);
  LP lds = lds_all + c * LROW;

  float2 v[P];
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = fs_load<SIN>(a, ib + (long)(t + T * q) * N2);
  if constexpr (MODE == 0) E::template fft<-1>(v, lds, t, rtw);
  else E::template fft<+1>(v, lds, t, rtw);
  if constexpr (MODE == 0 || MODE == 2) {
    if (a.tfull) { // lanes of consecutive columns read consecutive entries
      const float2 *tf = a.tfull + n2;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        float2 w = tf[(long)(t + T * q) * N2];
        v[q] = (MODE == 0) ? cmul(v[q], w) : cmulc(v[q], w);
      }
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const uint32_t k1 = (uint32_t)(t + T * q);
        const uint32_t m = (uint32_t)n2 * k1; // < N1*N2: no reduction needed
        float2 w = big_twiddle(a.tlo, a.thi, a.sbits, m);
        v[q] = (MODE == 0) ? cmul(v[q], w) : cmulc(v[q], w);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = cscale(v[q], a.scale);
  }
#pragma unroll
  for (int q = 0; q < P; ++q) fs_store<SOUT>(a, ob + (long)(t + T * q) * N2, v[q]);
}

// ---------------------------------------------------------------- row pass
// MODE 0: fastconv middle pass (in place allowed):
//         forward row FFT, * Wp[k1][k2], inverse row FFT, * W_N^{-n2 k1}
// MODE 1: forward row FFT, transposed store X[k1 + N1 k2], * scale
// MODE 2: inverse row FFT, transposed store,               * scale
// One group of C lines k1 in [group*C, group*C + C) of one row.
template <int L, int P, int C, int MODE, bool SOUT = false, typename LP>
__device__ __forceinline__ void fs_row_body(const FsArgs &a, long row, long group, int tid, LP lds_all)
{
  // a line's L/P threads are consecutive; when they fit one wavefront its
  // exchanges need no workgroup barrier
  using E = Engine<L, P, (L / P <= 64)>;
  constexpr int T = E::T;
  constexpr int LROW = FsLds<L>::LROW;
  const int c = tid / T; // consecutive threads walk along a line
  const int t = tid % T;
  const long k1 = group * C + c;
  const long N1 = a.n1;
  const rsrc_t rtw = make_rsrc(a.tw2, (uint32_t)(E::plan.twtotal + 1) * 8u);
  LP lds = lds_all + c * LROW;
  const float2 *src = a.in + row * a.ld_in + k1 * L;

  float2 v[P];
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = src[t + T * q];
  if constexpr (MODE == 0) {
    E::template fft<-1>(v, lds, t, rtw);
    const float2 *wp = a.wp + row * a.ldwp + k1 * L;
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = cmul(v[q], wp[t + T * q]);
    E::template fft<+1>(v, lds, t, rtw);
    if (a.tfull) {
      const float2 *tf = a.tfull + k1 * L;
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = cmulc(v[q], tf[t + T * q]);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const uint32_t n2 = (uint32_t)(t + T * q);
        const uint32_t m = n2 * (uint32_t)k1;
        v[q] = cmulc(v[q], big_twiddle(a.tlo, a.thi, a.sbits, m));
      }
    }
    const long ob = row * a.ld_out + k1 * L;
#pragma unroll
    for (int q = 0; q < P; ++q) fs_store<SOUT>(a, ob + t + T * q, v[q]);
  } else {
    if constexpr (MODE == 1) E::template fft<-1>(v, lds, t, rtw);
    else E::template fft<+1>(v, lds, t, rtw);
    if (a.scale != 1.0f) {
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = cscale(v[q], a.scale);
    }
    // transpose through LDS: element k2 of line c -> out[k1_0 + c + N1 k2]
#pragma unroll
    for (int q = 0; q < P; ++q) lds_st(lds, t + T * q, v[q]);
    __syncthreads();
    const int c2 = tid % C;
    const int t2 = tid / C;
    LP lsrc = lds_all + c2 * LROW;
    const long ob = row * a.ld_out + group * C + c2;
#pragma unroll
    for (int q = 0; q < P; ++q) fs_store<SOUT>(a, ob + (long)(t2 + T * q) * N1, lds_ld(lsrc, t2 + T * q));
  }
}

// Threads and LDS of one group (C lines of length L, P values per thread).
template <int L>
struct FsCfg {
  static constexpr int P = 16;
  static constexpr int C = (L / P) * 16 <= 1024 ? 16 : 1024 / (L / P);
  static constexpr int NT = C * (L / P);
  static constexpr int LDS = C * FsLds<L>::LROW;
};

} // namespace ovx
