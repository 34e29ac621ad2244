// ovx_row_dev.hpp -- device body of the register-resident Stockham fused
// fastconv (any compiled N <= 16384), shared by the standalone kernel and the
// persistent batch kernel.
#pragma once
#include "ovx_internal.hpp"

namespace ovx {

// Rows row0 + lrow of the job, thread t of the row's T = N/P threads, LDS row
// region `lds` (Engine LROW elements).  Every thread of the workgroup calls
// it; rows at or beyond `rows` compute on zeros and store nothing.  ix / iy
// access the thread's input and output line (interleaved IoC or split IoS),
// element e of the thread at ld(e) / st(v, e).
template <int N, int P, typename IX, typename IY, typename LP>
__device__ __forceinline__ void row_fastconv_io(const IX &ix, const IY &iy, const float2 *__restrict__ w, long ldw,
                                                long rows, const float2 *__restrict__ tw, float scale, long row0,
                                                int lrow, int t, LP lds)
{
  using E = Engine<N, P, (N / P <= 64)>;
  constexpr int T = E::T;
  const long row = row0 + lrow;
  const bool valid = row < rows;
  const rsrc_t rw = make_rsrc(w + row0 * ldw, 0xffffffffu);
  const rsrc_t rtw = make_rsrc(tw, (uint32_t)(E::plan.twtotal + 1) * 8u);
  const uint32_t vw = (uint32_t)(lrow * ldw + t) * 8u;

  float2 v[P];
  if (valid) {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = ix.ld((uint32_t)(T * q));
  } else {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = make_float2(0.f, 0.f);
  }
  E::template fft<-1>(v, lds, t, rtw);
  // multiply by the replica spectrum; slot q holds frequency t + T q
#pragma unroll
  for (int q = 0; q < P; ++q) {
    float2 ww = bload(rw, vw, (uint32_t)(T * q) * 8u);
    v[q] = cmul(v[q], ww);
  }
  if (scale != 1.0f) {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = cscale(v[q], scale);
  }
  E::template fft<+1>(v, lds, t, rtw);
  if (valid) {
#pragma unroll
    for (int q = 0; q < P; ++q) iy.st(v[q], (uint32_t)(T * q));
  }
}

template <int N, int P, typename LP>
__device__ __forceinline__ void row_fastconv_body(const float2 *__restrict__ x, long ldx,
                                                  const float2 *__restrict__ w, long ldw,
                                                  float2 *__restrict__ y, long ldy, long rows,
                                                  const float2 *__restrict__ tw, float scale, long row0, int lrow,
                                                  int t, LP lds)
{
  row_fastconv_io<N, P>(IoC{make_rsrc(x + row0 * ldx, 0xffffffffu), (uint32_t)(lrow * ldx + t) * 8u},
                        IoC{make_rsrc(y + row0 * ldy, 0xffffffffu), (uint32_t)(lrow * ldy + t) * 8u}, w, ldw,
                        rows, tw, scale, row0, lrow, t, lds);
}

// Forward (DIR = -1) or inverse (+1) transform of rows row0 + lrow, times
// scale (skipped when 1); same geometry contract as row_fastconv_io.
template <int N, int P, int DIR, typename IX, typename IY, typename LP>
__device__ __forceinline__ void row_fft_io(const IX &ix, const IY &iy, long rows, const float2 *__restrict__ tw,
                                           float scale, long row0, int lrow, int t, LP lds)
{
  using E = Engine<N, P, (N / P <= 64)>;
  constexpr int T = E::T;
  const long row = row0 + lrow;
  const bool valid = row < rows;
  const rsrc_t rtw = make_rsrc(tw, (uint32_t)(E::plan.twtotal + 1) * 8u);
  float2 v[P];
  if (valid) {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = ix.ld((uint32_t)(T * q));
  } else {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = make_float2(0.f, 0.f);
  }
  E::template fft<DIR>(v, lds, t, rtw);
  if (scale != 1.0f) {
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = cscale(v[q], scale);
  }
  if (valid) {
#pragma unroll
    for (int q = 0; q < P; ++q) iy.st(v[q], (uint32_t)(T * q));
  }
}

template <int N, int P, int DIR, typename LP>
__device__ __forceinline__ void row_fft_body(const float2 *__restrict__ in, long ld_in, float2 *__restrict__ out,
                                             long ld_out, long rows, const float2 *__restrict__ tw, float scale,
                                             long row0, int lrow, int t, LP lds)
{
  row_fft_io<N, P, DIR>(IoC{make_rsrc(in + row0 * ld_in, 0xffffffffu), (uint32_t)(lrow * ld_in + t) * 8u},
                        IoC{make_rsrc(out + row0 * ld_out, 0xffffffffu), (uint32_t)(lrow * ld_out + t) * 8u}, rows,
                        tw, scale, row0, lrow, t, lds);
}

} // namespace ovx
