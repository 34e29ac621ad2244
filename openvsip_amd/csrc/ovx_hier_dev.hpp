// ovx_hier_dev.hpp -- device bodies of the hierarchical transforms for
// N = 1024 * {1, 2, 4, 8, 16}; see ovx_hier.hip for the algorithm.
#pragma once
#include "ovx_hier.hpp"
#include "ovx_internal.hpp"

namespace ovx {

template <int N>
struct Hier {
  static constexpr int NW = 1024;        // points per wave transform
  static constexpr int R0 = N / NW;      // outer radix = waves per row
  static constexpr int TR = 64 * R0;     // threads per row
  static constexpr int M0 = 16 / R0;     // outer butterflies per thread
  using WE = Engine<NW, 16, true>;
  static constexpr int LSL = WE::LROW;   // LDS slice stride
  static constexpr int LROW = R0 * LSL;
  static constexpr int B = TR >= 256 ? 1 : 256 / TR; // rows per block
  static constexpr int NT = TR * B;
  // slot of row position pos (slice pos / NW, position pos % NW in it)
  __device__ static constexpr int lidx(int pos) { return (pos / NW) * LSL + WE::Layout::idx(pos % NW); }
  static_assert(R0 * NW == N && R0 >= 1 && R0 <= 16, "N must be 1024 * {1,2,4,8,16}");
};

// Row accessors of the input and output lines of rows row0 + lrow, thread t:
// interleaved (a.x, a.y) or split storage (a.xr/a.xi, a.yr/a.yi).
template <bool SPLIT>
__device__ __forceinline__ auto hier_in(const HierArgs &a, long row0, int lrow, int t)
{
  if constexpr (SPLIT)
    return IoS{make_rsrc(a.xr + row0 * a.ldx, 0xffffffffu), make_rsrc(a.xi + row0 * a.ldx, 0xffffffffu),
               (uint32_t)(lrow * a.ldx + t) * 4u};
  else
    return IoC{make_rsrc(a.x + row0 * a.ldx, 0xffffffffu), (uint32_t)(lrow * a.ldx + t) * 8u};
}
template <bool SPLIT>
__device__ __forceinline__ auto hier_out(const HierArgs &a, long row0, int lrow, int t)
{
  if constexpr (SPLIT)
    return IoS{make_rsrc(a.yr + row0 * a.ldy, 0xffffffffu), make_rsrc(a.yi + row0 * a.ldy, 0xffffffffu),
               (uint32_t)(lrow * a.ldy + t) * 4u};
  else
    return IoC{make_rsrc(a.y + row0 * a.ldy, 0xffffffffu), (uint32_t)(lrow * a.ldy + t) * 8u};
}

// Outer forward stage: load, DFT_R0, twiddle, write slices.
template <int N, typename IO, typename LP>
__device__ __forceinline__ void outer_forward(const IO &ix, bool valid, LP lds, const rsrc_t &rto, int t)
{
  using H = Hier<N>;
  constexpr int R0 = H::R0, M0 = H::M0, NW = H::NW;
  // the twiddles of all M0 butterflies are loaded up front, with the data,
  // so that their loads overlap instead of each waiting in turn
  float2 tw[M0 * R0];
#pragma unroll
  for (int m = 0; m < M0; ++m)
#pragma unroll
    for (int k1 = 1; k1 < R0; ++k1)
      tw[m * R0 + k1] = bload(rto, (uint32_t)(t + H::TR * m) * 8u, (uint32_t)((k1 - 1) * NW) * 8u);
#pragma unroll
  for (int m = 0; m < M0; ++m) {
    const int j = t + H::TR * m;
    float2 a[R0];
#pragma unroll
    for (int r = 0; r < R0; ++r)
      a[r] = valid ? ix.ld((uint32_t)(H::TR * m + NW * r)) : make_float2(0.f, 0.f);
    if constexpr (R0 > 1) {
      Dft<R0, -1>::run(a);
#pragma unroll
      for (int k1 = 1; k1 < R0; ++k1) a[k1] = cmul(a[k1], tw[m * R0 + k1]);
    }
#pragma unroll
    for (int k1 = 0; k1 < R0; ++k1) lds_st(lds, k1 * H::LSL + H::WE::Layout::idx(j), a[k1]);
  }
}

// Outer inverse stage: read slices, DFT_R0 (inverse), scale, store.
template <int N, typename IO, typename LP>
__device__ __forceinline__ void outer_inverse_store(const IO &iy, bool valid, LP lds, int t, float scale)
{
  using H = Hier<N>;
  constexpr int R0 = H::R0, M0 = H::M0, NW = H::NW;
#pragma unroll
  for (int m = 0; m < M0; ++m) {
    const int j = t + H::TR * m;
    float2 a[R0];
#pragma unroll
    for (int k1 = 0; k1 < R0; ++k1) a[k1] = lds_ld(lds, k1 * H::LSL + H::WE::Layout::idx(j));
    if constexpr (R0 > 1) Dft<R0, +1>::run(a);
    if (scale != 1.0f) {
#pragma unroll
      for (int r = 0; r < R0; ++r) a[r] = cscale(a[r], scale);
    }
    if (valid) {
#pragma unroll
      for (int r = 0; r < R0; ++r) iy.st(a[r], (uint32_t)(H::TR * m + NW * r));
    }
  }
}

// ---------------------------------------------------------------- fastconv
// w_perm: spectrum in [k1][k2] order (w_perm[k1*1024 + k2] = W[k1 + R0 k2]),
// one vector (ldw = 0) or one per row; when natural != 0 the spectrum is in
// natural order and gathered with stride R0 instead.
// Body shared by the standalone kernel and the persistent batch kernel:
// rows row0 + lrow of the job, thread t of the row, LDS row region `lds`.
// Every thread of the workgroup must call it (it contains workgroup
// barriers); rows at or beyond a.rows compute on zeros and store nothing.
// ONE_ROW: the caller's workgroup holds a single row that exists (one row
// per block, one block per row), so no row needs the check.
// SPLIT: X and Y in split storage (a.xr/a.xi, a.yr/a.yi); the spectrum is
// interleaved either way.
template <int N, bool ONE_ROW = false, bool SPLIT = false, typename LP>
__device__ __forceinline__ void hier_fastconv_body(const HierArgs &a, long row0, int lrow, int t, LP lds)
{
  using H = Hier<N>;
  constexpr int R0 = H::R0, NW = H::NW;
  const long row = row0 + lrow;
  const bool valid = ONE_ROW || row < a.rows;
  const auto ix = hier_in<SPLIT>(a, row0, lrow, t);
  const auto iy = hier_out<SPLIT>(a, row0, lrow, t);
  const rsrc_t rw = make_rsrc(a.w + row0 * a.ldw, 0xffffffffu);
  const rsrc_t rtw = make_rsrc(a.tw_wave, (uint32_t)(H::WE::plan.twtotal + 1) * 8u);
  const rsrc_t rto = make_rsrc(a.tw_outer, (uint32_t)((R0 - 1) * H::NW) * 8u); // outer twiddles

  if constexpr (R0 > 1) {
    outer_forward<N>(ix, valid, lds, rto, t);
    __syncthreads();
  }
  // ---- wave stage: wave k1 owns slice k1
  const int k1 = t >> 6;
  const int l = t & 63;
  const int ls = H::WE::Layout::idx(l); // slot of position l; l + 64 s is at ls + s * step(64)
  static_assert(64 % H::WE::Layout::period == 0, "constant-offset slice access");
  LP sl = lds + k1 * H::LSL;
  float2 u[16];
  if constexpr (R0 > 1) {
#pragma unroll
    for (int s = 0; s < 16; ++s) u[s] = lds_ld(sl, ls + s * H::WE::Layout::step(64));
  } else {
#pragma unroll
    for (int s = 0; s < 16; ++s) u[s] = valid ? ix.ld((uint32_t)(64 * s)) : make_float2(0.f, 0.f);
  }
  H::WE::template fft<-1>(u, sl, l, rtw);
  if (a.natural) {
    const uint32_t vw = (uint32_t)(lrow * a.ldw + k1 + R0 * l) * 8u;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      float2 ww = bload(rw, vw, (uint32_t)(R0 * 64 * s) * 8u);
      u[s] = cmul(u[s], ww);
    }
  } else {
    const uint32_t vw = (uint32_t)(lrow * a.ldw + k1 * NW + l) * 8u;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      float2 ww = bload(rw, vw, (uint32_t)(64 * s) * 8u);
      u[s] = cmul(u[s], ww);
    }
  }
  if (a.scale != 1.0f) {
#pragma unroll
    for (int s = 0; s < 16; ++s) u[s] = cscale(u[s], a.scale);
  }
  H::WE::template fft<+1>(u, sl, l, rtw);
  if constexpr (R0 > 1) {
    if (k1 > 0) {
#pragma unroll
      for (int s = 0; s < 16; ++s)
        u[s] = cmulc(u[s], bload(rto, (uint32_t)((k1 - 1) * NW + l) * 8u, (uint32_t)(64 * s) * 8u));
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) lds_st(sl, ls + s * H::WE::Layout::step(64), u[s]);
    __syncthreads();
    outer_inverse_store<N>(iy, valid, lds, t, 1.0f);
  } else {
    if (valid) {
#pragma unroll
      for (int s = 0; s < 16; ++s) iy.st(u[s], (uint32_t)(64 * s));
    }
  }
}

// ---------------------------------------------------------------- transforms
// Forward: outer stage, wave FFT, then the natural-order result
// X[k1 + R0 k2] is staged through LDS so that the store is coalesced.
template <int N, bool SPLIT = false, typename LP>
__device__ __forceinline__ void hier_fwd_body(const HierArgs &a, long row0, int lrow, int t, LP lds)
{
  using H = Hier<N>;
  constexpr int R0 = H::R0;
  const long row = row0 + lrow;
  const bool valid = row < a.rows;
  const auto ix = hier_in<SPLIT>(a, row0, lrow, t);
  const auto iy = hier_out<SPLIT>(a, row0, lrow, t);
  const rsrc_t rtw = make_rsrc(a.tw_wave, (uint32_t)(H::WE::plan.twtotal + 1) * 8u);
  const rsrc_t rto = make_rsrc(a.tw_outer, (uint32_t)((R0 - 1) * H::NW) * 8u); // outer twiddles

  if constexpr (R0 > 1) {
    outer_forward<N>(ix, valid, lds, rto, t);
    __syncthreads();
  }
  const int k1 = t >> 6;
  const int l = t & 63;
  const int ls = H::WE::Layout::idx(l); // slot of position l; l + 64 s is at ls + s * step(64)
  static_assert(64 % H::WE::Layout::period == 0, "constant-offset slice access");
  LP sl = lds + k1 * H::LSL;
  float2 u[16];
  if constexpr (R0 > 1) {
#pragma unroll
    for (int s = 0; s < 16; ++s) u[s] = lds_ld(sl, ls + s * H::WE::Layout::step(64));
  } else {
#pragma unroll
    for (int s = 0; s < 16; ++s)
      u[s] = valid ? ix.ld((uint32_t)(64 * s)) : make_float2(0.f, 0.f);
  }
  H::WE::template fft<-1>(u, sl, l, rtw);
  if (a.scale != 1.0f) {
#pragma unroll
    for (int s = 0; s < 16; ++s) u[s] = cscale(u[s], a.scale);
  }
  if constexpr (R0 > 1) {
    __syncthreads(); // every wave is done with its slice
#pragma unroll
    for (int s = 0; s < 16; ++s) lds_st(lds, H::lidx(k1 + R0 * (l + 64 * s)), u[s]);
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int q = 0; q < 16; ++q) iy.st(lds_ld(lds, H::lidx(t + H::TR * q)), (uint32_t)(H::TR * q));
    }
  } else {
    if (valid) {
#pragma unroll
      for (int s = 0; s < 16; ++s) iy.st(u[s], (uint32_t)(64 * s));
    }
  }
}

// Inverse: natural-order input staged through LDS into [k1][k2] slices, wave
// inverse FFT, inverse twiddle, outer inverse stage.
template <int N, bool SPLIT = false, typename LP>
__device__ __forceinline__ void hier_inv_body(const HierArgs &a, long row0, int lrow, int t, LP lds)
{
  using H = Hier<N>;
  constexpr int R0 = H::R0, NW = H::NW;
  const long row = row0 + lrow;
  const bool valid = row < a.rows;
  const auto ix = hier_in<SPLIT>(a, row0, lrow, t);
  const auto iy = hier_out<SPLIT>(a, row0, lrow, t);
  const rsrc_t rtw = make_rsrc(a.tw_wave, (uint32_t)(H::WE::plan.twtotal + 1) * 8u);
  const rsrc_t rto = make_rsrc(a.tw_outer, (uint32_t)((R0 - 1) * H::NW) * 8u); // outer twiddles
  const int k1 = t >> 6;
  const int l = t & 63;
  const int ls = H::WE::Layout::idx(l); // slot of position l; l + 64 s is at ls + s * step(64)
  static_assert(64 % H::WE::Layout::period == 0, "constant-offset slice access");
  LP sl = lds + k1 * H::LSL;
  float2 u[16];
  if constexpr (R0 > 1) {
#pragma unroll
    for (int q = 0; q < 16; ++q)
      lds_st(lds, H::lidx(t + H::TR * q), valid ? ix.ld((uint32_t)(H::TR * q)) : make_float2(0.f, 0.f));
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 16; ++s) u[s] = lds_ld(lds, H::lidx(k1 + R0 * (l + 64 * s)));
    __syncthreads(); // all natural-order reads done before slices are reused
  } else {
#pragma unroll
    for (int s = 0; s < 16; ++s)
      u[s] = valid ? ix.ld((uint32_t)(64 * s)) : make_float2(0.f, 0.f);
  }
  H::WE::template fft<+1>(u, sl, l, rtw);
  if constexpr (R0 > 1) {
    if (k1 > 0) {
#pragma unroll
      for (int s = 0; s < 16; ++s)
        u[s] = cmulc(u[s], bload(rto, (uint32_t)((k1 - 1) * NW + l) * 8u, (uint32_t)(64 * s) * 8u));
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) lds_st(sl, ls + s * H::WE::Layout::step(64), u[s]);
    __syncthreads();
    outer_inverse_store<N>(iy, valid, lds, t, a.scale);
  } else {
    if (a.scale != 1.0f) {
#pragma unroll
      for (int s = 0; s < 16; ++s) u[s] = cscale(u[s], a.scale);
    }
    if (valid) {
#pragma unroll
      for (int s = 0; s < 16; ++s) iy.st(u[s], (uint32_t)(64 * s));
    }
  }
}

} // namespace ovx
