// ovx_bf16.hip -- precision variant of the fused fastconv for N = 16384:
// the radix-16 DFT stages run as dense contractions on the MFMA cores with
// split-precision BF16 products accumulated in FP32.
//
// A radix-16 DFT of 16 butterflies at once is the real product
//   [Yr; Yi] (32 x 16) = [[Fr, -Fi], [Fi, Fr]] (32 x 32) . [Xr; Xi] (32 x 16)
// with F[q][r] = exp(DIR 2 pi i q r / 16).  Every operand is split into a
// bf16 high part and a bf16 low part (x = hi + lo, |x - hi - lo| <= 2^-16 |x|, bf16 keeping 8
// significant bits)
// and the product is hi.hi + hi.lo + lo.hi, three
// v_mfma_f32_16x16x32_bf16 per half (Yr rows, Yi rows), accumulated in FP32.
// The DFT matrix entries are the FP32-rounded values of cos and sin of
// multiples of 2 pi/16.  Twiddles, the last
// (radix-4) stage and the spectrum multiply stay in FP32 on the VALU.  Data
// between stages live in LDS in planar form (real and imaginary arrays).
// Stage structure (16, 16, 16, 4) and twiddle table are those of the
// register-Stockham path Engine<16384, 16>.
//
// MFMA operand layout (16x16x32 bf16): lane l holds A[l % 16][8 (l/16) + i]
// and B[8 (l/16) + i][l % 16], i < 8, and D[4 (l/16) + i][l % 16], i < 4.
#include "ovx_bf16.hpp"
#include "ovx_internal.hpp"

namespace ovx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int kN = 16384;
constexpr int kT = 1024;             // threads
constexpr int kGroups = kN / 16 / 16; // groups of 16 butterflies per radix-16 stage
constexpr int kGpw = kGroups / 16;    // groups per wave
using E = Engine<kN, 16>;             // stage plan (16, 16, 16, 4) and twiddles
constexpr int pad(int pos) { return pos + (pos >> 4); }
constexpr int kLds = kN + kN / 16;
} // namespace

struct SplitOp {
  bf16x8 hi, lo;
};

__device__ __forceinline__ SplitOp split8(const float (&v)[8])
{
  SplitOp s;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    __bf16 h = (__bf16)v[i];
    s.hi[i] = h;
    s.lo[i] = (__bf16)(v[i] - (float)h);
  }
  return s;
}

// An opaque copy of v: addresses derived from it are recomputed where they
// are used instead of being kept live across the kernel (the 1024-thread
// launch leaves 128 VGPRs per lane).
__device__ __forceinline__ int fresh(int v)
{
  asm volatile("" : "+v"(v));
  return v;
}

// cos(2 pi m / 16), m < 16, correctly rounded to FP32 from the exact values
// (cos(pi/8) = sqrt(2 + sqrt 2)/2, cos(pi/4) = sqrt(2)/2,
// cos(3 pi/8) = sqrt(2 - sqrt 2)/2).
__constant__ float c_cos16[16] = { 1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                                   0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                                   -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                                   0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f };

// A operands of the real 32x32 DFT-16 matrix for this lane, top half (Yr
// rows) and bottom half (Yi rows), split into bf16 hi/lo.
// F[q][r] = cos(2 pi m/16) + i DIR sin(2 pi m/16), m = q r mod 16, and
// sin(2 pi m/16) = cos(2 pi (m - 4)/16).
template <int DIR>
__device__ __forceinline__ void dft16_matrix(int lane_, SplitOp &top, SplitOp &bot)
{
  const int lane = fresh(lane_);
  const int q = lane & 15;        // output point (row)
  const int k0 = 8 * (lane >> 4); // first K index of this lane
  float t[8], b[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int kk = k0 + i;
    const int m = (q * (kk & 15)) & 15;
    const float cs = c_cos16[m];
    const float sn = DIR * c_cos16[(m + 12) & 15];
    if (kk < 16) { // coefficient on Xr[r]
      t[i] = cs;
      b[i] = sn;
    } else { // coefficient on Xi[r]
      t[i] = -sn;
      b[i] = cs;
    }
  }
  top = split8(t);
  bot = split8(b);
}

// One radix-16 stage (stage index S of E's plan) on the planar LDS row.
template <int S, int DIR>
__device__ __forceinline__ void mfma_stage(float *re, float *im, const float2 *__restrict__ tw, int wave, int lane_,
                                           const SplitOp &atop, const SplitOp &abot)
{
  const int lane = fresh(lane_);
  constexpr int Ns = E::plan.ns[S];
  constexpr int twoff = E::plan.twoff[S];
  static_assert(E::plan.radix[S] == 16, "MFMA stage must be radix 16");
  const int c = lane & 15;
  const int h = lane >> 4;
  const int p0 = 8 * (h & 1);
  const bool part_im = h >= 2;
  SplitOp bops[kGpw];
#pragma unroll
  for (int gi = 0; gi < kGpw; ++gi) {
    const int j = (wave * kGpw + gi) * 16 + c;
    const int k = j % Ns;
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = p0 + i;
      const int pos = j + r * (kN / 16);
      float2 x = make_float2(re[pad(pos)], im[pad(pos)]);
      if (Ns > 1 && r > 0) x = twmul<DIR>(x, tw[twoff + (r - 1) * Ns + k]);
      v[i] = part_im ? x.y : x.x;
    }
    bops[gi] = split8(v);
  }
  __syncthreads(); // every input read before any output is written
#pragma unroll
  for (int gi = 0; gi < kGpw; ++gi) {
    f32x4 yr = { 0.f, 0.f, 0.f, 0.f }, yi = { 0.f, 0.f, 0.f, 0.f };
    yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(atop.lo, bops[gi].hi, yr, 0, 0, 0);
    yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(atop.hi, bops[gi].lo, yr, 0, 0, 0);
    yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(atop.hi, bops[gi].hi, yr, 0, 0, 0);
    yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abot.lo, bops[gi].hi, yi, 0, 0, 0);
    yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abot.hi, bops[gi].lo, yi, 0, 0, 0);
    yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abot.hi, bops[gi].hi, yi, 0, 0, 0);
    const int j = (wave * kGpw + gi) * 16 + c;
    const int base = (j / Ns) * Ns * 16 + (j % Ns);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int q = 4 * h + i;
      const int pos = base + q * Ns;
      re[pad(pos)] = yr[i];
      im[pad(pos)] = yi[i];
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kT) void k_bf16x3_fastconv(const float2 *__restrict__ x, long ldx,
                                                         const float2 *__restrict__ w, float2 *__restrict__ y,
                                                         long ldy, long rows, const float2 *__restrict__ tw,
                                                         float scale)
{
  __shared__ float re[kLds], im[kLds];
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  const long row = blockIdx.x;
  if (row >= rows) return; // one row per workgroup: the whole group leaves together
  const float2 *xr = x + row * ldx;
  float2 *yr = y + row * ldy;
  const rsrc_t rtw = make_rsrc(tw, (uint32_t)(E::plan.twtotal + 1) * 8u);

#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int pos = t + kT * q;
    float2 v = xr[pos];
    re[pad(pos)] = v.x;
    im[pad(pos)] = v.y;
  }
  __syncthreads();
  SplitOp ftop, fbot;
  dft16_matrix<-1>(lane, ftop, fbot);
  mfma_stage<0, -1>(re, im, tw, wave, lane, ftop, fbot);
  mfma_stage<1, -1>(re, im, tw, wave, lane, ftop, fbot);
  mfma_stage<2, -1>(re, im, tw, wave, lane, ftop, fbot);
  // radix-4 stage (FP32, VALU) in slot form, spectrum multiply
  float2 v[16];
  int ts = fresh(t);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int pos = ts + kT * q;
    v[q] = make_float2(re[pad(pos)], im[pad(pos)]);
  }
  E::template stage<3, -1>(v, (float2 *)nullptr, ts, rtw);
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = cscale(cmul(v[q], w[ts + kT * q]), scale);
  __syncthreads(); // all slot reads done before the planar row is rewritten
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int pos = ts + kT * q;
    re[pad(pos)] = v[q].x;
    im[pad(pos)] = v[q].y;
  }
  __syncthreads();
  SplitOp itop, ibot;
  dft16_matrix<+1>(lane, itop, ibot);
  mfma_stage<0, +1>(re, im, tw, wave, lane, itop, ibot);
  mfma_stage<1, +1>(re, im, tw, wave, lane, itop, ibot);
  mfma_stage<2, +1>(re, im, tw, wave, lane, itop, ibot);
  ts = fresh(t);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int pos = ts + kT * q;
    v[q] = make_float2(re[pad(pos)], im[pad(pos)]);
  }
  E::template stage<3, +1>(v, (float2 *)nullptr, ts, rtw);
#pragma unroll
  for (int q = 0; q < 16; ++q) yr[ts + kT * q] = v[q];
}

hipError_t launch_bf16x3_fastconv(const float2 *x, long ldx, const float2 *w, float2 *y, long ldy, long rows,
                                  const float2 *tw, float scale, hipStream_t s)
{
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bf16x3_fastconv, dim3((unsigned)rows), dim3(kT), 0, s, x, ldx, w, y, ldy, rows, tw, scale);
  return hipGetLastError();
}

int bf16x3_length() { return kN; }
int bf16x3_twiddle_count() { return E::plan.twtotal; }

} // namespace ovx
