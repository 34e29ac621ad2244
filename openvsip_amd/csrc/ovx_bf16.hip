// ovx_bf16.hip -- precision variant of the fused fastconv for N = 16384:
// the radix-16 DFT stages run as dense contractions on the MFMA cores with
// split-precision BF16 products accumulated in FP32.
//
// Stage structure (radices 16, 16, 16, 4) and twiddle table are those of the
// register-Stockham path Engine<16384, 16>: stage s, butterfly j < 1024,
// k = j mod Ns, reads x[j + 1024 r] (r < 16), multiplies by W_{16 Ns}^{r k},
// takes the 16-point DFT and writes output q to
// (j / Ns) 16 Ns + k + q Ns.
//
// One MFMA group is 16 butterflies (the N = 16 columns).  Its DFT is the real
// product Y (32 x 16) = A (32 x 32) . B (32 x 16): the K index runs over
// (r, re/im) pairs, B[(r, re)] = Re x_r, B[(r, im)] = Im x_r, and the top and
// bottom halves of A give Re y_q and Im y_q:
//   A_top[q][(r, re)] = Re c_qr,  A_top[q][(r, im)] = -Im c_qr,
//   A_bot[q][(r, re)] = Im c_qr,  A_bot[q][(r, im)] =  Re c_qr,
// with c_qr the coefficient of x_r in y_q.  For the first stage c_qr is the
// DFT-16 entry exp(DIR 2 pi i q r / 16); in the second stage (Ns = 16) the 16
// butterflies of a group share one k, so the twiddle folds into A:
// c_qr = exp(DIR 2 pi i r (16 q + k) / 256), one wave per k; in the third
// stage (Ns = 256) the twiddles are applied to B on the VALU.
//
// Every operand is split into a bf16 high part and a bf16 low part
// (x = hi + lo, |x - hi - lo| <= 2^-16 |x|, bf16 keeping 8 significant bits)
// and the product is lo.hi + hi.lo + hi.hi, three v_mfma_f32_16x16x32_bf16
// per half, accumulated in FP32.  The last (radix-4) stage, the spectrum
// multiply and the third stage's twiddles stay in FP32 on the VALU.  The row
// lives in LDS (interleaved complex, swizzled so that the strides of every
// stage's accesses hit distinct banks).
//
// MFMA operand layout (16x16x32 bf16): lane l holds A[l % 16][8 (l/16) + i]
// and B[8 (l/16) + i][l % 16], i < 8, and D[4 (l/16) + i][l % 16], i < 4.
// Lane (c = l % 16, h = l / 16) therefore holds inputs r = 4h .. 4h+3 of
// butterfly column c, and outputs q = 4h .. 4h+3 of that column.
#include "ovx_bf16.hpp"
#include "ovx_internal.hpp"

namespace ovx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int kN = 16384;
constexpr int kT = 1024;      // threads: 16 waves
constexpr int kGpw = 4;       // groups of 16 butterflies per wave and stage
using E = Engine<kN, 16>;     // stage plan (16, 16, 16, 4) and twiddles
// element pos of the row -> LDS slot: strides 1, 16 and 256 between the 16
// lanes of a column group map to distinct banks.  Adding a multiple m Ns of
// Ns in {1, 16, 256, 1024} to a position whose low bits below Ns do not carry
// moves the slot by m (Ns + Ns/16 + Ns/256): slot(Ns) below.
__host__ __device__ constexpr int pad(int pos) { return pos + (pos >> 4) + (pos >> 8); }
__host__ __device__ constexpr int slot(int ns) { return ns + ns / 16 + ns / 256; }
constexpr int kLds = pad(kN - 1) + 1;
static_assert(pad(1024 * 15 + 1023) == pad(1023) + 15 * slot(1024), "slot stride of the input reads");
static_assert(E::plan.nst == 4 && E::plan.radix[0] == 16 && E::plan.radix[1] == 16 && E::plan.radix[2] == 16 &&
                  E::plan.radix[3] == 4,
              "stage plan (16, 16, 16, 4) expected");
} // namespace

struct SplitOp {
  bf16x8 hi, lo;
};

// Two floats to packed bf16 (round to nearest even; a in the low half).
__device__ __forceinline__ unsigned pack_bf16(float a, float b)
{
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{ a, b }, bf16x2));
}

// hi = bf16(v), lo = bf16(v - hi); the bf16 -> f32 widening of hi is a shift.
__device__ __forceinline__ SplitOp split8(const float (&v)[8])
{
  u32x4 hi, lo;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float a = v[2 * i], b = v[2 * i + 1];
    const unsigned p = pack_bf16(a, b);
    hi[i] = p;
    lo[i] = pack_bf16(a - __builtin_bit_cast(float, p << 16), b - __builtin_bit_cast(float, p & 0xffff0000u));
  }
  SplitOp s;
  s.hi = __builtin_bit_cast(bf16x8, hi);
  s.lo = __builtin_bit_cast(bf16x8, lo);
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

// A operands (top: Re rows, bottom: Im rows) from this lane's four
// coefficients c_{q, 4h + m}, m < 4.
__device__ __forceinline__ void a_operands(const float2 (&cf)[4], SplitOp &top, SplitOp &bot)
{
  float t[8], b[8];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    t[2 * m] = cf[m].x;
    t[2 * m + 1] = -cf[m].y;
    b[2 * m] = cf[m].y;
    b[2 * m + 1] = cf[m].x;
  }
  top = split8(t);
  bot = split8(b);
}

// Coefficients of stage S for this lane: row q = lane % 16, inputs
// r = 4 (lane / 16) + m.  u256[m] = exp(-2 pi i m / 256) (FP64 rounded once).
template <int S, int DIR>
__device__ __forceinline__ void stage_coefs(int lane, int wave, const float2 *__restrict__ u256, float2 (&cf)[4])
{
  const int q = lane & 15;
  const int h = lane >> 4;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int r = 4 * h + m;
    if constexpr (S == 1) {
      const float2 u = u256[(r * (16 * q + wave)) & 255];
      cf[m] = make_float2(u.x, DIR < 0 ? u.y : -u.y);
    } else {
      const int e = (q * r) & 15;
      const float s = c_cos16[(e + 12) & 15]; // sin(2 pi e / 16)
      cf[m] = make_float2(c_cos16[e], DIR < 0 ? -s : s);
    }
  }
}

// The coefficient matrices of the inverse direction are the conjugates of
// the forward ones: negate the Im c entries, i.e. the odd elements of the top
// operand (-Im c) and the even elements of the bottom operand (Im c), in both
// the hi and the lo parts (bf16 negation is exact).
__device__ __forceinline__ void conj_operands(SplitOp &top, SplitOp &bot)
{
  u32x4 th = __builtin_bit_cast(u32x4, top.hi), tl = __builtin_bit_cast(u32x4, top.lo);
  u32x4 bh = __builtin_bit_cast(u32x4, bot.hi), bl = __builtin_bit_cast(u32x4, bot.lo);
  th ^= 0x80000000u;
  tl ^= 0x80000000u;
  bh ^= 0x00008000u;
  bl ^= 0x00008000u;
  top.hi = __builtin_bit_cast(bf16x8, th);
  top.lo = __builtin_bit_cast(bf16x8, tl);
  bot.hi = __builtin_bit_cast(bf16x8, bh);
  bot.lo = __builtin_bit_cast(bf16x8, bl);
}

// Butterfly handled by column c of group gi of this wave in stage S.
template <int S>
__device__ __forceinline__ int butterfly(int wave, int gi, int c)
{
  if constexpr (S == 1) return wave + 16 * (16 * gi + c); // all with k = wave
  else return 16 * (kGpw * wave + gi) + c;
}

// One radix-16 stage on the LDS row.
// atop/abot: the stage's coefficient matrix (stage_coefs) as A operands.
template <int S, int DIR>
__device__ __forceinline__ void mfma_stage(float2 *lds, const float2 *__restrict__ tw, const SplitOp &atop,
                                           const SplitOp &abot, int wave, int lane_)
{
  constexpr int Ns = E::plan.ns[S];
  constexpr int twoff = E::plan.twoff[S];
  const int lane = fresh(lane_);
  const int c = lane & 15;
  const int h = lane >> 4;
  f32x4 yr[kGpw], yi[kGpw];
#pragma unroll
  for (int gi = 0; gi < kGpw; ++gi) {
    const int j = butterfly<S>(wave, gi, c);
    const int pj = pad(j) + 4 * h * slot(kN / 16);
    float v[8];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int r = 4 * h + m;
      float2 x = lds[pj + m * slot(kN / 16)];
      if constexpr (S == 2) {
        if (r > 0) x = twmul<DIR>(x, tw[twoff + (r - 1) * Ns + (j % Ns)]);
      }
      v[2 * m] = x.x;
      v[2 * m + 1] = x.y;
    }
    const SplitOp b = split8(v);
    f32x4 zr = { 0.f, 0.f, 0.f, 0.f }, zi = { 0.f, 0.f, 0.f, 0.f };
    zr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(atop.lo, b.hi, zr, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abot.lo, b.hi, zi, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(atop.hi, b.lo, zr, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abot.hi, b.lo, zi, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(atop.hi, b.hi, zr, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abot.hi, b.hi, zi, 0, 0, 0);
    yr[gi] = zr;
    yi[gi] = zi;
  }
  __syncthreads(); // every input read before any output is written
#pragma unroll
  for (int gi = 0; gi < kGpw; ++gi) {
    const int j = butterfly<S>(wave, gi, c);
    const int pb = pad((j / Ns) * Ns * 16 + (j % Ns)) + 4 * h * slot(Ns);
#pragma unroll
    for (int i = 0; i < 4; ++i) lds[pb + i * slot(Ns)] = make_float2(yr[gi][i], yi[gi][i]);
  }
  __syncthreads();
}

__global__ __launch_bounds__(kT) void k_bf16x3_fastconv(const float2 *__restrict__ x, long ldx,
                                                         const float2 *__restrict__ w, float2 *__restrict__ y,
                                                         long ldy, long rows, const float2 *__restrict__ tw,
                                                         const float2 *__restrict__ u256, float scale)
{
  __shared__ float2 lds[kLds];
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  const long row = blockIdx.x;
  if (row >= rows) return; // one row per workgroup: the whole group leaves together
  const rsrc_t rx = make_rsrc(x + row * ldx, (uint32_t)kN * 8u);
  const rsrc_t ry = make_rsrc(y + row * ldy, (uint32_t)kN * 8u);
  const rsrc_t rw = make_rsrc(w, (uint32_t)kN * 8u);
  const rsrc_t rtw = make_rsrc(tw, (uint32_t)(E::plan.twtotal + 1) * 8u);
  const uint32_t vt = (uint32_t)t * 8u;

  {
    const int pt = pad(t);
#pragma unroll
    for (int q = 0; q < 16; ++q) lds[pt + q * slot(kT)] = bload_nt(rx, vt, (uint32_t)(kT * q) * 8u);
  }
  // forward coefficient matrices: the DFT-16 (stages 0 and 2) and stage 1's
  // folded matrix; the inverse ones are their conjugates
  SplitOp ftop, fbot, gtop, gbot;
  {
    float2 cf[4];
    stage_coefs<0, -1>(lane, wave, u256, cf);
    a_operands(cf, ftop, fbot);
    stage_coefs<1, -1>(lane, wave, u256, cf);
    a_operands(cf, gtop, gbot);
  }
  __syncthreads();
  mfma_stage<0, -1>(lds, tw, ftop, fbot, wave, lane);
  mfma_stage<1, -1>(lds, tw, gtop, gbot, wave, lane);
  mfma_stage<2, -1>(lds, tw, ftop, fbot, wave, lane);
  conj_operands(ftop, fbot);
  conj_operands(gtop, gbot);
  // radix-4 stage (FP32, VALU) in slot form, spectrum multiply, back in place
  {
    float2 v[16];
    const int ts = fresh(t);
    const int pt = pad(ts);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = lds[pt + q * slot(kT)];
    E::template stage<3, -1>(v, (float2 *)nullptr, ts, rtw);
#pragma unroll
    for (int q = 0; q < 16; ++q)
      lds[pt + q * slot(kT)] = cscale(cmul(v[q], bload(rw, vt, (uint32_t)(kT * q) * 8u)), scale);
  }
  __syncthreads();
  mfma_stage<0, +1>(lds, tw, ftop, fbot, wave, lane);
  mfma_stage<1, +1>(lds, tw, gtop, gbot, wave, lane);
  mfma_stage<2, +1>(lds, tw, ftop, fbot, wave, lane);
  {
    float2 v[16];
    const int ts = fresh(t);
    const int pt = pad(ts);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = lds[pt + q * slot(kT)];
    E::template stage<3, +1>(v, (float2 *)nullptr, ts, rtw);
#pragma unroll
    for (int q = 0; q < 16; ++q) bstore_nt(ry, v[q], vt, (uint32_t)(kT * q) * 8u);
  }
}

hipError_t launch_bf16x3_fastconv(const float2 *x, long ldx, const float2 *w, float2 *y, long ldy, long rows,
                                  const float2 *tw, const float2 *u256, float scale, hipStream_t s)
{
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bf16x3_fastconv, dim3((unsigned)rows), dim3(kT), 0, s, x, ldx, w, y, ldy, rows, tw, u256,
                     scale);
  return hipGetLastError();
}

int bf16x3_length() { return kN; }
int bf16x3_twiddle_count() { return E::plan.twtotal; }

} // namespace ovx
