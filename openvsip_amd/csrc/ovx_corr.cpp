// ovx_corr.cpp -- FFT-based Correlation and Convolution plans (C-ABI part).
//
// Semantics follow the reference's direct implementations:
//   Correlation::correlate   src/ovxx/signal/corr.hpp:663-719
//   corr_full/same/min       src/ovxx/signal/corr.hpp:91-215
//   output sizes             support/test/ref/corr.hpp:24-36,
//                            src/ovxx/signal/conv.hpp:103-121
//   kernel symmetry          src/ovxx/signal/conv.hpp:144-175
//   conv_full/same/min       src/ovxx/signal/conv.hpp:232-330
// The arithmetic is overlap-save on the gfx950 FFT engine (ovx_ols.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <new>
#include <vector>

#include "ovx_hier.hpp"
#include "ovx_ols.hpp"
#include "ovx_plan.hpp"

namespace ovx {

struct CorrState {
  int is_conv = 0;
  int support = 0;
  long M = 0, N = 0, P = 0, D = 1;
  long out_off = 0;
  int L = 0;
  LineFFT fftL;
  DevBuf hpad, H, ref_stage, in_stage, out_stage;
};

// Segment length: a power of two with room for the kernel overlap.
static int choose_L(long M)
{
  long L = 64;
  while (L < 4 * (M - 1) && L < kMaxRowN) L *= 2;
  while (L - (M - 1) < 1 && L < kMaxRowN) L *= 2;
  return (int)L;
}

static int ols_init(CorrState &c)
{
  c.L = choose_L(c.M);
  if (c.M - 1 >= c.L) return fail(OVX_ERR_UNSUPPORTED, "kernel length %ld too long for overlap-save", c.M);
  // the overlap-save kernel is built on the register-Stockham engine: take
  // that path's configuration (variant 0 of it) and its twiddle table
  int st = c.fftL.init(c.L, hier_supported(c.L) ? 1 : 0);
  if (st) return st;
  if (c.fftL.kind != LineFFT::LDS || !c.fftL.tw.p)
    return fail(OVX_ERR_UNSUPPORTED, "no register-Stockham configuration for segment length %d", c.L);
  if ((st = c.hpad.alloc((size_t)c.L * 8))) return st;
  if ((st = c.H.alloc((size_t)c.L * 8))) return st;
  return OVX_OK;
}

// H = FFT_L(h padded); h is a device pointer of M values
static int ols_spectrum(CorrState &c, const float2 *h, int reverse, hipStream_t s)
{
  hipError_t e = launch_ols_prep(h, c.M, reverse, (float2 *)c.hpad.p, c.L, s);
  if (e != hipSuccess) return hip_fail(e, "kernel padding");
  return c.fftL.transform(-1, (const float2 *)c.hpad.p, c.L, (float2 *)c.H.p, c.L, 1, 1.0f, s);
}

static int ols_run(CorrState &c, const float2 *in, float2 *out, int conj_input, int unbiased, hipStream_t s)
{
  if (c.fftL.kind != LineFFT::LDS || !c.fftL.tw.p || !c.H.p)
    return fail(OVX_ERR_STATE, "overlap-save plan has no twiddle table");
  OlsArgs a;
  a.in = in;
  a.N = c.N;
  a.M = c.M;
  a.H = (const float2 *)c.H.p;
  a.tw = (const float2 *)c.fftL.tw.p;
  a.scale = 1.0f / (float)c.L;
  a.conj_input = conj_input;
  a.unbiased = unbiased;
  a.support = c.support;
  a.decimation = c.D;
  a.out_off = c.out_off;
  a.c_first = c.out_off;
  a.c_step = c.L - (c.M - 1);
  a.P = c.P;
  a.out = out;
  long span = (c.P - 1) * c.D + 1; // c indices needed
  long segments = (span + a.c_step - 1) / a.c_step;
  hipError_t e = launch_ols(c.L, a, segments, s);
  if (e != hipSuccess) return hip_fail(e, "overlap-save launch");
  return OVX_OK;
}

} // namespace ovx

using namespace ovx;

ovx_plan_s::~ovx_plan_s()
{
  delete corr;
}

const char *ovx_plan_s::name() const
{
  switch (kind) {
  case PLAN_FFTM: return line.kind == LineFFT::LDS ? "fftm:lds-resident" : line.kind == LineFFT::FOURSTEP ? "fftm:four-step" : line.kind == LineFFT::HIER ? "fftm:hierarchical-lds" : "fftm:none";
  case PLAN_FFT:
    return line.kind == LineFFT::LDS        ? "fft:lds-resident"
           : line.kind == LineFFT::FOURSTEP ? "fft:four-step"
           : line.kind == LineFFT::HIER     ? "fft:hierarchical-lds"
                                            : "fft:none";
  case PLAN_FASTCONV:
    if (bf16x3) return "fastconv:bf16x3-mfma";
    return line.kind == LineFFT::LDS        ? "fastconv:fused-lds-resident"
           : line.kind == LineFFT::FOURSTEP ? "fastconv:four-step-3pass"
           : line.kind == LineFFT::HIER     ? "fastconv:fused-hierarchical-lds"
                                            : "fastconv:none";
  case PLAN_FFTM_REAL:
    return dir == OVX_FFT_FWD ? "fftm-real:r2c-half-length" : "fftm-real:c2r-half-length";
  case PLAN_CORR: return "corr:overlap-save";
  case PLAN_CONV: return "conv:overlap-save";
  default: return "unknown";
  }
}

size_t ovx_plan_s::device_bytes() const
{
  size_t b = line.device_bytes() + stage_in.bytes + stage_out.bytes + stage_w.bytes + u256.bytes + wtab.bytes +
             split_ws.bytes;
  if (corr)
    b += corr->fftL.device_bytes() + corr->hpad.bytes + corr->H.bytes + corr->ref_stage.bytes +
         corr->in_stage.bytes + corr->out_stage.bytes;
  return b;
}

// ---------------------------------------------------------------- correlation
extern "C" long ovx_corr_output_size(int support, long m, long n)
{
  if (support == OVX_SUPPORT_FULL) return n + m - 1;
  if (support == OVX_SUPPORT_SAME) return n;
  return n - m + 1;
}

extern "C" int ovx_corr_plan_create(ovx_plan *plan, int support, long m, long n, unsigned flags)
{
  if (!plan) return fail(OVX_ERR_INVALID, "null plan pointer");
  *plan = nullptr;
  if (support < OVX_SUPPORT_FULL || support > OVX_SUPPORT_MIN) return fail(OVX_ERR_INVALID, "bad support");
  // corr_full requires ref_size <= in_size (corr.hpp:103)
  if (m < 1 || n < 1 || m > n) return fail(OVX_ERR_INVALID, "bad sizes m=%ld n=%ld (need 1 <= m <= n)", m, n);
  std::unique_ptr<ovx_plan_s> p(new (std::nothrow) ovx_plan_s());
  if (!p) return fail(OVX_ERR_NOMEM, "host allocation");
  p->kind = PLAN_CORR;
  p->flags = flags;
  p->corr = new (std::nothrow) CorrState();
  if (!p->corr) return fail(OVX_ERR_NOMEM, "host allocation");
  CorrState &c = *p->corr;
  c.support = support;
  c.M = m;
  c.N = n;
  c.P = ovx_corr_output_size(support, m, n);
  c.D = 1;
  long shift = support == OVX_SUPPORT_FULL ? m - 1 : support == OVX_SUPPORT_SAME ? m / 2 : 0;
  c.out_off = m - 1 - shift;
  int st = ols_init(c);
  if (st) return st;
  *plan = p.release();
  return OVX_OK;
}

extern "C" int ovx_corr_exec(ovx_plan p, int bias, const void *ref, const void *in, void *out, void *stream)
{
  if (!p || p->kind != PLAN_CORR) return fail(OVX_ERR_INVALID, "not a correlation plan");
  if (!ref || !in || !out) return fail(OVX_ERR_INVALID, "null data pointer");
  if (bias != OVX_BIASED && bias != OVX_UNBIASED) return fail(OVX_ERR_INVALID, "bad bias");
  CorrState &c = *p->corr;
  hipStream_t s = (hipStream_t)stream;
  const float2 *dref = (const float2 *)ref, *din = (const float2 *)in;
  float2 *dout = (float2 *)out;
  int st;
  bool host = p->flags & OVX_FLAG_HOST_POINTERS;
  if (host) {
    if ((st = c.ref_stage.alloc((size_t)c.M * 8)) || (st = c.in_stage.alloc((size_t)c.N * 8)) ||
        (st = c.out_stage.alloc((size_t)c.P * 8)))
      return st;
    hipError_t e = hipMemcpyAsync(c.ref_stage.p, ref, (size_t)c.M * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c.in_stage.p, in, (size_t)c.N * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "H2D copy");
    dref = (const float2 *)c.ref_stage.p;
    din = (const float2 *)c.in_stage.p;
    dout = (float2 *)c.out_stage.p;
  }
  if ((st = ols_spectrum(c, dref, 1, s))) return st;
  if ((st = ols_run(c, din, dout, 1, bias == OVX_UNBIASED, s))) return st;
  if (host) {
    hipError_t e = hipMemcpyAsync(out, c.out_stage.p, (size_t)c.P * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "D2H copy");
  }
  return OVX_OK;
}

// ---------------------------------------------------------------- convolution
extern "C" long ovx_conv_output_size(int support, long m, long n, long d)
{
  if (d < 1) return -1;
  if (support == OVX_SUPPORT_FULL) return ((n + m - 2) / d) + 1;
  if (support == OVX_SUPPORT_SAME) return ((n - 1) / d) + 1;
  // VSIP_IMPL_CONV_CORRECT_MIN_SUPPORT_SIZE == 1 (conv.hpp:36, 115-117)
  return ((n - m + 1) / d) + ((n - m + 1) % d == 0 ? 0 : 1);
}

extern "C" int ovx_conv_plan_create(ovx_plan *plan, int symmetry, int support, const void *coeff, long mc, long n,
                                    long decimation, unsigned flags)
{
  if (!plan) return fail(OVX_ERR_INVALID, "null plan pointer");
  *plan = nullptr;
  if (!coeff) return fail(OVX_ERR_INVALID, "null coefficients");
  if (support < OVX_SUPPORT_FULL || support > OVX_SUPPORT_MIN) return fail(OVX_ERR_INVALID, "bad support");
  if (symmetry < OVX_NONSYM || symmetry > OVX_SYM_EVEN_LEN_EVEN) return fail(OVX_ERR_INVALID, "bad symmetry");
  if (mc < 1 || n < 1 || decimation < 1) return fail(OVX_ERR_INVALID, "bad sizes");
  // kernel from coefficients and symmetry (conv.hpp:144-175); coefficients
  // are host data here, as the reference copies them at construction.
  const float2 *cf = (const float2 *)coeff;
  std::vector<float2> k;
  if (symmetry == OVX_SYM_EVEN_LEN_ODD) {
    k.assign(cf, cf + mc);
    for (long i = 0; i < mc - 1; ++i) k.push_back(cf[mc - 2 - i]);
  } else if (symmetry == OVX_SYM_EVEN_LEN_EVEN) {
    k.assign(cf, cf + mc);
    for (long i = 0; i < mc; ++i) k.push_back(cf[mc - 1 - i]);
  } else {
    k.assign(cf, cf + mc);
  }
  long m = (long)k.size();
  if (support == OVX_SUPPORT_MIN && m > n) return fail(OVX_ERR_INVALID, "kernel longer than input for min support");
  std::unique_ptr<ovx_plan_s> p(new (std::nothrow) ovx_plan_s());
  if (!p) return fail(OVX_ERR_NOMEM, "host allocation");
  p->kind = PLAN_CONV;
  p->flags = flags;
  p->corr = new (std::nothrow) CorrState();
  if (!p->corr) return fail(OVX_ERR_NOMEM, "host allocation");
  CorrState &c = *p->corr;
  c.is_conv = 1;
  c.support = support;
  c.M = m;
  c.N = n;
  c.D = decimation;
  c.P = ovx_conv_output_size(support, m, n, decimation);
  c.out_off = support == OVX_SUPPORT_FULL ? 0 : support == OVX_SUPPORT_SAME ? m / 2 : m - 1;
  int st = ols_init(c);
  if (st) return st;
  if ((st = c.ref_stage.alloc((size_t)m * 8))) return st;
  hipError_t e = hipMemcpy(c.ref_stage.p, k.data(), (size_t)m * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "coefficient upload");
  if ((st = ols_spectrum(c, (const float2 *)c.ref_stage.p, 0, nullptr))) return st;
  e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return hip_fail(e, "kernel spectrum");
  *plan = p.release();
  return OVX_OK;
}

extern "C" int ovx_conv_exec(ovx_plan p, const void *in, void *out, void *stream)
{
  if (!p || p->kind != PLAN_CONV) return fail(OVX_ERR_INVALID, "not a convolution plan");
  if (!in || !out) return fail(OVX_ERR_INVALID, "null data pointer");
  CorrState &c = *p->corr;
  hipStream_t s = (hipStream_t)stream;
  const float2 *din = (const float2 *)in;
  float2 *dout = (float2 *)out;
  int st;
  bool host = p->flags & OVX_FLAG_HOST_POINTERS;
  if (host) {
    if ((st = c.in_stage.alloc((size_t)c.N * 8)) || (st = c.out_stage.alloc((size_t)c.P * 8))) return st;
    hipError_t e = hipMemcpyAsync(c.in_stage.p, in, (size_t)c.N * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "H2D copy");
    din = (const float2 *)c.in_stage.p;
    dout = (float2 *)c.out_stage.p;
  }
  if ((st = ols_run(c, din, dout, 0, 0, s))) return st;
  if (host) {
    hipError_t e = hipMemcpyAsync(out, c.out_stage.p, (size_t)c.P * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "D2H copy");
  }
  return OVX_OK;
}
