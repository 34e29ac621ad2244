// ovx_capi.cpp -- the C-ABI (include/ovxhip.h): plans, replica cache,
// argument checking, host-pointer staging and status codes.  No exception
// crosses the boundary; every failure becomes an ovx_status plus a
// thread-local message (ovx_last_error).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ovxhip.h"
#include "ovx_bf16.hpp"
#include "ovx_real.hpp"
#include "ovx_fourstep.hpp"
#include "ovx_hier.hpp"
#include "ovx_internal.hpp"
#include "ovx_batch_host.hpp"
#include "ovx_plan.hpp"

using namespace ovx;

namespace {

thread_local std::string g_err;
thread_local int g_device = 0;

} // namespace

namespace ovx {

int fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what)
{
  return fail(e == hipErrorOutOfMemory ? OVX_ERR_NOMEM : OVX_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

int current_device() { return g_device; }

// ---------------------------------------------------------------- device memory
DevBuf::~DevBuf() { release(); }
void DevBuf::release()
{
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}
int DevBuf::alloc(size_t n)
{
  if (n <= bytes && p) return OVX_OK;
  release();
  if (n == 0) return OVX_OK;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    return hip_fail(e, "hipMalloc");
  }
  bytes = n;
  return OVX_OK;
}

// ---------------------------------------------------------------- line transforms
// Build the twiddle table for an LDS-resident configuration.
static int make_row_twiddles(const RowConfig &rc, DevBuf &tw)
{
  size_t n = (size_t)std::max(rc.plan.twtotal, 1);
  int st = tw.alloc(n * sizeof(float2));
  if (st) return st;
  hipError_t e = launch_twiddles(rc.plan, (float2 *)tw.p, nullptr);
  if (e != hipSuccess) return hip_fail(e, "twiddle kernel");
  e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return hip_fail(e, "twiddle kernel");
  return OVX_OK;
}

// Variant used when neither the caller nor the tuner picks one: the
// hierarchical kernel, except at 2048 and 4096, where the register-Stockham
// kernel (16 values per thread) measured faster at every row count tried.
int LineFFT::default_variant(long n) { return (n == 2048 || n == 4096) ? 1 : 0; }

int LineFFT::variants(long n)
{
  if (n >= 1 && n <= kMaxRowN) return (hier_supported(n) ? 1 : 0) + row_variants((int)n);
  return 1;
}

int LineFFT::init(long n_, int variant_)
{
  n = n_;
  variant = variant_;
  if (variant < 0 || variant >= variants(n)) variant = 0;
  if (hier_supported(n) && variant == 0) {
    int st;
    long outer = std::max<long>((n / 1024 - 1) * 1024, 1);
    if ((st = tw_outer.alloc(outer * 8)) || (st = tw_wave.alloc((size_t)hier_wave_twiddle_count() * 8))) return st;
    hipError_t e = launch_hier_twiddles(n, (float2 *)tw_outer.p, (float2 *)tw_wave.p, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "twiddle kernels");
    kind = HIER;
    return OVX_OK;
  }
  if (n >= 1 && n <= kMaxRowN && row_variants((int)n) > 0) {
    int v = hier_supported(n) ? variant - 1 : variant;
    if (!row_config((int)n, v, &rc)) return fail(OVX_ERR_UNSUPPORTED, "no LDS-resident configuration for n=%ld", n);
    kind = LDS;
    return make_row_twiddles(rc, tw);
  }
  if (n > kMaxRowN && (n & (n - 1)) == 0) {
    // four-step: n1 <= n2, both with compiled line kernels
    int lg = 0;
    while ((1l << lg) < n) ++lg;
    n1 = 1l << (lg / 2);
    n2 = n / n1;
    if (!fs_length_supported(n1) || !fs_length_supported(n2))
      return fail(OVX_ERR_UNSUPPORTED, "transform length %ld is not supported by the gfx950 backend", n);
    if (!row_config((int)n1, 0, &rc1) || !row_config((int)n2, 0, &rc2))
      return fail(OVX_ERR_UNSUPPORTED, "no line configuration for %ld x %ld", n1, n2);
    int st;
    if ((st = make_row_twiddles(rc1, tw1)) || (st = make_row_twiddles(rc2, tw2))) return st;
    sbits = (lg + 1) / 2;
    long lo = 1l << sbits, hi = (n + lo - 1) / lo;
    if ((st = tlo.alloc(lo * 8)) || (st = thi.alloc(hi * 8))) return st;
    hipError_t e = launch_unit_roots(n, lo, (float2 *)tlo.p, nullptr);
    if (e == hipSuccess) e = launch_unit_roots(n / lo, hi, (float2 *)thi.p, nullptr);
    if (e == hipSuccess && n <= kFsFullTwiddleMax) {
      // one coalesced load per element instead of two gathered ones and a
      // product (n entries, read from L2 by every row)
      if ((st = tfull.alloc(n * 8))) return st;
      e = launch_fs_full_twiddles(n1, n2, (float2 *)tfull.p, nullptr);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "four-step twiddles");
    kind = FOURSTEP;
    return OVX_OK;
  }
  return fail(OVX_ERR_UNSUPPORTED, "transform length %ld is not supported by the gfx950 backend", n);
}

static FsArgs fs_args(const LineFFT &f)
{
  FsArgs a{};
  a.n1 = f.n1;
  a.n2 = f.n2;
  a.nbig = f.n;
  a.tw1 = (const float2 *)f.tw1.p;
  a.tw2 = (const float2 *)f.tw2.p;
  a.tlo = (const float2 *)f.tlo.p;
  a.thi = (const float2 *)f.thi.p;
  a.sbits = f.sbits;
  a.tfull = (const float2 *)f.tfull.p;
  a.scale = 1.0f;
  return a;
}

const char *LineFFT::name() const
{
  switch (kind) {
  case LDS: return "lds-resident";
  case HIER: return "hierarchical-lds";
  case FOURSTEP: return "four-step";
  default: return "none";
  }
}

size_t LineFFT::device_bytes() const
{
  return tw.bytes + tw1.bytes + tw2.bytes + tlo.bytes + thi.bytes + tfull.bytes + work.bytes + wp.bytes + tw_outer.bytes +
         tw_wave.bytes;
}

bool LineFFT::tables_ready() const
{
  switch (kind) {
  case LDS: return tw.p != nullptr;
  case HIER: return tw_outer.p != nullptr && tw_wave.p != nullptr;
  case FOURSTEP: return tw1.p && tw2.p && tlo.p && thi.p;
  default: return false;
  }
}

static HierArgs hier_args(const LineFFT &f)
{
  HierArgs a{};
  a.tw_outer = (const float2 *)f.tw_outer.p;
  a.tw_wave = (const float2 *)f.tw_wave.p;
  a.scale = 1.0f;
  return a;
}

static bool offsets_fit(long rows_per_block, long ld)
{
  return (unsigned long)rows_per_block * (unsigned long)std::max(ld, 0l) * 8ul < (1ul << 32);
}

int LineFFT::transform(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long lines, float scale,
                       hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
  if (!tables_ready()) return fail(OVX_ERR_STATE, "transform plan has no twiddle tables");
  if (kind == HIER) {
    if (!offsets_fit(hier_rows_per_block(n), std::max(ld_in, ld_out)))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    HierArgs a = hier_args(*this);
    a.x = in;
    a.ldx = ld_in;
    a.y = out;
    a.ldy = ld_out;
    a.rows = lines;
    a.scale = scale;
    hipError_t e = launch_hier(dir < 0 ? HIER_FWD : HIER_INV, n, a, s);
    if (e != hipSuccess) return hip_fail(e, "row fft launch");
    return OVX_OK;
  }
  if (kind == LDS) {
    if ((unsigned long)rc.rows_per_block * (unsigned long)std::max(ld_in, ld_out) * 8ul >= (1ul << 32))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    hipError_t e = launch_row_fft(rc, dir, in, ld_in, out, ld_out, lines, (const float2 *)tw.p, scale, s);
    if (e != hipSuccess) return hip_fail(e, "row fft launch");
    return OVX_OK;
  }
  if (kind == FOURSTEP) {
    int st = work.alloc((size_t)lines * n * 8);
    if (st) return st;
    return transform_ws(dir, in, ld_in, out, ld_out, lines, scale, (float2 *)work.p, s);
  }
  return fail(OVX_ERR_UNSUPPORTED, "no transform path");
}

int LineFFT::transform_ws(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long lines, float scale,
                          float2 *ws, hipStream_t s) const
{
  if (kind != FOURSTEP) return transform(dir, in, ld_in, out, ld_out, lines, scale, s);
  if (lines <= 0) return OVX_OK;
  if (!tables_ready() || !ws) return fail(OVX_ERR_STATE, "four-step transform without tables or workspace");
  {
    if (lines > 65535) return fail(OVX_ERR_UNSUPPORTED, "more than 65535 lines of a four-step transform");
    FsArgs a = fs_args(*this);
    a.in = in;
    a.ld_in = ld_in;
    a.out = ws;
    a.ld_out = n;
    hipError_t e = launch_fs_col(dir < 0 ? 0 : 2, a, lines, s);
    if (e != hipSuccess) return hip_fail(e, "four-step column pass");
    a.in = ws;
    a.ld_in = n;
    a.out = out;
    a.ld_out = ld_out;
    a.scale = scale;
    e = launch_fs_row(dir < 0 ? 1 : 2, a, lines, s);
    if (e != hipSuccess) return hip_fail(e, "four-step row pass");
    return OVX_OK;
  }
}

// The hierarchical fastconv launch (interleaved or split X and Y, set in
// `a`): a single spectrum is permuted once into the kernel's [k1][k2] order
// with the scale folded in; a coefficient matrix is read in natural order.
int LineFFT::hier_fastconv(HierArgs a, int launch_kind, long ldx, const float2 *w, long ldw, long ldy, long lines,
                           float scale, hipStream_t s, const ExtScratch *ext) const
{
  if (!offsets_fit(hier_rows_per_block(n), std::max(std::max(ldx, ldy), ldw)))
    return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
  a.ldx = ldx;
  a.ldy = ldy;
  a.rows = lines;
  if (ldw == 0) {
    float2 *wpp;
    if (ext) {
      if (ext->wp_elems < (size_t)n) return fail(OVX_ERR_UNSUPPORTED, "lent scratch too small for the spectrum");
      wpp = ext->wp;
    } else {
      int st = wp.alloc((size_t)n * 8);
      if (st) return st;
      wpp = (float2 *)wp.p;
    }
    hipError_t e = launch_hier_permute(w, 0, wpp, n, 1, scale, s);
    if (e != hipSuccess) return hip_fail(e, "spectrum permutation");
    a.w = wpp;
    a.ldw = 0;
    a.natural = 0;
    a.scale = 1.0f;
  } else {
    a.w = w;
    a.ldw = ldw;
    a.natural = 1;
    a.scale = scale;
  }
  hipError_t e = launch_hier(launch_kind, n, a, s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "fastconv launch");
}

size_t LineFFT::fastconv_scratch_wp(long lines, long ldw) const
{
  if (kind == HIER && ldw == 0) return (size_t)n;
  if (kind == FOURSTEP) return (size_t)(ldw ? lines : 1) * (size_t)n;
  return 0;
}

int LineFFT::transform_split(int dir, const float *in_re, const float *in_im, long ld_in, float *out_re,
                             float *out_im, long ld_out, long lines, float scale, hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
  if (!native_split()) return fail(OVX_ERR_UNSUPPORTED, "no split-storage transform kernel on this path");
  if (!tables_ready()) return fail(OVX_ERR_STATE, "transform plan has no twiddle tables");
  if (kind == FOURSTEP) {
    // column pass from the split input into the workspace, row pass with
    // the transposed store into the split output
    if (lines > 65535) return fail(OVX_ERR_UNSUPPORTED, "more than 65535 lines of a four-step transform");
    int st = work.alloc((size_t)lines * n * 8);
    if (st) return st;
    FsArgs a = fs_args(*this);
    a.in_re = in_re;
    a.in_im = in_im;
    a.ld_in = ld_in;
    a.out = (float2 *)work.p;
    a.ld_out = n;
    hipError_t e = launch_fs_col(dir < 0 ? 0 : 2, a, lines, s);
    if (e != hipSuccess) return hip_fail(e, "four-step column pass");
    FsArgs b = fs_args(*this);
    b.in = (const float2 *)work.p;
    b.ld_in = n;
    b.out_re = out_re;
    b.out_im = out_im;
    b.ld_out = ld_out;
    b.scale = scale;
    e = launch_fs_row(dir < 0 ? 1 : 2, b, lines, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "four-step row pass");
  }
  if (kind == LDS) {
    if ((unsigned long)rc.rows_per_block * (unsigned long)std::max(ld_in, ld_out) * 8ul >= (1ul << 32))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    hipError_t e = launch_row_fft_split(rc, dir, in_re, in_im, ld_in, out_re, out_im, ld_out, lines,
                                        (const float2 *)tw.p, scale, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "row fft launch");
  }
  if (!offsets_fit(hier_rows_per_block(n), std::max(ld_in, ld_out)))
    return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
  HierArgs a = hier_args(*this);
  a.xr = in_re;
  a.xi = in_im;
  a.ldx = ld_in;
  a.yr = out_re;
  a.yi = out_im;
  a.ldy = ld_out;
  a.rows = lines;
  a.scale = scale;
  hipError_t e = launch_hier(dir < 0 ? HIER_FWD_SPLIT : HIER_INV_SPLIT, n, a, s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "row fft launch");
}

// Split X, Y and a single split spectrum vector: one kernel permutes and
// scales the spectrum straight from its two arrays, then the split fastconv.
int LineFFT::fastconv_split_vec(const float *xr, const float *xi, long ldx, const float *wr, const float *wi,
                                float *yr, float *yi, long ldy, long lines, float scale, hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
  if (!native_split()) return fail(OVX_ERR_UNSUPPORTED, "no split-storage fastconv kernel on this path");
  if (!tables_ready()) return fail(OVX_ERR_STATE, "fastconv plan has no twiddle tables");
  if (kind == FOURSTEP) {
    // interleave the spectrum into the workspace's tail; fastconv_split
    // permutes it before its first pass writes the head
    int st = work.alloc((size_t)(lines + 1) * n * 8);
    if (st) return st;
    float2 *wi2 = (float2 *)work.p + (size_t)lines * n;
    hipError_t e = launch_interleave(wr, wi, 0, wi2, n, 1, n, s);
    if (e != hipSuccess) return hip_fail(e, "interleave");
    return fastconv_split(xr, xi, ldx, wi2, 0, yr, yi, ldy, lines, scale, s);
  }
  int st = wp.alloc((size_t)n * 8);
  if (st) return st;
  if (kind == LDS) {
    // the register-Stockham kernel reads the spectrum in natural order
    hipError_t e = launch_interleave(wr, wi, 0, (float2 *)wp.p, n, 1, n, s);
    if (e != hipSuccess) return hip_fail(e, "interleave");
    return fastconv_split(xr, xi, ldx, (const float2 *)wp.p, 0, yr, yi, ldy, lines, scale, s);
  }
  if (!offsets_fit(hier_rows_per_block(n), std::max(ldx, ldy)))
    return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
  hipError_t e = launch_hier_permute_split(wr, wi, (float2 *)wp.p, n, scale, s);
  if (e != hipSuccess) return hip_fail(e, "spectrum permutation");
  HierArgs a = hier_args(*this);
  a.xr = xr;
  a.xi = xi;
  a.ldx = ldx;
  a.yr = yr;
  a.yi = yi;
  a.ldy = ldy;
  a.rows = lines;
  a.w = (const float2 *)wp.p;
  a.ldw = 0;
  a.natural = 0;
  a.scale = 1.0f;
  e = launch_hier(HIER_FASTCONV_SPLIT, n, a, s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "fastconv launch");
}

int LineFFT::fastconv_split(const float *xr, const float *xi, long ldx, const float2 *w, long ldw, float *yr,
                            float *yi, long ldy, long lines, float scale, hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
  if (!native_split()) return fail(OVX_ERR_UNSUPPORTED, "no split-storage fastconv kernel on this path");
  if (!tables_ready()) return fail(OVX_ERR_STATE, "fastconv plan has no twiddle tables");
  if (kind == FOURSTEP) {
    // the three passes of fs_fastconv with X read split by the first and Y
    // written split by the last; the middle pass works in the workspace
    if (lines > 65535) return fail(OVX_ERR_UNSUPPORTED, "more than 65535 lines of a four-step fastconv");
    const long wrows = ldw ? lines : 1;
    int st = wp.alloc((size_t)wrows * n * 8);
    if (st) return st;
    hipError_t e = launch_fs_permute(w, ldw, (float2 *)wp.p, n1, n2, wrows, scale, s);
    if (e != hipSuccess) return hip_fail(e, "spectrum permutation");
    if ((st = work.alloc((size_t)lines * n * 8))) return st;
    float2 *ws = (float2 *)work.p;
    FsArgs a = fs_args(*this);
    a.wp = (const float2 *)wp.p;
    a.ldwp = ldw ? n : 0;
    a.in_re = xr;
    a.in_im = xi;
    a.ld_in = ldx;
    a.out = ws;
    a.ld_out = n;
    if ((e = launch_fs_col(0, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step column pass");
    a.in_re = a.in_im = nullptr;
    a.in = ws;
    a.ld_in = n;
    if ((e = launch_fs_row(0, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step row pass");
    a.out = nullptr;
    a.out_re = yr;
    a.out_im = yi;
    a.ld_out = ldy;
    if ((e = launch_fs_col(1, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step column pass");
    return OVX_OK;
  }
  if (kind == LDS) {
    long ld = std::max(std::max(ldx, ldy), ldw);
    if ((unsigned long)rc.rows_per_block * (unsigned long)ld * 8ul >= (1ul << 32))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    hipError_t e = launch_row_fastconv_split(rc, xr, xi, ldx, w, ldw, yr, yi, ldy, lines, (const float2 *)tw.p,
                                             scale, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "fastconv launch");
  }
  HierArgs a = hier_args(*this);
  a.xr = xr;
  a.xi = xi;
  a.yr = yr;
  a.yi = yi;
  return hier_fastconv(a, HIER_FASTCONV_SPLIT, ldx, w, ldw, ldy, lines, scale, s);
}

int LineFFT::fastconv(const float2 *x, long ldx, const float2 *w, long ldw, float2 *y, long ldy, long lines,
                      float scale, hipStream_t s, const ExtScratch *ext) const
{
  if (lines <= 0) return OVX_OK;
  if (!tables_ready()) return fail(OVX_ERR_STATE, "fastconv plan has no twiddle tables");
  if (kind == HIER) {
    HierArgs a = hier_args(*this);
    a.x = x;
    a.y = y;
    return hier_fastconv(a, HIER_FASTCONV, ldx, w, ldw, ldy, lines, scale, s, ext);
  }
  if (kind == LDS) {
    long ld = std::max(std::max(ldx, ldy), ldw);
    if ((unsigned long)rc.rows_per_block * (unsigned long)ld * 8ul >= (1ul << 32))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    hipError_t e = launch_row_fastconv(rc, x, ldx, w, ldw, y, ldy, lines, (const float2 *)tw.p, scale, s);
    if (e != hipSuccess) return hip_fail(e, "fastconv launch");
    return OVX_OK;
  }
  if (kind == FOURSTEP) {
    if (lines > 65535) return fail(OVX_ERR_UNSUPPORTED, "more than 65535 lines of a four-step fastconv");
    long wrows = ldw ? lines : 1;
    float2 *wpp;
    if (ext) {
      if (ext->wp_elems < (size_t)wrows * n) return fail(OVX_ERR_UNSUPPORTED, "lent scratch too small for the spectrum");
      wpp = ext->wp;
    } else {
      int st = wp.alloc((size_t)wrows * n * 8);
      if (st) return st;
      wpp = (float2 *)wp.p;
    }
    hipError_t e = launch_fs_permute(w, ldw, wpp, n1, n2, wrows, scale, s);
    if (e != hipSuccess) return hip_fail(e, "spectrum permutation");
    return fs_fastconv(x, ldx, wpp, ldw ? n : 0, y, ldy, lines, s);
  }
  return fail(OVX_ERR_UNSUPPORTED, "no fastconv path");
}

// Four-step fastconv, three passes over all rows: column FFT + twiddle
// (X -> Y), row FFT x spectrum, inverse row FFT, inverse twiddle (Y in
// place), inverse column FFT (Y in place).  Running the passes chunk by chunk
// of rows, to keep a chunk's Y in the Infinity Cache between passes, measured
// slower at every chunk size from 16 to 128 MiB of X (the passes lose
// parallelism faster than they gain cache hits), so one launch per pass
// covers every row.
int LineFFT::fs_fastconv(const float2 *x, long ldx, const float2 *wprep, long ldwp, float2 *y, long ldy, long lines,
                         hipStream_t s) const
{
  FsArgs a = fs_args(*this);
  a.wp = wprep;
  a.ldwp = ldwp;
  a.in = x;
  a.ld_in = ldx;
  a.out = y;
  a.ld_out = ldy;
  hipError_t e;
  if ((e = launch_fs_col(0, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step column pass");
  a.in = y;
  a.ld_in = ldy;
  if ((e = launch_fs_row(0, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step row pass");
  if ((e = launch_fs_col(1, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step column pass");
  return OVX_OK;
}

int LineFFT::prepare_spectrum(const float2 *w, float2 *wp_out, float scale, hipStream_t s) const
{
  hipError_t e;
  if (kind == HIER) e = launch_hier_permute(w, 0, wp_out, n, 1, scale, s);
  else if (kind == FOURSTEP) e = launch_fs_permute(w, 0, wp_out, n1, n2, 1, scale, s);
  else e = launch_scale_copy(w, wp_out, n, scale, s);
  if (e != hipSuccess) return hip_fail(e, "spectrum preparation");
  return OVX_OK;
}

int LineFFT::fastconv_prepared(const float2 *x, long ldx, const float2 *wprep, float2 *y, long ldy, long lines,
                               hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
  if (!tables_ready() || !wprep) return fail(OVX_ERR_STATE, "fastconv plan has no twiddle tables or spectrum");
  if (kind == HIER) {
    if (!offsets_fit(hier_rows_per_block(n), std::max(ldx, ldy)))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    HierArgs a = hier_args(*this);
    a.x = x;
    a.ldx = ldx;
    a.y = y;
    a.ldy = ldy;
    a.rows = lines;
    a.w = wprep;
    a.ldw = 0;
    a.natural = 0;
    a.scale = 1.0f;
    hipError_t e = launch_hier(HIER_FASTCONV, n, a, s);
    if (e != hipSuccess) return hip_fail(e, "fastconv launch");
    return OVX_OK;
  }
  if (kind == FOURSTEP) {
    if (lines > 65535) return fail(OVX_ERR_UNSUPPORTED, "more than 65535 lines of a four-step fastconv");
    return fs_fastconv(x, ldx, wprep, 0, y, ldy, lines, s);
  }
  return fastconv(x, ldx, wprep, 0, y, ldy, lines, 1.0f, s);
}

} // namespace ovx

// ---------------------------------------------------------------- helpers
static int check_plan(ovx_plan p, int kind)
{
  if (!p) return fail(OVX_ERR_INVALID, "null plan");
  if (p->kind != kind) return fail(OVX_ERR_INVALID, "plan of the wrong kind");
  return OVX_OK;
}

static bool is_dir(int d) { return d == OVX_FFT_FWD || d == OVX_FFT_INV; }
static int sign_of(int d) { return d == OVX_FFT_FWD ? -1 : +1; }

// `count` lines of `width` bytes, `pitch` bytes apart, in a host buffer and
// at the same layout in its device staging copy.
struct HostLines {
  size_t count, pitch, width;
  size_t extent() const { return count ? (count - 1) * pitch + width : 0; }
};

static HostLines host_lines(long count, long ld, long len, size_t elem)
{
  const size_t w = (size_t)len * elem;
  return {(size_t)count, count > 1 ? (size_t)ld * elem : w, w};
}

// Only the lines' own bytes move: the gaps between lines stay as they are
// on the host, and uninitialised staging memory is never copied back.
static hipError_t copy_lines(void *dst, const void *src, HostLines g, hipMemcpyKind kind, hipStream_t s)
{
  if (g.count <= 1 || g.pitch == g.width) return hipMemcpyAsync(dst, src, g.extent(), kind, s);
  return hipMemcpy2DAsync(dst, g.pitch, src, g.pitch, g.width, g.count, kind, s);
}

// Host-pointer staging: copy the input lines in, run `body` on device
// pointers, copy the output lines out, wait.
template <class F>
static int with_staging(ovx_plan p, const void *in, HostLines gi, void *out, HostLines go, hipStream_t s, F body)
{
  int st = p->stage_in.alloc(gi.extent());
  if (st) return st;
  st = p->stage_out.alloc(go.extent());
  if (st) return st;
  hipError_t e = copy_lines(p->stage_in.p, in, gi, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "H2D copy");
  st = body(p->stage_in.p, p->stage_out.p);
  if (st) return st;
  e = copy_lines(out, p->stage_out.p, go, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_fail(e, "D2H copy");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "stream synchronize");
  return OVX_OK;
}

// ---------------------------------------------------------------- no-allocation scratch
// The scratch of ovx_fastconv_exec_noalloc: a fixed device array (each GPU
// the library runs on holds its own copy until the process ends).  Host code
// never uses the array's name as an address: its device address is looked up
// once per device (noalloc_setup, from ovx_init or the entry point's first
// call on that device) and kept in g_noalloc_ptr.
static __device__ __align__(256) unsigned char scratch[OVX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};
} // namespace

// With `device` current: look up this device's copy of the scratch array.
static int noalloc_setup(int device)
{
  if (device < 0 || device >= kMaxDevices) return fail(OVX_ERR_UNSUPPORTED, "device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return OVX_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!ptr) return fail(OVX_ERR_DEVICE, "no device address for the scratch array");
  g_noalloc_ptr[device] = ptr;
  return OVX_OK;
}

// ---------------------------------------------------------------- library
extern "C" int ovx_init(int device)
{
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (device < 0 || device >= count) return fail(OVX_ERR_INVALID, "device %d out of range (%d devices)", device, count);
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  g_device = device;
  return noalloc_setup(device);
}

extern "C" const char *ovx_last_error(void) { return g_err.c_str(); }
extern "C" const char *ovx_version(void) { return "ovxhip 0.1 (gfx950)"; }
extern "C" long ovx_max_length(void) { return 2048l * 2048l; }
extern "C" int ovx_length_variants(long n)
{
  if (!ovx_supported_length(n)) return 0;
  return LineFFT::variants(n);
}

namespace ovx {
int tuned_variant(long n, long rows);
}

// Explicit variant from the flags, else the autotuner's cached choice for
// (n, rows), else the default (variant 0).
static int variant_of(unsigned flags, long n = 0, long rows = 0)
{
  int v = (int)((flags >> 8) & 0xffu);
  if (v) return v - 1;
  if (n > 0 && rows > 0) {
    int t = tuned_variant(n, rows);
    if (t >= 0 && t < LineFFT::variants(n)) return t;
  }
  return LineFFT::default_variant(n);
}

extern "C" int ovx_supported_length(long n)
{
  if (n >= 1 && n <= kMaxRowN) return row_variants((int)n) > 0 ? 1 : 0;
  if (n > kMaxRowN && n <= ovx_max_length() && (n & (n - 1)) == 0) return 1;
  return 0;
}

// ---------------------------------------------------------------- Fftm
extern "C" int ovx_fftm_plan_create(ovx_plan *plan, long rows, long cols, int axis, int direction, float scale,
                                    unsigned flags)
{
  if (!plan) return fail(OVX_ERR_INVALID, "null plan pointer");
  *plan = nullptr;
  if (rows < 0 || cols < 0) return fail(OVX_ERR_INVALID, "negative size");
  if (axis != OVX_ROW && axis != OVX_COL) return fail(OVX_ERR_INVALID, "axis must be OVX_ROW or OVX_COL");
  if (!is_dir(direction)) return fail(OVX_ERR_INVALID, "direction must be OVX_FFT_FWD or OVX_FFT_INV");
  std::unique_ptr<ovx_plan_s> p(new (std::nothrow) ovx_plan_s());
  if (!p) return fail(OVX_ERR_NOMEM, "host allocation");
  p->kind = PLAN_FFTM;
  p->rows = rows;
  p->cols = cols;
  p->axis = axis;
  p->dir = direction;
  p->scale = scale;
  p->flags = flags;
  long n = axis == OVX_ROW ? cols : rows;
  if (n > 0) {
    int st = p->line.init(n, variant_of(flags, n, axis == OVX_ROW ? rows : cols));
    if (st) return st;
  }
  *plan = p.release();
  return OVX_OK;
}

// lines of length n at line stride `ld` with unit element stride
static int fftm_geometry(ovx_plan p, long s0, long s1, long rows, long cols, long *n, long *lines, long *ld)
{
  if (rows != p->rows || cols != p->cols)
    return fail(OVX_ERR_INVALID, "size %ldx%ld does not match the plan's %ldx%ld", rows, cols, p->rows, p->cols);
  if (p->axis == OVX_ROW) {
    if (s1 != 1 && cols > 1) return fail(OVX_ERR_UNSUPPORTED, "column stride must be 1 for a row Fftm");
    *n = cols; *lines = rows; *ld = s0;
  } else {
    if (s0 != 1 && rows > 1) return fail(OVX_ERR_UNSUPPORTED, "row stride must be 1 for a column Fftm");
    *n = rows; *lines = cols; *ld = s1;
  }
  if (*lines > 1 && *ld < *n) return fail(OVX_ERR_INVALID, "line stride %ld shorter than the transform %ld", *ld, *n);
  return OVX_OK;
}

// The Fftm and Fft entry points' common body (the plan's kind already checked).
static int fftm_exec(ovx_plan p, const void *in, long in_s0, long in_s1, void *out, long out_s0, long out_s1,
                     long rows, long cols, void *stream)
{
  int st;
  long n, lines, ldi, ldo;
  if ((st = fftm_geometry(p, in_s0, in_s1, rows, cols, &n, &lines, &ldi))) return st;
  if ((st = fftm_geometry(p, out_s0, out_s1, rows, cols, &n, &lines, &ldo))) return st;
  if (n == 0 || lines == 0) return OVX_OK; // empty views may carry null pointers
  if (!in || !out) return fail(OVX_ERR_INVALID, "null data pointer");
  hipStream_t s = (hipStream_t)stream;
  int dir = sign_of(p->dir);
  if (p->flags & OVX_FLAG_HOST_POINTERS) {
    const HostLines gi = host_lines(lines, ldi, n, 8), go = host_lines(lines, ldo, n, 8);
    return with_staging(p, in, gi, out, go, s, [&](void *din, void *dout) {
      return p->line.transform(dir, (const float2 *)din, ldi, (float2 *)dout, ldo, lines, p->scale, s);
    });
  }
  return p->line.transform(dir, (const float2 *)in, ldi, (float2 *)out, ldo, lines, p->scale, s);
}

extern "C" int ovx_fftm_out_of_place(ovx_plan p, const void *in, long in_s0, long in_s1, void *out, long out_s0,
                                     long out_s1, long rows, long cols, void *stream)
{
  int st = check_plan(p, PLAN_FFTM);
  if (st) return st;
  return fftm_exec(p, in, in_s0, in_s1, out, out_s0, out_s1, rows, cols, stream);
}

extern "C" int ovx_fftm_in_place(ovx_plan p, void *inout, long s0, long s1, long rows, long cols, void *stream)
{
  return ovx_fftm_out_of_place(p, inout, s0, s1, inout, s0, s1, rows, cols, stream);
}

// ---------------------------------------------------------------- Fft
extern "C" int ovx_fft_plan_create(ovx_plan *plan, long n, int direction, float scale, unsigned flags)
{
  if (!plan) return fail(OVX_ERR_INVALID, "null plan pointer");
  int st = ovx_fftm_plan_create(plan, 1, n, OVX_ROW, direction, scale, flags);
  if (st) return st;
  (*plan)->kind = PLAN_FFT;
  return OVX_OK;
}

extern "C" int ovx_fft_out_of_place(ovx_plan p, const void *in, long in_stride, void *out, long out_stride, long n,
                                    void *stream)
{
  int st = check_plan(p, PLAN_FFT);
  if (st) return st;
  if (n != p->cols) return fail(OVX_ERR_INVALID, "length %ld does not match the plan's %ld", n, p->cols);
  if ((in_stride != 1 || out_stride != 1) && n > 1) return fail(OVX_ERR_UNSUPPORTED, "strides must be 1");
  return fftm_exec(p, in, n, 1, out, n, 1, 1, n, stream);
}

extern "C" int ovx_fft_in_place(ovx_plan p, void *inout, long stride, long n, void *stream)
{
  return ovx_fft_out_of_place(p, inout, stride, inout, stride, n, stream);
}

// ---------------------------------------------------------------- real Fftm
extern "C" int ovx_fftm_real_plan_create(ovx_plan *plan, long rows, long cols, int axis, int direction, float scale,
                                         unsigned flags)
{
  if (!plan) return fail(OVX_ERR_INVALID, "null plan pointer");
  *plan = nullptr;
  if (rows < 0 || cols < 0) return fail(OVX_ERR_INVALID, "negative size");
  if (axis != OVX_ROW && axis != OVX_COL) return fail(OVX_ERR_INVALID, "axis must be OVX_ROW or OVX_COL");
  if (!is_dir(direction)) return fail(OVX_ERR_INVALID, "direction must be OVX_FFT_FWD or OVX_FFT_INV");
  const long n = axis == OVX_ROW ? cols : rows;
  if (n % 2) return fail(OVX_ERR_UNSUPPORTED, "real transform length %ld is odd", n);
  std::unique_ptr<ovx_plan_s> p(new (std::nothrow) ovx_plan_s());
  if (!p) return fail(OVX_ERR_NOMEM, "host allocation");
  p->kind = PLAN_FFTM_REAL;
  p->rows = rows;
  p->cols = cols;
  p->axis = axis;
  p->dir = direction;
  p->scale = scale;
  p->flags = flags;
  if (n > 0) {
    const long h = n / 2;
    int st = p->line.init(h, variant_of(flags, h, axis == OVX_ROW ? rows : cols));
    if (st) return st;
    if ((st = p->wtab.alloc((size_t)(h / 2 + 1) * 8))) return st;
    hipError_t e = launch_unit_roots(n, h / 2 + 1, (float2 *)p->wtab.p, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "real Fftm tables");
  }
  *plan = p.release();
  return OVX_OK;
}

// line length / count / stride of one side of a real Fftm: `len` values
// per line along the plan's axis, unit stride along it.
static int real_side(ovx_plan p, long s0, long s1, long len, long *lines, long *ld)
{
  if (p->axis == OVX_ROW) {
    if (s1 != 1 && len > 1) return fail(OVX_ERR_UNSUPPORTED, "column stride must be 1 for a row Fftm");
    *lines = p->rows;
    *ld = s0;
  } else {
    if (s0 != 1 && len > 1) return fail(OVX_ERR_UNSUPPORTED, "row stride must be 1 for a column Fftm");
    *lines = p->cols;
    *ld = s1;
  }
  if (*lines > 1 && *ld < len) return fail(OVX_ERR_INVALID, "line stride %ld shorter than the line %ld", *ld, len);
  return OVX_OK;
}

namespace {
// Lines of a real Fftm call: n reals per line on the real side (line stride
// ldr floats, read as ldz complex pairs), h + 1 values on the complex side.
struct RealLines {
  bool fwd;
  long n, h, lines, ldr, ldc, ldz;
};
} // namespace

// Geometry checks shared by the interleaved and split real entry points;
// `real_ptr` is the real side's data (8-byte aligned: read as pairs).
static int real_lines(ovx_plan p, long in_s0, long in_s1, long out_s0, long out_s1, long rows, long cols,
                      RealLines *g)
{
  if (rows != p->rows || cols != p->cols)
    return fail(OVX_ERR_INVALID, "size %ldx%ld does not match the plan's %ldx%ld", rows, cols, p->rows, p->cols);
  g->fwd = p->dir == OVX_FFT_FWD;
  g->n = p->axis == OVX_ROW ? cols : rows;
  g->h = g->n / 2;
  int st;
  // real side: n floats per line; complex side: h + 1 values per line
  if ((st = real_side(p, g->fwd ? in_s0 : out_s0, g->fwd ? in_s1 : out_s1, g->n, &g->lines, &g->ldr))) return st;
  if ((st = real_side(p, g->fwd ? out_s0 : in_s0, g->fwd ? out_s1 : in_s1, g->h + 1, &g->lines, &g->ldc))) return st;
  if (g->n > 0 && g->lines > 1 && (g->ldr % 2))
    return fail(OVX_ERR_UNSUPPORTED, "real line stride %ld must be even (lines are read as complex pairs)", g->ldr);
  g->ldz = g->lines > 1 ? g->ldr / 2 : g->h; // real line as h complex values
  if (g->n && g->lines && (!p->line.tables_ready() || !p->wtab.p))
    return fail(OVX_ERR_STATE, "real Fftm tables missing");
  return OVX_OK;
}

// The device body: forward = h-point complex transform of the real line read
// as pairs, then the r2c post-pass; inverse = c2r pre-pass, then the inverse
// h-point transform in place.  `ldc` is the complex side's line stride.
static int real_run(ovx_plan p, const RealLines &g, const void *din, void *dout, long ldc, hipStream_t s)
{
  if (g.fwd) {
    int r = p->line.transform(-1, (const float2 *)din, g.ldz, (float2 *)dout, ldc, g.lines, 1.0f, s);
    if (r) return r;
    hipError_t e = launch_r2c_post((float2 *)dout, ldc, g.lines, g.h, (const float2 *)p->wtab.p, p->scale, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "r2c pass");
  }
  hipError_t e = launch_c2r_pre((const float2 *)din, ldc, (float2 *)dout, g.ldz, g.lines, g.h,
                                (const float2 *)p->wtab.p, s);
  if (e != hipSuccess) return hip_fail(e, "c2r pass");
  return p->line.transform(+1, (const float2 *)dout, g.ldz, (float2 *)dout, g.ldz, g.lines, p->scale, s);
}

extern "C" int ovx_fftm_real_out_of_place(ovx_plan p, const void *in, long in_s0, long in_s1, void *out, long out_s0,
                                          long out_s1, long rows, long cols, void *stream)
{
  int st = check_plan(p, PLAN_FFTM_REAL);
  if (st) return st;
  RealLines g;
  if ((st = real_lines(p, in_s0, in_s1, out_s0, out_s1, rows, cols, &g))) return st;
  if (g.n == 0 || g.lines == 0) return OVX_OK; // empty views may carry null pointers
  if (!in || !out) return fail(OVX_ERR_INVALID, "null data pointer");
  if (in == out) return fail(OVX_ERR_UNSUPPORTED, "real Fftm runs out of place (input and output differ in type)");
  if (((uintptr_t)(g.fwd ? in : out)) % 8) return fail(OVX_ERR_UNSUPPORTED, "real data must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  auto run = [&](const void *din, void *dout) -> int { return real_run(p, g, din, dout, g.ldc, s); };
  if (p->flags & OVX_FLAG_HOST_POINTERS) {
    const HostLines gr = host_lines(g.lines, g.ldr, g.n, 4), gc = host_lines(g.lines, g.ldc, g.h + 1, 8);
    return with_staging(p, in, g.fwd ? gr : gc, out, g.fwd ? gc : gr, s, run);
  }
  return run(in, out);
}

// ---------------------------------------------------------------- fastconv
extern "C" int ovx_fastconv_plan_create(ovx_plan *plan, long rows, long n, float scale, unsigned flags)
{
  if (!plan) return fail(OVX_ERR_INVALID, "null plan pointer");
  *plan = nullptr;
  if (rows < 0 || n < 1) return fail(OVX_ERR_INVALID, "bad size %ldx%ld", rows, n);
  std::unique_ptr<ovx_plan_s> p(new (std::nothrow) ovx_plan_s());
  if (!p) return fail(OVX_ERR_NOMEM, "host allocation");
  p->kind = PLAN_FASTCONV;
  p->rows = rows;
  p->cols = n;
  p->scale = scale;
  p->flags = flags;
  if (flags & OVX_FLAG_BF16X3) {
    if (n != bf16x3_length() || (flags & OVX_FLAG_HOST_POINTERS))
      return fail(OVX_ERR_UNSUPPORTED, "the BF16x3 MFMA variant runs n = %d on device pointers", bf16x3_length());
    p->bf16x3 = true;
    // its stage plan and twiddles are those of the register-Stockham
    // configuration with 16 values per thread
    int st = p->line.init(n, hier_supported(n) ? 1 : 0);
    if (st) return st;
    if (p->line.kind != LineFFT::LDS || p->line.rc.p != 16 || p->line.rc.plan.twtotal != bf16x3_twiddle_count())
      return fail(OVX_ERR_STATE, "BF16x3 variant: unexpected stage plan");
    if ((st = p->u256.alloc(256 * 8))) return st;
    hipError_t e = launch_unit_roots(256, 256, (float2 *)p->u256.p, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "BF16x3 tables");
    *plan = p.release();
    return OVX_OK;
  }
  int st = p->line.init(n, variant_of(flags, n, rows));
  if (st) return st;
  *plan = p.release();
  return OVX_OK;
}

// A plan's device scratch is shared by its calls, which are ordered on one
// stream.  A call that uses the scratch on another stream than the last such
// call is refused (OVX_ERR_STATE) while that stream still has work queued,
// instead of racing on the scratch; calls on the same stream, the common
// case, cost nothing extra (an event recorded after every call cost about
// 3 us between back-to-back launches).  The check is conservative: it also
// refuses when the other stream's pending work is not this plan's.  Calls
// that use no plan scratch (a prepared replica spectrum, the
// register-Stockham and BF16x3 kernels reading the spectrum as given) skip
// it.
static int claim_plan(ovx_plan p, hipStream_t s)
{
  if (!p->scratch_used || p->scratch_stream == s) return OVX_OK;
  hipError_t q = hipStreamQuery(p->scratch_stream);
  if (q == hipErrorNotReady)
    return fail(OVX_ERR_STATE, "plan still in use by a call on another stream (a plan is not safe for concurrent calls)");
  (void)hipGetLastError(); // a destroyed stream has no work left: its error is not ours
  return OVX_OK;
}

static int mark_plan(ovx_plan p, hipStream_t s, int status)
{
  if (status) return status;
  p->scratch_used = true;
  p->scratch_stream = s;
  return OVX_OK;
}

static int fastconv_exec_body(ovx_plan p, const void *x, long ldx, const void *w, long ldw, void *y, long ldy,
                                long rows, hipStream_t s);

extern "C" int ovx_fastconv_exec(ovx_plan p, const void *x, long ldx, const void *w, long ldw, void *y, long ldy,
                                 long rows, void *stream)
{
  int st = check_plan(p, PLAN_FASTCONV);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  if (p->flags & OVX_FLAG_HOST_POINTERS) return fastconv_exec_body(p, x, ldx, w, ldw, y, ldy, rows, s);
  // the hierarchical and four-step kernels read the spectrum permuted into
  // the plan's buffer
  const bool scratch = !p->bf16x3 && (p->line.kind == LineFFT::HIER || p->line.kind == LineFFT::FOURSTEP);
  if (!scratch) return fastconv_exec_body(p, x, ldx, w, ldw, y, ldy, rows, s);
  if ((st = claim_plan(p, s))) return st;
  return mark_plan(p, s, fastconv_exec_body(p, x, ldx, w, ldw, y, ldy, rows, s));
}

static int fastconv_exec_body(ovx_plan p, const void *x, long ldx, const void *w, long ldw, void *y, long ldy,
                                long rows, hipStream_t s)
{
  int st;
  if (rows < 0 || rows > p->rows) return fail(OVX_ERR_INVALID, "rows %ld outside the plan's %ld", rows, p->rows);
  long n = p->cols;
  if (rows > 1 && (ldx < n || ldy < n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  if (ldw != 0 && rows > 1 && ldw < n) return fail(OVX_ERR_INVALID, "coefficient row stride shorter than n");
  if (rows == 0) return OVX_OK; // empty views may carry null pointers
  if (!x || !w || !y) return fail(OVX_ERR_INVALID, "null data pointer");
  if (p->bf16x3) {
    if (ldw != 0) return fail(OVX_ERR_UNSUPPORTED, "BF16x3 variant takes one spectrum vector");
    if (!p->line.tables_ready() || !p->u256.p) return fail(OVX_ERR_STATE, "BF16x3 tables missing");
    hipError_t e = launch_bf16x3_fastconv((const float2 *)x, ldx, (const float2 *)w, (float2 *)y, ldy, rows,
                                          (const float2 *)p->line.tw.p, (const float2 *)p->u256.p, p->scale, s);
    if (e != hipSuccess) return hip_fail(e, "bf16x3 fastconv launch");
    return OVX_OK;
  }
  if (p->flags & OVX_FLAG_HOST_POINTERS) {
    // stage x and w together, compute into stage_out
    const HostLines gw = host_lines(ldw ? rows : 1, ldw, n, 8);
    if ((st = p->stage_w.alloc(gw.extent()))) return st;
    hipError_t e = copy_lines(p->stage_w.p, w, gw, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "H2D copy");
    return with_staging(p, x, host_lines(rows, ldx, n, 8), y, host_lines(rows, ldy, n, 8), s, [&](void *dx, void *dy) {
      return p->line.fastconv((const float2 *)dx, ldx, (const float2 *)p->stage_w.p, ldw, (float2 *)dy, ldy, rows,
                              p->scale, s);
    });
  }
  return p->line.fastconv((const float2 *)x, ldx, (const float2 *)w, ldw, (float2 *)y, ldy, rows, p->scale, s);
}

// ovx_fastconv_exec with the call's scratch -- the permuted spectrum of a
// hierarchical or four-step plan -- in the library's fixed device array.
// Nothing is allocated (the plan's tables were made when the plan was
// created).
extern "C" int ovx_fastconv_exec_noalloc(ovx_plan p, const void *x, long ldx, const void *w, long ldw, void *y,
                                         long ldy, long rows, void *stream)
{
  int st = check_plan(p, PLAN_FASTCONV);
  if (st) return st;
  if (p->flags & OVX_FLAG_HOST_POINTERS)
    return fail(OVX_ERR_UNSUPPORTED, "the no-allocation fastconv takes device pointers (staging would allocate)");
  if (rows < 0 || rows > p->rows) return fail(OVX_ERR_INVALID, "rows %ld outside the plan's %ld", rows, p->rows);
  const long n = p->cols;
  if (rows > 1 && (ldx < n || ldy < n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  if (ldw != 0 && rows > 1 && ldw < n) return fail(OVX_ERR_INVALID, "coefficient row stride shorter than n");
  if (rows == 0) return OVX_OK;
  if (!x || !w || !y) return fail(OVX_ERR_INVALID, "null data pointer");
  hipStream_t s = (hipStream_t)stream;
  if (p->bf16x3) return fastconv_exec_body(p, x, ldx, w, ldw, y, ldy, rows, s); // reads no scratch
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if ((st = noalloc_setup(dev))) return st;
  const size_t wp_bytes = p->line.fastconv_scratch_wp(rows, ldw) * sizeof(float2);
  if (wp_bytes > OVX_NOALLOC_SCRATCH_BYTES)
    return fail(OVX_ERR_UNSUPPORTED, "the call needs %zu bytes of scratch; the no-allocation limit is %u", wp_bytes,
                (unsigned)OVX_NOALLOC_SCRATCH_BYTES);
  ExtScratch ext;
  ext.wp = (float2 *)g_noalloc_ptr[dev];
  ext.wp_elems = OVX_NOALLOC_SCRATCH_BYTES / sizeof(float2);
  return p->line.fastconv((const float2 *)x, ldx, (const float2 *)w, ldw, (float2 *)y, ldy, rows, p->scale, s, &ext);
}

// ---------------------------------------------------------------- split complex
// Split-complex storage (OpenVSIP's split_complex format: real and imaginary
// parts in two float arrays with the same strides), the fftm_backend
// overloads on std::pair<T*, T*> (backend.hpp:741-783; the FFTW backend's
// split DFTs, fftw.cpp:945-1001).  The kernels read and write split arrays
// themselves (two 32-bit accesses per value instead of one 64-bit one); the
// four-step path reads them in its first pass and writes them in its last.
// A plan without a transform path goes through the plan's workspace: the
// lines are interleaved, the interleaved path runs there in place, and the
// result is split into the output arrays -- two extra passes over the data.
// Either way the floats equal the interleaved call's on the same values.
namespace {

struct SplitIn {
  const float *re, *im;
};
struct SplitOut {
  float *re, *im;
};

// Host-pointer staging of one split side: both parts copied line by line into
// one device buffer, the imaginary part `extent` bytes after the real part.
hipError_t stage_split_in(void *dst, const void *re, const void *im, HostLines g, hipStream_t s)
{
  hipError_t e = copy_lines(dst, re, g, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = copy_lines((char *)dst + g.extent(), im, g, hipMemcpyHostToDevice, s);
  return e;
}

SplitIn staged_in(void *buf, HostLines g) { return {(const float *)buf, (const float *)((char *)buf + g.extent())}; }
SplitOut staged_out(void *buf, HostLines g) { return {(float *)buf, (float *)((char *)buf + g.extent())}; }

} // namespace

static int fftm_split_exec(ovx_plan p, const void *in_re, const void *in_im, long in_s0, long in_s1, void *out_re,
                           void *out_im, long out_s0, long out_s1, long rows, long cols, void *stream)
{
  int st;
  long n, lines, ldi, ldo;
  if ((st = fftm_geometry(p, in_s0, in_s1, rows, cols, &n, &lines, &ldi))) return st;
  if ((st = fftm_geometry(p, out_s0, out_s1, rows, cols, &n, &lines, &ldo))) return st;
  if (n == 0 || lines == 0) return OVX_OK;
  if (!in_re || !in_im || !out_re || !out_im) return fail(OVX_ERR_INVALID, "null data pointer");
  hipStream_t s = (hipStream_t)stream;
  const int dir = sign_of(p->dir);
  auto run = [&](SplitIn a, SplitOut b) -> int {
    if (p->line.native_split())
      return p->line.transform_split(dir, a.re, a.im, ldi, b.re, b.im, ldo, lines, p->scale, s);
    int r = p->split_ws.alloc((size_t)lines * n * 8);
    if (r) return r;
    float2 *ws = (float2 *)p->split_ws.p;
    hipError_t e = launch_interleave(a.re, a.im, ldi, ws, n, lines, n, s);
    if (e != hipSuccess) return hip_fail(e, "interleave");
    if ((r = p->line.transform(dir, ws, n, ws, n, lines, p->scale, s))) return r;
    e = launch_deinterleave(ws, n, b.re, b.im, ldo, lines, n, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "deinterleave");
  };
  if (!(p->flags & OVX_FLAG_HOST_POINTERS))
    return run({(const float *)in_re, (const float *)in_im}, {(float *)out_re, (float *)out_im});
  const HostLines gi = host_lines(lines, ldi, n, 4), go = host_lines(lines, ldo, n, 4);
  if ((st = p->stage_in.alloc(2 * gi.extent())) || (st = p->stage_out.alloc(2 * go.extent()))) return st;
  hipError_t e = stage_split_in(p->stage_in.p, in_re, in_im, gi, s);
  if (e != hipSuccess) return hip_fail(e, "H2D copy");
  if ((st = run(staged_in(p->stage_in.p, gi), staged_out(p->stage_out.p, go)))) return st;
  SplitOut so = staged_out(p->stage_out.p, go);
  e = copy_lines(out_re, so.re, go, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = copy_lines(out_im, so.im, go, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "D2H copy");
}

extern "C" int ovx_fftm_split_out_of_place(ovx_plan p, const void *in_re, const void *in_im, long in_s0, long in_s1,
                                           void *out_re, void *out_im, long out_s0, long out_s1, long rows,
                                           long cols, void *stream)
{
  int st = check_plan(p, PLAN_FFTM);
  if (st) return st;
  return fftm_split_exec(p, in_re, in_im, in_s0, in_s1, out_re, out_im, out_s0, out_s1, rows, cols, stream);
}

extern "C" int ovx_fftm_split_in_place(ovx_plan p, void *re, void *im, long s0, long s1, long rows, long cols,
                                       void *stream)
{
  return ovx_fftm_split_out_of_place(p, re, im, s0, s1, re, im, s0, s1, rows, cols, stream);
}

extern "C" int ovx_fft_split_out_of_place(ovx_plan p, const void *in_re, const void *in_im, long in_stride,
                                          void *out_re, void *out_im, long out_stride, long n, void *stream)
{
  int st = check_plan(p, PLAN_FFT);
  if (st) return st;
  if (n != p->cols) return fail(OVX_ERR_INVALID, "length %ld does not match the plan's %ld", n, p->cols);
  if ((in_stride != 1 || out_stride != 1) && n > 1) return fail(OVX_ERR_UNSUPPORTED, "strides must be 1");
  return fftm_split_exec(p, in_re, in_im, n, 1, out_re, out_im, n, 1, 1, n, stream);
}

extern "C" int ovx_fft_split_in_place(ovx_plan p, void *re, void *im, long stride, long n, void *stream)
{
  return ovx_fft_split_out_of_place(p, re, im, stride, re, im, stride, n, stream);
}

extern "C" int ovx_fastconv_split_exec(ovx_plan p, const void *x_re, const void *x_im, long ldx, const void *w_re,
                                       const void *w_im, long ldw, void *y_re, void *y_im, long ldy, long rows,
                                       void *stream)
{
  int st = check_plan(p, PLAN_FASTCONV);
  if (st) return st;
  if (rows < 0 || rows > p->rows) return fail(OVX_ERR_INVALID, "rows %ld outside the plan's %ld", rows, p->rows);
  const long n = p->cols;
  if (rows > 1 && (ldx < n || ldy < n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  if (ldw != 0 && rows > 1 && ldw < n) return fail(OVX_ERR_INVALID, "coefficient row stride shorter than n");
  if (rows == 0) return OVX_OK;
  if (!x_re || !x_im || !w_re || !w_im || !y_re || !y_im) return fail(OVX_ERR_INVALID, "null data pointer");
  if (p->bf16x3) return fail(OVX_ERR_UNSUPPORTED, "the BF16x3 variant takes interleaved data");
  hipStream_t s = (hipStream_t)stream;
  const long wlines = ldw ? rows : 1;
  auto run = [&](SplitIn x, SplitIn w, SplitOut y) -> int {
    if (p->line.native_split() && ldw == 0)
      return p->line.fastconv_split_vec(x.re, x.im, ldx, w.re, w.im, y.re, y.im, ldy, rows, p->scale, s);
    if (p->line.native_split()) {
      // the kernel reads and writes the split arrays itself; only the
      // coefficient matrix is interleaved first
      int r = p->split_ws.alloc((size_t)wlines * n * 8);
      if (r) return r;
      float2 *ww = (float2 *)p->split_ws.p;
      hipError_t e = launch_interleave(w.re, w.im, ldw, ww, n, wlines, n, s);
      if (e != hipSuccess) return hip_fail(e, "interleave");
      return p->line.fastconv_split(x.re, x.im, ldx, ww, ldw ? n : 0, y.re, y.im, ldy, rows, p->scale, s);
    }
    int r = p->split_ws.alloc((size_t)(rows + wlines) * n * 8);
    if (r) return r;
    float2 *wx = (float2 *)p->split_ws.p, *ww = wx + (size_t)rows * n;
    hipError_t e = launch_interleave(x.re, x.im, ldx, wx, n, rows, n, s);
    if (e == hipSuccess) e = launch_interleave(w.re, w.im, ldw, ww, n, wlines, n, s);
    if (e != hipSuccess) return hip_fail(e, "interleave");
    if ((r = p->line.fastconv(wx, n, ww, ldw ? n : 0, wx, n, rows, p->scale, s))) return r;
    e = launch_deinterleave(wx, n, y.re, y.im, ldy, rows, n, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "deinterleave");
  };
  if (!(p->flags & OVX_FLAG_HOST_POINTERS)) {
    if ((st = claim_plan(p, s))) return st;
    return mark_plan(p, s,
                     run({(const float *)x_re, (const float *)x_im}, {(const float *)w_re, (const float *)w_im},
                         {(float *)y_re, (float *)y_im}));
  }
  const HostLines gx = host_lines(rows, ldx, n, 4), gw = host_lines(wlines, ldw, n, 4),
                  gy = host_lines(rows, ldy, n, 4);
  if ((st = p->stage_in.alloc(2 * gx.extent())) || (st = p->stage_w.alloc(2 * gw.extent())) ||
      (st = p->stage_out.alloc(2 * gy.extent())))
    return st;
  hipError_t e = stage_split_in(p->stage_in.p, x_re, x_im, gx, s);
  if (e == hipSuccess) e = stage_split_in(p->stage_w.p, w_re, w_im, gw, s);
  if (e != hipSuccess) return hip_fail(e, "H2D copy");
  if ((st = run(staged_in(p->stage_in.p, gx), staged_in(p->stage_w.p, gw), staged_out(p->stage_out.p, gy))))
    return st;
  SplitOut so = staged_out(p->stage_out.p, gy);
  e = copy_lines(y_re, so.re, gy, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = copy_lines(y_im, so.im, gy, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "D2H copy");
}

// Real Fftm with a split complex side (fftm_backend<T, complex<T>> and
// <complex<T>, T> split overloads, backend.hpp:600-694; FFTW's split r2c /
// c2r, fftw.cpp:778-903): the r2c post-pass writes, and the c2r pre-pass
// reads, the split arrays themselves.
extern "C" int ovx_fftm_r2c_split(ovx_plan p, const void *in, long in_s0, long in_s1, void *out_re, void *out_im,
                                  long out_s0, long out_s1, long rows, long cols, void *stream)
{
  int st = check_plan(p, PLAN_FFTM_REAL);
  if (st) return st;
  if (p->dir != OVX_FFT_FWD) return fail(OVX_ERR_INVALID, "an inverse real Fftm plan runs complex -> real");
  RealLines g;
  if ((st = real_lines(p, in_s0, in_s1, out_s0, out_s1, rows, cols, &g))) return st;
  if (g.n == 0 || g.lines == 0) return OVX_OK;
  if (!in || !out_re || !out_im) return fail(OVX_ERR_INVALID, "null data pointer");
  const bool host = p->flags & OVX_FLAG_HOST_POINTERS;
  if (!host && ((uintptr_t)in) % 8) return fail(OVX_ERR_UNSUPPORTED, "real data must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const long hc = g.h + 1;
  auto run = [&](const void *din, SplitOut o) -> int {
    // the h-point transform into the workspace, then the post-pass writes
    // the split spectrum straight from it
    int r = p->split_ws.alloc((size_t)g.lines * hc * 8);
    if (r) return r;
    float2 *ws = (float2 *)p->split_ws.p;
    if ((r = p->line.transform(-1, (const float2 *)din, g.ldz, ws, hc, g.lines, 1.0f, s))) return r;
    hipError_t e = launch_r2c_post_split(ws, hc, o.re, o.im, g.ldc, g.lines, g.h, (const float2 *)p->wtab.p,
                                         p->scale, s);
    return e == hipSuccess ? OVX_OK : hip_fail(e, "r2c pass");
  };
  if (!host) return run(in, {(float *)out_re, (float *)out_im});
  const HostLines gr = host_lines(g.lines, g.ldr, g.n, 4), gc = host_lines(g.lines, g.ldc, hc, 4);
  if ((st = p->stage_in.alloc(gr.extent())) || (st = p->stage_out.alloc(2 * gc.extent()))) return st;
  hipError_t e = copy_lines(p->stage_in.p, in, gr, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "H2D copy");
  if ((st = run(p->stage_in.p, staged_out(p->stage_out.p, gc)))) return st;
  SplitOut so = staged_out(p->stage_out.p, gc);
  e = copy_lines(out_re, so.re, gc, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = copy_lines(out_im, so.im, gc, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "D2H copy");
}

extern "C" int ovx_fftm_c2r_split(ovx_plan p, const void *in_re, const void *in_im, long in_s0, long in_s1, void *out,
                                  long out_s0, long out_s1, long rows, long cols, void *stream)
{
  int st = check_plan(p, PLAN_FFTM_REAL);
  if (st) return st;
  if (p->dir != OVX_FFT_INV) return fail(OVX_ERR_INVALID, "a forward real Fftm plan runs real -> complex");
  RealLines g;
  if ((st = real_lines(p, in_s0, in_s1, out_s0, out_s1, rows, cols, &g))) return st;
  if (g.n == 0 || g.lines == 0) return OVX_OK;
  if (!in_re || !in_im || !out) return fail(OVX_ERR_INVALID, "null data pointer");
  const bool host = p->flags & OVX_FLAG_HOST_POINTERS;
  if (!host && ((uintptr_t)out) % 8) return fail(OVX_ERR_UNSUPPORTED, "real data must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const long hc = g.h + 1;
  auto run = [&](SplitIn i, void *dout) -> int {
    // the pre-pass reads the split spectrum, then the inverse h-point
    // transform in place on the output
    hipError_t e = launch_c2r_pre_split(i.re, i.im, g.ldc, (float2 *)dout, g.ldz, g.lines, g.h,
                                        (const float2 *)p->wtab.p, s);
    if (e != hipSuccess) return hip_fail(e, "c2r pass");
    return p->line.transform(+1, (const float2 *)dout, g.ldz, (float2 *)dout, g.ldz, g.lines, p->scale, s);
  };
  if (!host) return run({(const float *)in_re, (const float *)in_im}, out);
  const HostLines gr = host_lines(g.lines, g.ldr, g.n, 4), gc = host_lines(g.lines, g.ldc, hc, 4);
  if ((st = p->stage_in.alloc(2 * gc.extent())) || (st = p->stage_out.alloc(gr.extent()))) return st;
  hipError_t e = stage_split_in(p->stage_in.p, in_re, in_im, gc, s);
  if (e != hipSuccess) return hip_fail(e, "H2D copy");
  if ((st = run(staged_in(p->stage_in.p, gc), p->stage_out.p))) return st;
  e = copy_lines(out, p->stage_out.p, gr, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? OVX_OK : hip_fail(e, "D2H copy");
}

// ---------------------------------------------------------------- replicas
// A replica spectrum kept resident on the device, in natural order and, for
// the default kernel path of its length, prepared once in the layout that
// path's fastconv kernel reads (permuted, scaled by 1/n).
namespace {
struct Replica {
  long n = 0;
  DevBuf spec, prep, nat; // nat: natural order, 1/n folded in (register-Stockham paths)
  int prep_kind = 0;
};
std::mutex g_rep_mu;
std::map<int, std::unique_ptr<Replica>> g_replicas;
int g_next_rep = 1;
} // namespace

extern "C" int ovx_replica_create(long n, const void *replica, long m, int is_spectrum, int *id)
{
  if (!replica || !id) return fail(OVX_ERR_INVALID, "null pointer");
  if (n < 1 || m < 0 || m > n) return fail(OVX_ERR_INVALID, "bad replica length m=%ld n=%ld", m, n);
  std::unique_ptr<Replica> r(new (std::nothrow) Replica());
  if (!r) return fail(OVX_ERR_NOMEM, "host allocation");
  r->n = n;
  int st = r->spec.alloc((size_t)n * 8);
  if (st) return st;
  LineFFT f;
  if ((st = f.init(n, 0))) return st;
  if (is_spectrum) {
    hipError_t e = hipMemcpy(r->spec.p, replica, (size_t)n * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "H2D copy");
  } else {
    // zero-pad to n and take the forward FFT once (t2f, tests/fastconv.cpp:20-26)
    DevBuf tmp;
    if ((st = tmp.alloc((size_t)n * 8))) return st;
    hipError_t e = hipMemset(tmp.p, 0, (size_t)n * 8);
    if (e == hipSuccess && m > 0) e = hipMemcpy(tmp.p, replica, (size_t)m * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "replica upload");
    if ((st = f.transform(-1, (const float2 *)tmp.p, n, (float2 *)r->spec.p, n, 1, 1.0f, nullptr))) return st;
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "replica transform");
  }
  if ((st = r->prep.alloc((size_t)n * 8))) return st;
  if ((st = f.prepare_spectrum((const float2 *)r->spec.p, (float2 *)r->prep.p, 1.0f / (float)n, nullptr))) return st;
  hipError_t e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return hip_fail(e, "replica preparation");
  r->prep_kind = f.kind;
  if (f.kind != LineFFT::LDS && f.kind != LineFFT::FOURSTEP) {
    // a tuned plan may run this length on the register-Stockham path, which
    // reads the spectrum in natural order
    if ((st = r->nat.alloc((size_t)n * 8))) return st;
    hipError_t e2 = launch_scale_copy((const float2 *)r->spec.p, (float2 *)r->nat.p, n, 1.0f / (float)n, nullptr);
    if (e2 == hipSuccess) e2 = hipStreamSynchronize(nullptr);
    if (e2 != hipSuccess) return hip_fail(e2, "replica preparation");
  }
  std::lock_guard<std::mutex> g(g_rep_mu);
  *id = g_next_rep++;
  g_replicas[*id] = std::move(r);
  return OVX_OK;
}

extern "C" int ovx_replica_destroy(int id)
{
  std::lock_guard<std::mutex> g(g_rep_mu);
  auto it = g_replicas.find(id);
  if (it == g_replicas.end()) return fail(OVX_ERR_INVALID, "unknown replica id %d", id);
  g_replicas.erase(it);
  return OVX_OK;
}

extern "C" int ovx_replica_spectrum(int id, const void **spectrum, long *n)
{
  if (!spectrum) return fail(OVX_ERR_INVALID, "null pointer");
  std::lock_guard<std::mutex> g(g_rep_mu);
  auto it = g_replicas.find(id);
  if (it == g_replicas.end()) return fail(OVX_ERR_INVALID, "unknown replica id %d", id);
  *spectrum = it->second->spec.p;
  if (n) *n = it->second->n;
  return OVX_OK;
}

namespace ovx {
int replica_prepared(int id, const float2 **prep, long *n, const LineFFT *f)
{
  std::lock_guard<std::mutex> g(g_rep_mu);
  auto it = g_replicas.find(id);
  if (it == g_replicas.end()) return fail(OVX_ERR_INVALID, "unknown replica id %d", id);
  // register-Stockham paths read the spectrum in natural order whatever
  // their variant; the hierarchical kernels read the [k1][k2] order, the
  // four-step kernels theirs: one layout per kind
  const Replica &r = *it->second;
  if (f->kind == LineFFT::LDS && r.prep_kind != LineFFT::LDS && r.nat.p) {
    *prep = (const float2 *)r.nat.p;
    *n = r.n;
    return OVX_OK;
  }
  bool ok = r.prep_kind == f->kind;
  if (!ok) return fail(OVX_ERR_STATE, "replica %d is not prepared for this kernel path", id);
  *prep = (const float2 *)r.prep.p;
  *n = it->second->n;
  return OVX_OK;
}
} // namespace ovx

extern "C" int ovx_fastconv_exec_replica(ovx_plan p, const void *x, long ldx, int replica, void *y, long ldy,
                                         long rows, void *stream)
{
  int st = check_plan(p, PLAN_FASTCONV);
  if (st) return st;
  if (!x || !y) return fail(OVX_ERR_INVALID, "null data pointer");
  if (p->flags & OVX_FLAG_HOST_POINTERS)
    return fail(OVX_ERR_UNSUPPORTED, "resident-replica fastconv takes device pointers");
  if (rows < 0 || rows > p->rows) return fail(OVX_ERR_INVALID, "rows %ld outside the plan's %ld", rows, p->rows);
  long n = p->cols;
  if (rows > 1 && (ldx < n || ldy < n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  const float2 *spec, *prep = nullptr;
  long rn = 0;
  {
    std::lock_guard<std::mutex> g(g_rep_mu);
    auto it = g_replicas.find(replica);
    if (it == g_replicas.end()) return fail(OVX_ERR_INVALID, "unknown replica id %d", replica);
    if (it->second->n != n) return fail(OVX_ERR_INVALID, "replica length %ld != plan length %ld", it->second->n, n);
    spec = (const float2 *)it->second->spec.p;
  }
  hipStream_t s = (hipStream_t)stream;
  if (rows == 0) return OVX_OK;
  if (p->bf16x3) return ovx_fastconv_exec(p, x, ldx, spec, 0, y, ldy, rows, stream);
  // the prepared spectrum is the replica's: no plan scratch
  if (p->scale == 1.0f / (float)n && replica_prepared(replica, &prep, &rn, &p->line) == OVX_OK)
    return p->line.fastconv_prepared((const float2 *)x, ldx, prep, (float2 *)y, ldy, rows, s);
  return ovx_fastconv_exec(p, x, ldx, spec, 0, y, ldy, rows, stream);
}

// ---------------------------------------------------------------- plans
extern "C" int ovx_plan_destroy(ovx_plan p)
{
  if (!p) return fail(OVX_ERR_INVALID, "null plan");
  delete p;
  return OVX_OK;
}

extern "C" const char *ovx_plan_name(ovx_plan p)
{
  if (!p) return "null";
  return p->name();
}

extern "C" size_t ovx_plan_device_bytes(ovx_plan p)
{
  if (!p) return 0;
  return p->device_bytes();
}
