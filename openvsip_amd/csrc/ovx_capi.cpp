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
#include "ovx_fourstep.hpp"
#include "ovx_internal.hpp"
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

int LineFFT::init(long n_, int variant)
{
  n = n_;
  if (n >= 1 && n <= kMaxRowN && row_variants((int)n) > 0) {
    if (!row_config((int)n, variant < row_variants((int)n) ? variant : 0, &rc))
      return fail(OVX_ERR_UNSUPPORTED, "no LDS-resident configuration for n=%ld", n);
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
  a.scale = 1.0f;
  return a;
}

const char *LineFFT::name() const
{
  switch (kind) {
  case LDS: return "lds-resident";
  case FOURSTEP: return "four-step";
  default: return "none";
  }
}

size_t LineFFT::device_bytes() const
{
  return tw.bytes + tw1.bytes + tw2.bytes + tlo.bytes + thi.bytes + work.bytes + wp.bytes;
}

int LineFFT::transform(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long lines, float scale,
                       hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
  if (kind == LDS) {
    if ((unsigned long)rc.rows_per_block * (unsigned long)std::max(ld_in, ld_out) * 8ul >= (1ul << 32))
      return fail(OVX_ERR_UNSUPPORTED, "row stride too large for 32-bit buffer offsets");
    hipError_t e = launch_row_fft(rc, dir, in, ld_in, out, ld_out, lines, (const float2 *)tw.p, scale, s);
    if (e != hipSuccess) return hip_fail(e, "row fft launch");
    return OVX_OK;
  }
  if (kind == FOURSTEP) {
    if (lines > 65535) return fail(OVX_ERR_UNSUPPORTED, "more than 65535 lines of a four-step transform");
    int st = work.alloc((size_t)lines * n * 8);
    if (st) return st;
    FsArgs a = fs_args(*this);
    a.in = in;
    a.ld_in = ld_in;
    a.out = (float2 *)work.p;
    a.ld_out = n;
    hipError_t e = launch_fs_col(dir < 0 ? 0 : 2, a, lines, s);
    if (e != hipSuccess) return hip_fail(e, "four-step column pass");
    a.in = (const float2 *)work.p;
    a.ld_in = n;
    a.out = out;
    a.ld_out = ld_out;
    a.scale = scale;
    e = launch_fs_row(dir < 0 ? 1 : 2, a, lines, s);
    if (e != hipSuccess) return hip_fail(e, "four-step row pass");
    return OVX_OK;
  }
  return fail(OVX_ERR_UNSUPPORTED, "no transform path");
}

int LineFFT::fastconv(const float2 *x, long ldx, const float2 *w, long ldw, float2 *y, long ldy, long lines,
                      float scale, hipStream_t s) const
{
  if (lines <= 0) return OVX_OK;
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
    int st = wp.alloc((size_t)wrows * n * 8);
    if (st) return st;
    hipError_t e = launch_fs_permute(w, ldw, (float2 *)wp.p, n1, n2, wrows, scale, s);
    if (e != hipSuccess) return hip_fail(e, "spectrum permutation");
    FsArgs a = fs_args(*this);
    a.wp = (const float2 *)wp.p;
    a.ldwp = ldw ? n : 0;
    a.in = x;
    a.ld_in = ldx;
    a.out = y;
    a.ld_out = ldy;
    if ((e = launch_fs_col(0, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step column pass");
    a.in = y;
    a.ld_in = ldy;
    if ((e = launch_fs_row(0, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step row pass");
    if ((e = launch_fs_col(1, a, lines, s)) != hipSuccess) return hip_fail(e, "four-step column pass");
    return OVX_OK;
  }
  return fail(OVX_ERR_UNSUPPORTED, "no fastconv path");
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

// Host-pointer staging: copy in, run `body` on device pointers, copy out,
// wait.  in_bytes/out_bytes cover the strided extents.
template <class F>
static int with_staging(ovx_plan p, const void *in, size_t in_bytes, void *out, size_t out_bytes, hipStream_t s,
                        F body)
{
  int st = p->stage_in.alloc(in_bytes);
  if (st) return st;
  st = p->stage_out.alloc(out_bytes);
  if (st) return st;
  hipError_t e = hipMemcpyAsync(p->stage_in.p, in, in_bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "H2D copy");
  st = body(p->stage_in.p, p->stage_out.p);
  if (st) return st;
  e = hipMemcpyAsync(out, p->stage_out.p, out_bytes, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_fail(e, "D2H copy");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "stream synchronize");
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
  return OVX_OK;
}

extern "C" const char *ovx_last_error(void) { return g_err.c_str(); }
extern "C" const char *ovx_version(void) { return "ovxhip 0.1 (gfx950)"; }
extern "C" long ovx_max_length(void) { return 2048l * 2048l; }
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
    int st = p->line.init(n, 0);
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

extern "C" int ovx_fftm_out_of_place(ovx_plan p, const void *in, long in_s0, long in_s1, void *out, long out_s0,
                                     long out_s1, long rows, long cols, void *stream)
{
  int st = check_plan(p, PLAN_FFTM);
  if (st) return st;
  if (!in || !out) return fail(OVX_ERR_INVALID, "null data pointer");
  long n, lines, ldi, ldo;
  if ((st = fftm_geometry(p, in_s0, in_s1, rows, cols, &n, &lines, &ldi))) return st;
  if ((st = fftm_geometry(p, out_s0, out_s1, rows, cols, &n, &lines, &ldo))) return st;
  if (n == 0 || lines == 0) return OVX_OK;
  hipStream_t s = (hipStream_t)stream;
  int dir = sign_of(p->dir);
  if (p->flags & OVX_FLAG_HOST_POINTERS) {
    size_t ib = ((size_t)(lines - 1) * ldi + n) * 8, ob = ((size_t)(lines - 1) * ldo + n) * 8;
    return with_staging(p, in, ib, out, ob, s, [&](void *din, void *dout) {
      return p->line.transform(dir, (const float2 *)din, ldi, (float2 *)dout, ldo, lines, p->scale, s);
    });
  }
  return p->line.transform(dir, (const float2 *)in, ldi, (float2 *)out, ldo, lines, p->scale, s);
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
  p->kind = PLAN_FFTM;
  st = ovx_fftm_out_of_place(p, in, n, 1, out, n, 1, 1, n, stream);
  p->kind = PLAN_FFT;
  return st;
}

extern "C" int ovx_fft_in_place(ovx_plan p, void *inout, long stride, long n, void *stream)
{
  return ovx_fft_out_of_place(p, inout, stride, inout, stride, n, stream);
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
  int st = p->line.init(n, 0);
  if (st) return st;
  *plan = p.release();
  return OVX_OK;
}

extern "C" int ovx_fastconv_exec(ovx_plan p, const void *x, long ldx, const void *w, long ldw, void *y, long ldy,
                                 long rows, void *stream)
{
  int st = check_plan(p, PLAN_FASTCONV);
  if (st) return st;
  if (!x || !w || !y) return fail(OVX_ERR_INVALID, "null data pointer");
  if (rows < 0 || rows > p->rows) return fail(OVX_ERR_INVALID, "rows %ld outside the plan's %ld", rows, p->rows);
  long n = p->cols;
  if (rows > 1 && (ldx < n || ldy < n)) return fail(OVX_ERR_INVALID, "row stride shorter than n");
  if (ldw != 0 && rows > 1 && ldw < n) return fail(OVX_ERR_INVALID, "coefficient row stride shorter than n");
  if (rows == 0) return OVX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (p->flags & OVX_FLAG_HOST_POINTERS) {
    // stage x and w together, compute into stage_out
    size_t xb = ((size_t)(rows - 1) * ldx + n) * 8;
    size_t wb = ldw ? ((size_t)(rows - 1) * ldw + n) * 8 : (size_t)n * 8;
    size_t yb = ((size_t)(rows - 1) * ldy + n) * 8;
    if ((st = p->stage_w.alloc(wb))) return st;
    hipError_t e = hipMemcpyAsync(p->stage_w.p, w, wb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "H2D copy");
    return with_staging(p, x, xb, y, yb, s, [&](void *dx, void *dy) {
      return p->line.fastconv((const float2 *)dx, ldx, (const float2 *)p->stage_w.p, ldw, (float2 *)dy, ldy, rows,
                              p->scale, s);
    });
  }
  return p->line.fastconv((const float2 *)x, ldx, (const float2 *)w, ldw, (float2 *)y, ldy, rows, p->scale, s);
}

// ---------------------------------------------------------------- replicas
namespace {
struct Replica {
  long n = 0;
  DevBuf spec;
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
    LineFFT f;
    if ((st = f.init(n, 0))) return st;
    if ((st = f.transform(-1, (const float2 *)tmp.p, n, (float2 *)r->spec.p, n, 1, 1.0f, nullptr))) return st;
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "replica transform");
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
