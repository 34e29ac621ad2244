// ovx_plan.hpp -- host-side plan objects behind the C-ABI handle.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

#include "../../include/ovxhip.h"
#include "ovx_hier.hpp"
#include "ovx_internal.hpp"

namespace ovx {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int current_device();

// Owning device allocation (hipMalloc / hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf();
  int alloc(size_t n); // grow-only; keeps the buffer when already large enough
  void release();
};

// Device scratch a caller lends to a fastconv call in place of the path's
// own buffer (allocated on first use) for the permuted spectrum.  A call
// that needs more than it lends fails with OVX_ERR_UNSUPPORTED.
struct ExtScratch {
  float2 *wp = nullptr;
  size_t wp_elems = 0;
};

// A line transform of length n: picks the kernel path (LDS-resident rows,
// or the two-pass four-step for longer lines) and owns its twiddles.
struct LineFFT {
  enum Kind { NONE, LDS, FOURSTEP, HIER } kind = NONE;
  long n = 0;
  int variant = 0;
  RowConfig rc;
  DevBuf tw;
  // hierarchical (outer stage + per-wave 1024-point transforms)
  DevBuf tw_outer, tw_wave;
  // two-pass four-step (n = n1 * n2)
  long n1 = 0, n2 = 0;
  int sbits = 0;
  RowConfig rc1, rc2;
  DevBuf tw1, tw2, tlo, thi, tfull;
  mutable DevBuf work, wp; // Fftm workspace, permuted spectrum

  int init(long n, int variant);
  const char *name() const;
  size_t device_bytes() const;
  // lines x n at line strides ld_in / ld_out; dir = -1 forward, +1 inverse
  int transform(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long lines, float scale,
                hipStream_t s) const;
  // same, with a caller-provided four-step workspace of lines * n elements
  // (so that concurrent callers of one path do not share `work`)
  int transform_ws(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long lines, float scale,
                   float2 *ws, hipStream_t s) const;
  int fastconv(const float2 *x, long ldx, const float2 *w, long ldw, float2 *y, long ldy, long lines, float scale,
               hipStream_t s, const ExtScratch *ext = nullptr) const;
  // permuted-spectrum elements fastconv() needs for `lines` lines and
  // spectrum row stride ldw (0 when it permutes nothing)
  size_t fastconv_scratch_wp(long lines, long ldw) const;
  // fastconv with X and Y in split storage (line strides in floats) and an
  // interleaved spectrum, and the plain transform on split lines, read and
  // written by the kernels themselves (native_split(): every transform path)
  bool native_split() const { return kind == HIER || kind == LDS || kind == FOURSTEP; }
  int fastconv_split(const float *xr, const float *xi, long ldx, const float2 *w, long ldw, float *yr, float *yi,
                     long ldy, long lines, float scale, hipStream_t s) const;
  int transform_split(int dir, const float *in_re, const float *in_im, long ld_in, float *out_re, float *out_im,
                      long ld_out, long lines, float scale, hipStream_t s) const;
  int fastconv_split_vec(const float *xr, const float *xi, long ldx, const float *wr, const float *wi, float *yr,
                         float *yi, long ldy, long lines, float scale, hipStream_t s) const;
  int hier_fastconv(HierArgs a, int launch_kind, long ldx, const float2 *w, long ldw, long ldy, long lines,
                    float scale, hipStream_t s, const ExtScratch *ext = nullptr) const;
  int fs_fastconv(const float2 *x, long ldx, const float2 *wprep, long ldwp, float2 *y, long ldy, long lines,
                  hipStream_t s) const;
  // Spectrum layout the fastconv kernels of this path read: wp = scale * w
  // permuted for the path (n values).  fastconv_prepared() takes such a
  // spectrum (one vector for all lines) and applies no further scale.
  int prepare_spectrum(const float2 *w, float2 *wp, float scale, hipStream_t s) const;
  int fastconv_prepared(const float2 *x, long ldx, const float2 *wp, float2 *y, long ldy, long lines,
                        hipStream_t s) const;
  static int variants(long n);
  static int default_variant(long n);
  // every table the kernels of this path read has been allocated
  bool tables_ready() const;
};

enum PlanKind { PLAN_FFTM = 1, PLAN_FFT, PLAN_FASTCONV, PLAN_CORR, PLAN_CONV, PLAN_FFTM_REAL };

struct CorrState; // ovx_corr.cpp

} // namespace ovx

struct ovx_plan_s {
  int kind = 0;
  long rows = 0, cols = 0;
  int axis = OVX_ROW;
  int dir = OVX_FFT_FWD;
  float scale = 1.0f;
  unsigned flags = 0;
  ovx::LineFFT line;
  ovx::DevBuf stage_in, stage_out, stage_w;
  ovx::CorrState *corr = nullptr;
  bool bf16x3 = false; // fastconv radix-16 stages on MFMA (precision variant)
  ovx::DevBuf u256;    // its 256th roots of unity
  ovx::DevBuf wtab;    // real Fftm: exp(-2 pi i k / n), k <= n/4
  ovx::DevBuf split_ws; // split-complex calls: interleaved copy of the lines
  // the stream of the plan's last device-pointer call that used its scratch
  // (the permuted spectrum, shared by the plan's calls), once there was one
  bool scratch_used = false;
  hipStream_t scratch_stream = nullptr;

  ~ovx_plan_s();
  const char *name() const;
  size_t device_bytes() const;
};
