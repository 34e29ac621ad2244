// ovx_plan.hpp -- host-side plan objects behind the C-ABI handle.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

#include "../../include/ovxhip.h"
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

// A line transform of length n: picks the kernel path (LDS-resident rows,
// or the two-pass four-step for longer lines) and owns its twiddles.
struct LineFFT {
  enum Kind { NONE, LDS, FOURSTEP } kind = NONE;
  long n = 0;
  RowConfig rc;
  DevBuf tw;
  // two-pass four-step (n = n1 * n2)
  long n1 = 0, n2 = 0;
  int sbits = 0;
  RowConfig rc1, rc2;
  DevBuf tw1, tw2, tlo, thi;
  mutable DevBuf work, wp; // Fftm workspace, permuted spectrum

  int init(long n, int variant);
  const char *name() const;
  size_t device_bytes() const;
  // lines x n at line strides ld_in / ld_out; dir = -1 forward, +1 inverse
  int transform(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long lines, float scale,
                hipStream_t s) const;
  int fastconv(const float2 *x, long ldx, const float2 *w, long ldw, float2 *y, long ldy, long lines, float scale,
               hipStream_t s) const;
};

enum PlanKind { PLAN_FFTM = 1, PLAN_FFT, PLAN_FASTCONV, PLAN_CORR, PLAN_CONV };

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

  ~ovx_plan_s();
  const char *name() const;
  size_t device_bytes() const;
};
