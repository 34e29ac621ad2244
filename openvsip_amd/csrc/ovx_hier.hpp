// ovx_hier.hpp -- launch interface of the hierarchical (outer stage + per-wave
// 1024-point) transforms for N = 1024 * {1, 2, 4, 8, 16}.
#pragma once
#include <hip/hip_runtime.h>

namespace ovx {

enum { HIER_FASTCONV = 0, HIER_FWD = 1, HIER_INV = 2, HIER_FASTCONV_SPLIT = 3, HIER_FWD_SPLIT = 4, HIER_INV_SPLIT = 5 };

struct HierArgs {
  const float2 *x;
  long ldx;
  float2 *y;
  long ldy;
  const float2 *w;        // fastconv spectrum ([k1][k2] order unless natural)
  long ldw;               // 0: one spectrum for all rows
  long rows;
  const float2 *tw_outer; // (R0-1)*1024 entries
  const float2 *tw_wave;  // 1024-point engine table
  float scale;            // applied unless 1
  int natural;            // spectrum in natural order (gathered)
  // HIER_FASTCONV_SPLIT: X and Y as split real / imaginary arrays (line
  // strides ldx, ldy in floats); x and y are unused
  const float *xr = nullptr, *xi = nullptr;
  float *yr = nullptr, *yi = nullptr;
};

bool hier_supported(long n);
int hier_rows_per_block(long n);
int hier_wave_twiddle_count();
hipError_t launch_hier(int kind, long n, const HierArgs &a, hipStream_t s);
hipError_t launch_hier_permute(const float2 *w, long ldw, float2 *wp, long n, long rows, float scale,
                               hipStream_t s);
// one split-storage spectrum vector (n values) into the same order
hipError_t launch_hier_permute_split(const float *wr, const float *wi, float2 *wp, long n, float scale,
                                     hipStream_t s);
hipError_t launch_hier_twiddles(long n, float2 *two, float2 *tw_wave, hipStream_t s);

} // namespace ovx
