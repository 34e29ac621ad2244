// ovx_real.hip -- real <-> complex Fftm by a half-length complex transform.
//
// A real line x of even length n is read as the complex line
// z[m] = x[2m] + i x[2m+1] of length h = n/2.  With Z = FFT_h(z),
//   E[k] = (Z[k] + conj Z[h-k]) / 2,   O[k] = -i (Z[k] - conj Z[h-k]) / 2
// are the transforms of the even and odd samples, and
//   X[k] = E[k] + W^k O[k],   X[h-k] = conj(E[k] - W^k O[k]),   W = e^{-2 pi i/n},
// for k = 0 .. h/2 (one thread per pair, so the pass runs in place).
// The inverse (unnormalised, the imaginary parts of X[0] and X[h] ignored)
// builds Z[k] = A + i C, Z[h-k] = conj(A - i C) with
// A = X[k] + conj X[h-k], C = W^-k (X[k] - conj X[h-k]), then
// z = IFFT_h(Z) holds x[2m] + i x[2m+1].
#include "ovx_real.hpp"

namespace ovx {

__global__ void k_r2c_post(float2 *__restrict__ y, long ldy, long lines, long h, const float2 *__restrict__ wt,
                           float scale)
{
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long line = blockIdx.y;
  if (k > h / 2 || line >= lines) return;
  float2 *row = y + line * ldy;
  const float2 zk = row[k];
  const float2 zm = k == 0 ? zk : row[h - k];
  const float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  const float2 o = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
  const float2 wo = cmul(wt[k], o);
  row[k] = cscale(cadd(e, wo), scale);
  const float2 d = csub(e, wo);
  if (k == 0) row[h] = cscale(make_float2(d.x, -d.y), scale);
  else if (k != h - k) row[h - k] = cscale(make_float2(d.x, -d.y), scale);
}

__global__ void k_c2r_pre(const float2 *__restrict__ x, long ldx, float2 *__restrict__ z, long ldz, long lines, long h,
                          const float2 *__restrict__ wt)
{
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long line = blockIdx.y;
  if (k > h / 2 || line >= lines) return;
  const float2 *in = x + line * ldx;
  float2 *out = z + line * ldz;
  float2 xk = in[k], xm = in[h - k];
  if (k == 0) { // DC and Nyquist carry no imaginary part in a real signal
    xk.y = 0.f;
    xm.y = 0.f;
  }
  const float2 a = make_float2(xk.x + xm.x, xk.y - xm.y);        // X[k] + conj X[h-k]
  const float2 c = cmulc(make_float2(xk.x - xm.x, xk.y + xm.y), wt[k]); // W^-k (X[k] - conj X[h-k])
  out[k] = make_float2(a.x - c.y, a.y + c.x);                    // A + i C
  if (k != 0 && k != h - k) out[h - k] = make_float2(a.x + c.y, c.x - a.y); // conj(A - i C)
}

// Split-storage forms: the post-pass reads Z from an interleaved workspace
// and writes X as separate real / imaginary arrays; the pre-pass reads X
// from them.  Same arithmetic as the interleaved passes.
__global__ void k_r2c_post_split(const float2 *__restrict__ z, long ldz, float *__restrict__ yr,
                                 float *__restrict__ yi, long ldy, long lines, long h, const float2 *__restrict__ wt,
                                 float scale)
{
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long line = blockIdx.y;
  if (k > h / 2 || line >= lines) return;
  const float2 *row = z + line * ldz;
  float *re = yr + line * ldy, *im = yi + line * ldy;
  const float2 zk = row[k];
  const float2 zm = k == 0 ? zk : row[h - k];
  const float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  const float2 o = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
  const float2 wo = cmul(wt[k], o);
  const float2 xk = cscale(cadd(e, wo), scale);
  re[k] = xk.x;
  im[k] = xk.y;
  const float2 d = csub(e, wo);
  const float2 xm = cscale(make_float2(d.x, -d.y), scale);
  if (k == 0) {
    re[h] = xm.x;
    im[h] = xm.y;
  } else if (k != h - k) {
    re[h - k] = xm.x;
    im[h - k] = xm.y;
  }
}

__global__ void k_c2r_pre_split(const float *__restrict__ xr, const float *__restrict__ xi, long ldx,
                                float2 *__restrict__ z, long ldz, long lines, long h, const float2 *__restrict__ wt)
{
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long line = blockIdx.y;
  if (k > h / 2 || line >= lines) return;
  const float *re = xr + line * ldx, *im = xi + line * ldx;
  float2 *out = z + line * ldz;
  float2 xk = make_float2(re[k], im[k]), xm = make_float2(re[h - k], im[h - k]);
  if (k == 0) { // DC and Nyquist carry no imaginary part in a real signal
    xk.y = 0.f;
    xm.y = 0.f;
  }
  const float2 a = make_float2(xk.x + xm.x, xk.y - xm.y);
  const float2 c = cmulc(make_float2(xk.x - xm.x, xk.y + xm.y), wt[k]);
  out[k] = make_float2(a.x - c.y, a.y + c.x);
  if (k != 0 && k != h - k) out[h - k] = make_float2(a.x + c.y, c.x - a.y);
}

hipError_t launch_r2c_post_split(const float2 *z, long ldz, float *yr, float *yi, long ldy, long lines, long h,
                                 const float2 *wt, float scale, hipStream_t s)
{
  if (lines <= 0 || h <= 0) return hipSuccess;
  dim3 grid((unsigned)((h / 2 + 1 + 255) / 256), (unsigned)lines);
  hipLaunchKernelGGL(k_r2c_post_split, grid, dim3(256), 0, s, z, ldz, yr, yi, ldy, lines, h, wt, scale);
  return hipGetLastError();
}

hipError_t launch_c2r_pre_split(const float *xr, const float *xi, long ldx, float2 *z, long ldz, long lines, long h,
                                const float2 *wt, hipStream_t s)
{
  if (lines <= 0 || h <= 0) return hipSuccess;
  dim3 grid((unsigned)((h / 2 + 1 + 255) / 256), (unsigned)lines);
  hipLaunchKernelGGL(k_c2r_pre_split, grid, dim3(256), 0, s, xr, xi, ldx, z, ldz, lines, h, wt);
  return hipGetLastError();
}

hipError_t launch_r2c_post(float2 *y, long ldy, long lines, long h, const float2 *wt, float scale, hipStream_t s)
{
  if (lines <= 0 || h <= 0) return hipSuccess;
  dim3 grid((unsigned)((h / 2 + 1 + 255) / 256), (unsigned)lines);
  hipLaunchKernelGGL(k_r2c_post, grid, dim3(256), 0, s, y, ldy, lines, h, wt, scale);
  return hipGetLastError();
}

hipError_t launch_c2r_pre(const float2 *x, long ldx, float2 *z, long ldz, long lines, long h, const float2 *wt,
                          hipStream_t s)
{
  if (lines <= 0 || h <= 0) return hipSuccess;
  dim3 grid((unsigned)((h / 2 + 1 + 255) / 256), (unsigned)lines);
  hipLaunchKernelGGL(k_c2r_pre, grid, dim3(256), 0, s, x, ldx, z, ldz, lines, h, wt);
  return hipGetLastError();
}

} // namespace ovx
