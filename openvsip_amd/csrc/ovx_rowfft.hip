// ovx_rowfft.hip -- LDS-resident row transforms for N <= 16384 on gfx950:
// forward / inverse Fftm by row and the fused fast convolution
// (forward FFT, multiply by the replica spectrum, inverse FFT, scale) in one
// pass over HBM: each line is read once and written once.
//
// Reference behaviour being replaced:
//   Fftm row loop        src/ovxx/fftw/fftw.cpp:965-977 (one execute per row)
//   scale pass           src/ovxx/signal/fft/workspace.hpp:474-478
//   vmmul<0> row loop    src/vsip/impl/vmmul.hpp:79-85
#include "ovx_internal.hpp"
#include "ovx_row_dev.hpp"

namespace ovx {

// Which lines a block handles: B lines of T threads each, lines contiguous in
// memory at stride ld.  Out-of-range lines (the tail block) compute on zeros
// and store nothing, but still take part in every barrier.
template <int N, int P, int B>
struct RowGeom {
  static constexpr int T = N / P;
  static constexpr int NT = T * B;
};

template <int N, int P, int B, int DIR>
__global__ __launch_bounds__(N / P * B) void k_row_fft(const float2 *__restrict__ in, long ld_in,
                                                      float2 *__restrict__ out, long ld_out, long rows,
                                                      const float2 *__restrict__ tw, float scale)
{
  using E = Engine<N, P, (N / P <= 64)>;
  __shared__ float2 lds_all[B * E::LROW];
  const int lrow = (B == 1) ? 0 : (int)threadIdx.x / E::T;
  const int t = (B == 1) ? (int)threadIdx.x : (int)threadIdx.x % E::T;
  row_fft_body<N, P, DIR>(in, ld_in, out, ld_out, rows, tw, scale, (long)blockIdx.x * B, lrow, t,
                          lds_all + lrow * E::LROW);
}

template <int N, int P, int B>
__global__ __launch_bounds__(N / P * B) void k_row_fastconv(const float2 *__restrict__ x, long ldx,
                                                           const float2 *__restrict__ w, long ldw,
                                                           float2 *__restrict__ y, long ldy, long rows,
                                                           const float2 *__restrict__ tw, float scale)
{
  using E = Engine<N, P, (N / P <= 64)>;
  constexpr int T = E::T;
  __shared__ float2 lds_all[B * E::LROW];
  const int lrow = (B == 1) ? 0 : (int)threadIdx.x / T;
  const int t = (B == 1) ? (int)threadIdx.x : (int)threadIdx.x % T;
  row_fastconv_body<N, P>(x, ldx, w, ldw, y, ldy, rows, tw, scale, (long)blockIdx.x * B, lrow, t,
                          lds_all + lrow * E::LROW);
}

// Split-storage forms: real and imaginary parts of the lines in two float
// arrays with the same strides (the spectrum stays interleaved).
template <int N, int P, int B, int DIR>
__global__ __launch_bounds__(N / P * B) void k_row_fft_split(const float *__restrict__ in_re,
                                                            const float *__restrict__ in_im, long ld_in,
                                                            float *__restrict__ out_re, float *__restrict__ out_im,
                                                            long ld_out, long rows, const float2 *__restrict__ tw,
                                                            float scale)
{
  using E = Engine<N, P, (N / P <= 64)>;
  __shared__ float2 lds_all[B * E::LROW];
  const int lrow = (B == 1) ? 0 : (int)threadIdx.x / E::T;
  const int t = (B == 1) ? (int)threadIdx.x : (int)threadIdx.x % E::T;
  const long row0 = (long)blockIdx.x * B;
  row_fft_io<N, P, DIR>(IoS{make_rsrc(in_re + row0 * ld_in, 0xffffffffu), make_rsrc(in_im + row0 * ld_in, 0xffffffffu),
                            (uint32_t)(lrow * ld_in + t) * 4u},
                        IoS{make_rsrc(out_re + row0 * ld_out, 0xffffffffu),
                            make_rsrc(out_im + row0 * ld_out, 0xffffffffu), (uint32_t)(lrow * ld_out + t) * 4u},
                        rows, tw, scale, row0, lrow, t, lds_all + lrow * E::LROW);
}

template <int N, int P, int B>
__global__ __launch_bounds__(N / P * B) void k_row_fastconv_split(const float *__restrict__ xr,
                                                                 const float *__restrict__ xi, long ldx,
                                                                 const float2 *__restrict__ w, long ldw,
                                                                 float *__restrict__ yr, float *__restrict__ yi,
                                                                 long ldy, long rows, const float2 *__restrict__ tw,
                                                                 float scale)
{
  using E = Engine<N, P, (N / P <= 64)>;
  __shared__ float2 lds_all[B * E::LROW];
  const int lrow = (B == 1) ? 0 : (int)threadIdx.x / E::T;
  const int t = (B == 1) ? (int)threadIdx.x : (int)threadIdx.x % E::T;
  const long row0 = (long)blockIdx.x * B;
  row_fastconv_io<N, P>(IoS{make_rsrc(xr + row0 * ldx, 0xffffffffu), make_rsrc(xi + row0 * ldx, 0xffffffffu),
                            (uint32_t)(lrow * ldx + t) * 4u},
                        IoS{make_rsrc(yr + row0 * ldy, 0xffffffffu), make_rsrc(yi + row0 * ldy, 0xffffffffu),
                            (uint32_t)(lrow * ldy + t) * 4u},
                        w, ldw, rows, tw, scale, row0, lrow, t, lds_all + lrow * E::LROW);
}

// ---------------------------------------------------------------- twiddles
struct TwStageDesc {
  int nst;
  int radix[16];
  int ns[16];
  int off[16];
};

__global__ void k_twiddles(TwStageDesc d, float2 *tw, int total)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  // find the stage holding entry i
  int s = 0;
  for (int k = 0; k < d.nst; ++k)
    if (d.ns[k] > 1 && i >= d.off[k]) s = k;
  const int Ns = d.ns[s];
  const int rel = i - d.off[s];
  const int r = rel / Ns + 1;
  const int k = rel % Ns;
  const long L = (long)Ns * d.radix[s];
  const long m = ((long)r * k) % L;
  double sn, cs;
  sincospi(-2.0 * (double)m / (double)L, &sn, &cs);
  tw[i] = make_float2((float)cs, (float)sn);
}

__global__ void k_unit_roots(long n, long count, float2 *out)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  long m = i % n;
  double sn, cs;
  sincospi(-2.0 * (double)m / (double)n, &sn, &cs);
  out[i] = make_float2((float)cs, (float)sn);
}

hipError_t launch_twiddles(const RadixPlan &plan, float2 *tw, hipStream_t s)
{
  if (plan.twtotal == 0) return hipSuccess;
  TwStageDesc d{};
  d.nst = plan.nst;
  for (int k = 0; k < plan.nst; ++k) { d.radix[k] = plan.radix[k]; d.ns[k] = plan.ns[k]; d.off[k] = plan.twoff[k]; }
  int blocks = (plan.twtotal + 255) / 256;
  hipLaunchKernelGGL(k_twiddles, dim3(blocks), dim3(256), 0, s, d, tw, plan.twtotal);
  return hipGetLastError();
}

__global__ void k_scale_copy(const float2 *__restrict__ in, float2 *__restrict__ out, long n, float scale)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cscale(in[i], scale);
}

hipError_t launch_scale_copy(const float2 *in, float2 *out, long n, float scale, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, scale);
  return hipGetLastError();
}

hipError_t launch_unit_roots(long n, long count, float2 *out, hipStream_t s)
{
  if (count <= 0) return hipSuccess;
  long blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_unit_roots, dim3((unsigned)blocks), dim3(256), 0, s, n, count, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- registry
// Compiled configurations: (N, P).  Rows per block B makes blocks of at
// least 256 threads for short lines.
template <int N, int P>
struct RowCfg {
  static constexpr int T = N / P;
  static constexpr int B = T >= 256 ? 1 : 256 / T;
  static constexpr int NT = T * B;
};

template <int N, int P>
static void fill_cfg(RowConfig *c, int variant)
{
  using C = RowCfg<N, P>;
  c->n = N;
  c->p = P;
  c->rows_per_block = C::B;
  c->threads = C::NT;
  c->lds_bytes = (size_t)C::B * Engine<N, P>::LROW * sizeof(float2);
  c->plan = Engine<N, P>::plan;
  c->variant = variant;
}

template <int N, int P>
static hipError_t do_fft(int dir, const float2 *in, long ld_in, float2 *out, long ld_out, long rows,
                         const float2 *tw, float scale, hipStream_t s)
{
  using C = RowCfg<N, P>;
  dim3 grid((unsigned)((rows + C::B - 1) / C::B)), block(C::NT);
  if (dir < 0)
    hipLaunchKernelGGL((k_row_fft<N, P, C::B, -1>), grid, block, 0, s, in, ld_in, out, ld_out, rows, tw, scale);
  else
    hipLaunchKernelGGL((k_row_fft<N, P, C::B, +1>), grid, block, 0, s, in, ld_in, out, ld_out, rows, tw, scale);
  return hipGetLastError();
}

template <int N, int P>
static hipError_t do_fastconv(const float2 *x, long ldx, const float2 *w, long ldw, float2 *y, long ldy,
                              long rows, const float2 *tw, float scale, hipStream_t s)
{
  using C = RowCfg<N, P>;
  dim3 grid((unsigned)((rows + C::B - 1) / C::B)), block(C::NT);
  hipLaunchKernelGGL((k_row_fastconv<N, P, C::B>), grid, block, 0, s, x, ldx, w, ldw, y, ldy, rows, tw, scale);
  return hipGetLastError();
}

template <int N, int P>
static hipError_t do_fft_split(int dir, const float *ir, const float *ii, long ld_in, float *orr, float *oi,
                               long ld_out, long rows, const float2 *tw, float scale, hipStream_t s)
{
  using C = RowCfg<N, P>;
  dim3 grid((unsigned)((rows + C::B - 1) / C::B)), block(C::NT);
  if (dir < 0)
    hipLaunchKernelGGL((k_row_fft_split<N, P, C::B, -1>), grid, block, 0, s, ir, ii, ld_in, orr, oi, ld_out, rows,
                       tw, scale);
  else
    hipLaunchKernelGGL((k_row_fft_split<N, P, C::B, +1>), grid, block, 0, s, ir, ii, ld_in, orr, oi, ld_out, rows,
                       tw, scale);
  return hipGetLastError();
}

template <int N, int P>
static hipError_t do_fastconv_split(const float *xr, const float *xi, long ldx, const float2 *w, long ldw, float *yr,
                                    float *yi, long ldy, long rows, const float2 *tw, float scale, hipStream_t s)
{
  using C = RowCfg<N, P>;
  dim3 grid((unsigned)((rows + C::B - 1) / C::B)), block(C::NT);
  hipLaunchKernelGGL((k_row_fastconv_split<N, P, C::B>), grid, block, 0, s, xr, xi, ldx, w, ldw, yr, yi, ldy, rows,
                     tw, scale);
  return hipGetLastError();
}

// X(N, P): every compiled configuration; the first listed for an N is
// variant 0 (the default).
#define OVX_ROW_CONFIGS(X)                                                      \
  X(2, 2) X(4, 4) X(8, 8) X(16, 16) X(32, 16) X(64, 16) X(128, 16) X(256, 16)   \
  X(512, 16) X(1024, 16) X(2048, 16) X(4096, 16) X(8192, 16) X(16384, 16)       \
  X(3, 3) X(6, 6) X(12, 12) X(24, 24) X(48, 24) X(96, 24) X(192, 24)            \
  X(384, 24) X(768, 24) X(1536, 24) X(3072, 24) X(6144, 24) X(12288, 24)       \
  X(3072, 12) X(6144, 12) X(12288, 12)                                          \
  X(5, 5) X(10, 10) X(20, 20) X(40, 40) X(80, 40) X(160, 40) X(320, 40)         \
  X(640, 40) X(1280, 40) X(2560, 40) X(5120, 40) X(10240, 40)                 \
  X(9, 9) X(18, 18) X(36, 36) X(72, 36) X(144, 36) X(288, 36) X(576, 36)       \
  X(1152, 36) X(2304, 36) X(4608, 36) X(9216, 36)

struct Entry {
  int n, p;
  void (*fill)(RowConfig *, int);
  hipError_t (*fft)(int, const float2 *, long, float2 *, long, long, const float2 *, float, hipStream_t);
  hipError_t (*fconv)(const float2 *, long, const float2 *, long, float2 *, long, long, const float2 *, float,
                      hipStream_t);
  hipError_t (*fft_split)(int, const float *, const float *, long, float *, float *, long, long, const float2 *,
                          float, hipStream_t);
  hipError_t (*fconv_split)(const float *, const float *, long, const float2 *, long, float *, float *, long, long,
                            const float2 *, float, hipStream_t);
};

#define OVX_ENTRY(N, P)                                                                              \
  { N, P, &fill_cfg<N, P>, &do_fft<N, P>, &do_fastconv<N, P>, &do_fft_split<N, P>, &do_fastconv_split<N, P> },
static const Entry kEntries[] = { OVX_ROW_CONFIGS(OVX_ENTRY) };
static const int kNumEntries = (int)(sizeof(kEntries) / sizeof(kEntries[0]));

static const Entry *find_entry(int n, int variant)
{
  int seen = 0;
  for (int i = 0; i < kNumEntries; ++i)
    if (kEntries[i].n == n) {
      if (seen == variant) return &kEntries[i];
      ++seen;
    }
  return nullptr;
}

int row_variants(int n)
{
  int c = 0;
  for (int i = 0; i < kNumEntries; ++i)
    if (kEntries[i].n == n) ++c;
  return c;
}

bool row_config(int n, int variant, RowConfig *out)
{
  const Entry *e = find_entry(n, variant);
  if (!e) return false;
  e->fill(out, variant);
  return true;
}

hipError_t launch_row_fft(const RowConfig &cfg, int dir, const float2 *in, long ld_in, float2 *out, long ld_out,
                          long rows, const float2 *tw, float scale, hipStream_t s)
{
  const Entry *e = find_entry(cfg.n, cfg.variant);
  if (!e) return hipErrorInvalidValue;
  if (rows <= 0) return hipSuccess;
  return e->fft(dir, in, ld_in, out, ld_out, rows, tw, scale, s);
}

hipError_t launch_row_fastconv(const RowConfig &cfg, const float2 *x, long ldx, const float2 *w, long ldw,
                               float2 *y, long ldy, long rows, const float2 *tw, float scale, hipStream_t s)
{
  const Entry *e = find_entry(cfg.n, cfg.variant);
  if (!e) return hipErrorInvalidValue;
  if (rows <= 0) return hipSuccess;
  return e->fconv(x, ldx, w, ldw, y, ldy, rows, tw, scale, s);
}

hipError_t launch_row_fft_split(const RowConfig &cfg, int dir, const float *in_re, const float *in_im, long ld_in,
                                float *out_re, float *out_im, long ld_out, long rows, const float2 *tw, float scale,
                                hipStream_t s)
{
  const Entry *e = find_entry(cfg.n, cfg.variant);
  if (!e) return hipErrorInvalidValue;
  if (rows <= 0) return hipSuccess;
  return e->fft_split(dir, in_re, in_im, ld_in, out_re, out_im, ld_out, rows, tw, scale, s);
}

hipError_t launch_row_fastconv_split(const RowConfig &cfg, const float *xr, const float *xi, long ldx,
                                     const float2 *w, long ldw, float *yr, float *yi, long ldy, long rows,
                                     const float2 *tw, float scale, hipStream_t s)
{
  const Entry *e = find_entry(cfg.n, cfg.variant);
  if (!e) return hipErrorInvalidValue;
  if (rows <= 0) return hipSuccess;
  return e->fconv_split(xr, xi, ldx, w, ldw, yr, yi, ldy, rows, tw, scale, s);
}

} // namespace ovx
