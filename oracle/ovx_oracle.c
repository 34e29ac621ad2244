/*
 * ovx_oracle.c -- CPU restatement of OpenVSIP's fast-convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (openvsip_amd/, the HIP
 * library, the C-ABI) links, loads or calls this file.  The tests and
 * __graft_entry__.smoke() compare GPU results against it, and bench.py's
 * cpu_baseline leg times it (the FP32 "port" of the reference CPU path).
 *
 * Every function restates one piece of the reference and cites the
 * reference file:line it follows (paths relative to the reference tree).
 *
 * Conventions: complex data is interleaved (re, im) -- the reference's
 * default storage (src/ovxx/storage/block.hpp:27-42, interleaved unless
 * configured split).  Sign: dir = -1 forward, +1 inverse
 * (src/ovxx/signal/fft/util.hpp:43-51: exponent -1 for fft_fwd (-2),
 * +1 for fft_inv (-1)).  Inverse transforms are unnormalised; the caller
 * passes the scale (tests/fastconv.cpp:133-134 passes 1/N).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   ovo_dft_ref      reference golden vectors tests/fft_ext/data/ccfftop_f*
 *                    (committed as tests/golden/fft_ext_cc.npz), closed forms.
 *   ovo_fft_f64      brute-force ovo_dft_ref on all n <= 64 and composite
 *                    sizes, numpy.fft (pocketfft) at large n, Parseval,
 *                    inverse(forward(x)) = N x.
 *   ovo_fastconv_*   reference tests/fastconv.cpp shift-kernel closed form,
 *                    direct circular convolution (brute force) on small n,
 *                    cfg fixture produced by ovo_fastconv_ref_dft.
 *   ovo_corr_*       reference support/test/ref/corr.hpp semantics
 *                    re-derived independently in numpy (np.correlate with
 *                    explicit shifts) + reference tests/correlation.cpp
 *                    bias rules.
 *   ovo_conv_*       numpy.convolve with explicit decimation/support slices.
 *   ovo_diff_db      hand-computed values.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double re, im; } cd;
typedef struct { float re, im; } cf;

/* ------------------------------------------------------------------ */
/* The reference's builtin DFT backend.                                */
/* src/ovxx/signal/fft/dft.hpp:25-31 (sin_cos), 43-122 (dft<1,c,c,S>:  */
/* double accumulator, twiddle sin_cos(phi*k*w) with                   */
/* phi = exponent*2*pi/l, one rounding to float at the end) and        */
/* src/ovxx/signal/fft/workspace.hpp:474-478 (separate *= scale pass   */
/* in float when the backend cannot scale).                            */
/* ------------------------------------------------------------------ */
void ovo_dft_ref(const float *in, long in_stride, float *out, long out_stride,
                 long l, int dir, double scale)
{
  const double phi = dir * 2.0 * M_PI / (double)l;
  float fscale = (float)scale;
  for (long w = 0; w < l; ++w) {
    double sr = 0.0, si = 0.0;
    for (long k = 0; k < l; ++k) {
      double a = phi * (double)k * (double)w;
      double c = cos(a), s = sin(a);
      double xr = in[2 * k * in_stride], xi = in[2 * k * in_stride + 1];
      sr += xr * c - xi * s;
      si += xr * s + xi * c;
    }
    float fr = (float)sr, fi = (float)si;
    if (scale != 1.0) { fr *= fscale; fi *= fscale; }
    out[2 * w * out_stride] = fr;
    out[2 * w * out_stride + 1] = fi;
  }
}

/* dftm<c,c,A,D>::out_of_place row loop, src/ovxx/signal/fft/dft.hpp:1053-1068 */
void ovo_dftm_ref(const float *in, long ld_in, float *out, long ld_out,
                  long rows, long cols, int dir, double scale)
{
  for (long r = 0; r < rows; ++r)
    ovo_dft_ref(in + 2 * r * ld_in, 1, out + 2 * r * ld_out, 1, cols, dir, scale);
}

/* ------------------------------------------------------------------ */
/* FP64 mixed-radix FFT (accuracy oracle for sizes where the O(N^2)    */
/* DFT is too slow).  Same transform as ovo_dft_ref (sign convention   */
/* src/ovxx/signal/fft/util.hpp:43-51), computed by recursive          */
/* decimation in time over the prime factors of n with one FP64        */
/* twiddle table W_N^j = exp(dir*2*pi*i*j/N).                          */
/* ------------------------------------------------------------------ */
static long smallest_factor(long n)
{
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (long p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

typedef struct {
  long N;
  const cd *tw;  /* tw[j] = W_N^j, j < N */
  cd *scratch;   /* >= 2*N */
} fft_ctx;

static inline cd cmul(cd a, cd b)
{
  cd r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
  return r;
}

static void fft_rec(const fft_ctx *c, const cd *in, long is, cd *out, long n)
{
  if (n == 1) { out[0] = in[0]; return; }
  long p = smallest_factor(n), m = n / p;
  for (long r = 0; r < p; ++r)
    fft_rec(c, in + r * is, is * p, out + r * m, m);
  long step = c->N / n;        /* W_n^j = W_N^(j*step) */
  long pstep = c->N / p;       /* W_p^j = W_N^(j*pstep) */
  cd tbuf[64], obuf[64];
  cd *t = p <= 64 ? tbuf : (cd *)malloc(sizeof(cd) * p);
  cd *o = p <= 64 ? obuf : (cd *)malloc(sizeof(cd) * p);
  for (long k = 0; k < m; ++k) {
    for (long r = 0; r < p; ++r)
      t[r] = cmul(out[r * m + k], c->tw[(r * k * step) % c->N]);
    for (long s = 0; s < p; ++s) {
      cd acc = { 0.0, 0.0 };
      for (long r = 0; r < p; ++r) {
        cd v = cmul(t[r], c->tw[((r * s) % p) * pstep]);
        acc.re += v.re; acc.im += v.im;
      }
      o[s] = acc;
    }
    for (long s = 0; s < p; ++s) out[k + s * m] = o[s];
  }
  if (p > 64) { free(t); free(o); }
}

static cd *make_table(long N, int dir)
{
  cd *tw = (cd *)malloc(sizeof(cd) * N);
  for (long j = 0; j < N; ++j) {
    double a = dir * 2.0 * M_PI * (double)j / (double)N;
    tw[j].re = cos(a);
    tw[j].im = sin(a);
  }
  return tw;
}

/* One FP64 transform of length n; in/out interleaved double, contiguous. */
void ovo_fft_f64(const double *in, double *out, long n, int dir)
{
  fft_ctx c;
  cd *tw = make_table(n, dir);
  c.N = n; c.tw = tw; c.scratch = 0;
  fft_rec(&c, (const cd *)in, 1, (cd *)out, n);
  free(tw);
}

/* Row loop of FP64 transforms with scale (workspace.hpp:462-479). */
void ovo_fftm_f64(const double *in, long ld_in, double *out, long ld_out,
                  long rows, long n, int dir, double scale)
{
  fft_ctx c;
  cd *tw = make_table(n, dir);
  c.N = n; c.tw = tw; c.scratch = 0;
  for (long r = 0; r < rows; ++r) {
    cd *o = (cd *)(out + 2 * r * ld_out);
    fft_rec(&c, (const cd *)(in + 2 * r * ld_in), 1, o, n);
    if (scale != 1.0)
      for (long k = 0; k < n; ++k) { o[k].re *= scale; o[k].im *= scale; }
  }
  free(tw);
}

/* ------------------------------------------------------------------ */
/* FP64 fastconv oracle.                                               */
/* Y = scale * IFFT_row( W (.) FFT_row(X) ), the separate_multi form   */
/* of tests/fastconv.cpp:118-150: forward Fftm by row, vmmul<0>        */
/* (src/ovxx/expr/vmmul.hpp:53-59: row i, col j -> v(j)*m(i,j)),       */
/* inverse Fftm with scale 1/N (workspace.hpp:474-478).                */
/* x: float complex, rows x n, row stride ldx (elements).              */
/* w: float complex spectrum; ldw = 0 -> one vector for every row      */
/*    (vmmul<0>), ldw > 0 -> coefficient matrix row stride             */
/*    (tests/fastconv.cpp:215-245 fused_m_multi, elementwise product). */
/* y: double complex, rows x n contiguous.                             */
/* ------------------------------------------------------------------ */
void ovo_fastconv_f64(const float *x, long ldx, const float *w, long ldw,
                      double *y, long rows, long n, double scale)
{
  cd *twf = make_table(n, -1), *twi = make_table(n, +1);
  fft_ctx cf_ = { n, twf, 0 }, ci_ = { n, twi, 0 };
  cd *a = (cd *)malloc(sizeof(cd) * n), *b = (cd *)malloc(sizeof(cd) * n);
  for (long r = 0; r < rows; ++r) {
    for (long k = 0; k < n; ++k) {
      a[k].re = x[2 * (r * ldx + k)];
      a[k].im = x[2 * (r * ldx + k) + 1];
    }
    fft_rec(&cf_, a, 1, b, n);
    const float *wr = w + 2 * r * ldw;
    for (long k = 0; k < n; ++k) {
      cd ww = { wr[2 * k], wr[2 * k + 1] };
      b[k] = cmul(ww, b[k]);
    }
    cd *o = (cd *)(y + 2 * r * n);
    fft_rec(&ci_, b, 1, o, n);
    for (long k = 0; k < n; ++k) { o[k].re *= scale; o[k].im *= scale; }
  }
  free(a); free(b); free(twf); free(twi);
}

/* The reference's own fastconv on its builtin dft backend, step for   */
/* step: dftm forward (float out), vmmul<0> in float                   */
/* (src/vsip/impl/vmmul.hpp:79-85 row loop of v * m.row(r)), dftm      */
/* inverse (float out), then the float scale pass (workspace.hpp:474). */
/* O(N^2) per row: small sizes only (fixture generation).              */
void ovo_fastconv_ref_dft(const float *x, long ldx, const float *w, long ldw,
                          float *y, long ldy, long rows, long n, double scale)
{
  float *t = (float *)malloc(sizeof(float) * 2 * n);
  for (long r = 0; r < rows; ++r) {
    ovo_dft_ref(x + 2 * r * ldx, 1, t, 1, n, -1, 1.0);
    const float *wr = w + 2 * r * ldw;
    for (long k = 0; k < n; ++k) {
      float ar = wr[2 * k], ai = wr[2 * k + 1], br = t[2 * k], bi = t[2 * k + 1];
      t[2 * k] = ar * br - ai * bi;
      t[2 * k + 1] = ar * bi + ai * br;
    }
    ovo_dft_ref(t, 1, y + 2 * r * ldy, 1, n, +1, scale);
  }
  free(t);
}

/* ------------------------------------------------------------------ */
/* FP32 "port" of the reference CPU fastconv (the CPU baseline).       */
/* Four full-matrix passes exactly as the reference runs them          */
/* in the reference's CPU fastconv: per-row forward FFT               */
/* (src/ovxx/fftw/fftw.cpp:965-977, one execute per row, out of place),*/
/* the vmmul row loop (src/vsip/impl/vmmul.hpp:79-85), per-row inverse */
/* FFT, then the separate *= 1/N pass (workspace.hpp:474-478).         */
/* The FFT engine is this file's own FP32 Stockham FFT (radix 4/2/3),  */
/* not FFTW (FFTW is not available).                                   */
/* ------------------------------------------------------------------ */
typedef struct {
  long n;
  int nst;
  int radix[64];
  cf *tw;        /* concatenated per-stage twiddles */
  long twoff[64];
} f32plan;

static void f32plan_init(f32plan *p, long n, int dir)
{
  p->n = n; p->nst = 0;
  long m = n;
  while (m % 4 == 0) { p->radix[p->nst++] = 4; m /= 4; }
  while (m % 2 == 0) { p->radix[p->nst++] = 2; m /= 2; }
  while (m % 3 == 0) { p->radix[p->nst++] = 3; m /= 3; }
  while (m % 5 == 0) { p->radix[p->nst++] = 5; m /= 5; }
  if (m != 1) p->radix[p->nst++] = (int)m; /* generic odd radix */
  long total = 0, ns = 1;
  for (int s = 0; s < p->nst; ++s) { p->twoff[s] = total; total += ns * p->radix[s]; ns *= p->radix[s]; }
  p->tw = (cf *)malloc(sizeof(cf) * (total + 1));
  ns = 1;
  for (int s = 0; s < p->nst; ++s) {
    long R = p->radix[s];
    for (long k = 0; k < ns; ++k)
      for (long r = 0; r < R; ++r) {
        double a = dir * 2.0 * M_PI * (double)(r * k) / (double)(ns * R);
        p->tw[p->twoff[s] + k * R + r].re = (float)cos(a);
        p->tw[p->twoff[s] + k * R + r].im = (float)sin(a);
      }
    ns *= R;
  }
}

static void f32plan_free(f32plan *p) { free(p->tw); }

static inline cf fmulc(cf a, cf b)
{
  cf r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
  return r;
}

/* Stockham autosort: x -> y per stage, ping-pong; result lands in out. */
static void f32fft(const f32plan *p, const cf *in, cf *out, cf *work, int dir)
{
  long n = p->n;
  const cf *x = in;
  cf *bufs[2] = { out, work };
  /* choose the first destination so that the last stage writes out */
  int dst = (p->nst % 2 == 1) ? 0 : 1;
  long ns = 1;
  for (int s = 0; s < p->nst; ++s) {
    long R = p->radix[s];
    cf *y = bufs[dst];
    long nr = n / R;
    const cf *tw = p->tw + p->twoff[s];
    for (long j = 0; j < nr; ++j) {
      long k = j % ns;
      long base = (j / ns) * ns * R + k;
      const cf *t = tw + k * R;
      if (R == 4) {
        cf a0 = x[j], a1 = fmulc(x[j + nr], t[1]), a2 = fmulc(x[j + 2 * nr], t[2]), a3 = fmulc(x[j + 3 * nr], t[3]);
        cf s0 = { a0.re + a2.re, a0.im + a2.im }, d0 = { a0.re - a2.re, a0.im - a2.im };
        cf s1 = { a1.re + a3.re, a1.im + a3.im }, d1 = { a1.re - a3.re, a1.im - a3.im };
        /* rotate d1 by dir*i: dir=-1 -> -i*d1 */
        cf rd1 = dir < 0 ? (cf){ d1.im, -d1.re } : (cf){ -d1.im, d1.re };
        y[base] = (cf){ s0.re + s1.re, s0.im + s1.im };
        y[base + ns] = (cf){ d0.re + rd1.re, d0.im + rd1.im };
        y[base + 2 * ns] = (cf){ s0.re - s1.re, s0.im - s1.im };
        y[base + 3 * ns] = (cf){ d0.re - rd1.re, d0.im - rd1.im };
      } else if (R == 2) {
        cf a0 = x[j], a1 = fmulc(x[j + nr], t[1]);
        y[base] = (cf){ a0.re + a1.re, a0.im + a1.im };
        y[base + ns] = (cf){ a0.re - a1.re, a0.im - a1.im };
      } else {
        cf v[64];
        for (long r = 0; r < R; ++r) v[r] = r ? fmulc(x[j + r * nr], t[r]) : x[j];
        for (long q = 0; q < R; ++q) {
          float sr = 0, si = 0;
          for (long r = 0; r < R; ++r) {
            double a = dir * 2.0 * M_PI * (double)((r * q) % R) / (double)R;
            cf w = { (float)cos(a), (float)sin(a) };
            cf m = fmulc(v[r], w);
            sr += m.re; si += m.im;
          }
          y[base + q * ns] = (cf){ sr, si };
        }
      }
    }
    x = y;
    dst ^= 1;
    ns *= R;
  }
  if (p->nst == 0) memcpy(out, in, sizeof(cf) * n);
}

typedef struct {
  const cf *x; long ldx; const cf *w; long ldw; cf *y; long ldy;
  long r0, r1, n; float scale;
  const f32plan *pf, *pi;
} port_job;

static void *port_worker(void *arg)
{
  port_job *j = (port_job *)arg;
  long n = j->n;
  cf *t = (cf *)malloc(sizeof(cf) * n), *work = (cf *)malloc(sizeof(cf) * n);
  /* pass 1: forward FFT per row into y (used as the tmp matrix) */
  for (long r = j->r0; r < j->r1; ++r)
    f32fft(j->pf, j->x + r * j->ldx, j->y + r * j->ldy, work, -1);
  /* pass 2: vmmul<0> row loop */
  for (long r = j->r0; r < j->r1; ++r) {
    cf *row = j->y + r * j->ldy;
    const cf *wr = j->w + r * j->ldw;
    for (long k = 0; k < n; ++k) row[k] = fmulc(wr[k], row[k]);
  }
  /* pass 3: inverse FFT per row (out of place via t, copied back) */
  for (long r = j->r0; r < j->r1; ++r) {
    f32fft(j->pi, j->y + r * j->ldy, t, work, +1);
    memcpy(j->y + r * j->ldy, t, sizeof(cf) * n);
  }
  /* pass 4: scale */
  for (long r = j->r0; r < j->r1; ++r) {
    cf *row = j->y + r * j->ldy;
    for (long k = 0; k < n; ++k) { row[k].re *= j->scale; row[k].im *= j->scale; }
  }
  free(t); free(work);
  return 0;
}

/* Returns 0 on success.  nthreads = 1 is the reference's behaviour      */
/* (FFTW threads are off by default, m4/fft.m4:22-25; rows run serially).*/
int ovo_fastconv_port_f32(const float *x, long ldx, const float *w, long ldw,
                          float *y, long ldy, long rows, long n, double scale,
                          int nthreads)
{
  f32plan pf, pi;
  f32plan_init(&pf, n, -1);
  f32plan_init(&pi, n, +1);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > rows) nthreads = (int)rows;
  port_job jobs[256];
  pthread_t th[256];
  if (nthreads > 256) nthreads = 256;
  for (int i = 0; i < nthreads; ++i) {
    jobs[i].x = (const cf *)x; jobs[i].ldx = ldx;
    jobs[i].w = (const cf *)w; jobs[i].ldw = ldw;
    jobs[i].y = (cf *)y; jobs[i].ldy = ldy;
    jobs[i].r0 = rows * i / nthreads; jobs[i].r1 = rows * (i + 1) / nthreads;
    jobs[i].n = n; jobs[i].scale = (float)scale; jobs[i].pf = &pf; jobs[i].pi = &pi;
  }
  if (nthreads == 1) port_worker(&jobs[0]);
  else {
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], 0, port_worker, &jobs[i]);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], 0);
  }
  f32plan_free(&pf); f32plan_free(&pi);
  return 0;
}

/* FP32 forward/inverse Fftm with the port engine (used to check the    */
/* port's own accuracy against the FP64 oracle).                        */
void ovo_fftm_port_f32(const float *x, long ldx, float *y, long ldy,
                       long rows, long n, int dir, double scale)
{
  f32plan p;
  f32plan_init(&p, n, dir);
  cf *work = (cf *)malloc(sizeof(cf) * n);
  for (long r = 0; r < rows; ++r) {
    cf *o = (cf *)y + r * ldy;
    f32fft(&p, (const cf *)x + r * ldx, o, work, dir);
    if (scale != 1.0)
      for (long k = 0; k < n; ++k) { o[k].re *= (float)scale; o[k].im *= (float)scale; }
  }
  free(work);
  f32plan_free(&p);
}

/* ------------------------------------------------------------------ */
/* Correlation, src/ovxx/signal/corr.hpp:91-215 (corr_full, corr_same, */
/* corr_min): out[n] = sum_k ref[k] * conj(in[n + k - shift]), with    */
/* shift = M-1 (full), M/2 (same), 0 (min); unbiased divisors per      */
/* position; support_same uses the C-VSIPL end divisor                 */
/* (corr.hpp:65 VSIP_IMPL_CORR_CORRECT_SAME_SUPPORT_SCALING 0 and      */
/* 165-177).  support: 0 full, 1 same, 2 min                           */
/* (src/vsip/impl/signal/types.hpp:22-27).  bias: 0 biased, 1 unbiased */
/* (types.hpp:37-40).  The _ref_f32 variant accumulates in             */
/* complex<float> as the reference does (corr_accum_trait<T>::sum_type */
/* = T, corr.hpp:80-84); _f64 accumulates in double.                   */
/* ------------------------------------------------------------------ */
long ovo_corr_output_size(int support, long M, long N)
{
  /* support/test/ref/corr.hpp:24-36 */
  if (support == 0) return N + M - 1;
  if (support == 1) return N;
  return N - M + 1;
}

#define CORR_BODY(ACC)                                                         \
  long P = ovo_corr_output_size(support, M, N);                                \
  for (long n = 0; n < P; ++n) {                                               \
    ACC sr = 0, si = 0;                                                        \
    for (long k = 0; k < M; ++k) {                                             \
      long pos;                                                                \
      int ok;                                                                  \
      if (support == 0) { pos = n + k - (M - 1); ok = (n + k >= M - 1) && (n + k < N + M - 1); } \
      else if (support == 1) { pos = n + k - M / 2; ok = (n + k >= M / 2) && (n + k < N + M / 2); } \
      else { pos = n + k; ok = 1; }                                            \
      if (!ok) continue;                                                       \
      ACC rr = ref[2 * k], ri = ref[2 * k + 1];                                \
      ACC xr = in[2 * pos], xi = -in[2 * pos + 1];                             \
      sr += rr * xr - ri * xi;                                                 \
      si += rr * xi + ri * xr;                                                 \
    }                                                                          \
    if (bias == 1) {                                                           \
      ACC d;                                                                   \
      if (support == 0) {                                                      \
        if (n < M - 1) d = (ACC)(n + 1);                                       \
        else if (n >= N) d = (ACC)(N + M - 1 - n);                             \
        else d = (ACC)M;                                                       \
      } else if (support == 1) {                                               \
        if (n < M / 2) d = (ACC)(n + (M + 1) / 2);                             \
        else if (n >= N - M / 2) d = (ACC)(N - 1 + (M + 1) / 2 - n);           \
        else d = (ACC)M;                                                       \
      } else d = (ACC)M;                                                       \
      sr /= d; si /= d;                                                        \
    }                                                                          \
    out[2 * n] = sr; out[2 * n + 1] = si;                                      \
  }

void ovo_corr_ref_f32(int bias, int support, const float *ref, long M,
                      const float *in, long N, float *out)
{
  CORR_BODY(float)
}

void ovo_corr_f64(int bias, int support, const float *ref, long M,
                  const float *in, long N, double *out)
{
  CORR_BODY(double)
}

/* ------------------------------------------------------------------ */
/* Convolution, src/ovxx/signal/conv.hpp:232-330 (conv_full, conv_same,*/
/* conv_min with VSIP_IMPL_CONV_CORRECT_MIN_SUPPORT_SIZE 1, conv.hpp:36)*/
/* and output sizes conv.hpp:103-121.  Kernel symmetry expansion       */
/* conv.hpp:144-175 (sym 0 nonsym, 1 even-len-odd, 2 even-len-even     */
/* follows src/vsip/impl/signal/types.hpp symmetry_type order).        */
/* ------------------------------------------------------------------ */
long ovo_conv_output_size(int support, long M, long N, long D)
{
  if (support == 0) return ((N + M - 2) / D) + 1;
  if (support == 1) return ((N - 1) / D) + 1;
  return ((N - M + 1) / D) + ((N - M + 1) % D == 0 ? 0 : 1);
}

/* Expand coefficients by symmetry; returns kernel length.  kernel must */
/* hold 2*Mc entries.                                                   */
long ovo_conv_kernel(int sym, const float *coeff, long Mc, float *kernel)
{
  if (sym == 1) {
    for (long i = 0; i < Mc; ++i) { kernel[2 * i] = coeff[2 * i]; kernel[2 * i + 1] = coeff[2 * i + 1]; }
    for (long i = 0; i < Mc - 1; ++i) {
      kernel[2 * (Mc + i)] = coeff[2 * (Mc - 2 - i)];
      kernel[2 * (Mc + i) + 1] = coeff[2 * (Mc - 2 - i) + 1];
    }
    return 2 * Mc - 1;
  }
  if (sym == 2) {
    for (long i = 0; i < Mc; ++i) { kernel[2 * i] = coeff[2 * i]; kernel[2 * i + 1] = coeff[2 * i + 1]; }
    for (long i = 0; i < Mc; ++i) {
      kernel[2 * (Mc + i)] = coeff[2 * (Mc - 1 - i)];
      kernel[2 * (Mc + i) + 1] = coeff[2 * (Mc - 1 - i) + 1];
    }
    return 2 * Mc;
  }
  memcpy(kernel, coeff, sizeof(float) * 2 * Mc);
  return Mc;
}

#define CONV_BODY(ACC)                                                         \
  long P = ovo_conv_output_size(support, M, N, D);                             \
  for (long n = 0; n < P; ++n) {                                               \
    ACC sr = 0, si = 0;                                                        \
    for (long k = 0; k < M; ++k) {                                             \
      long pos;                                                                \
      if (support == 0) {                                                      \
        if (!(n * D >= k && n * D - k < N)) continue;                          \
        pos = n * D - k;                                                       \
      } else if (support == 1) {                                               \
        if (!(n * D + M / 2 >= k && n * D + M / 2 - k < N)) continue;          \
        pos = n * D + M / 2 - k;                                               \
      } else {                                                                 \
        pos = n * D + (M - 1) - k;                                             \
      }                                                                        \
      ACC cr = kernel[2 * k], ci = kernel[2 * k + 1];                          \
      ACC xr = in[2 * pos], xi = in[2 * pos + 1];                              \
      sr += cr * xr - ci * xi;                                                 \
      si += cr * xi + ci * xr;                                                 \
    }                                                                          \
    out[2 * n] = sr; out[2 * n + 1] = si;                                      \
  }

void ovo_conv_ref_f32(int support, const float *kernel, long M, const float *in,
                      long N, long D, float *out)
{
  CONV_BODY(float)
}

void ovo_conv_f64(int support, const float *kernel, long M, const float *in,
                  long N, long D, double *out)
{
  CONV_BODY(double)
}

/* ------------------------------------------------------------------ */
/* test::diff, support/test/diff.hpp:62-96: max over i of              */
/* 10*log10(|a_i-b_i|^2 / (2*max(max|a|^2, max|b|^2))), -201 where     */
/* |a_i-b_i|^2 < 1e-20, 201 if either input holds a NaN.               */
/* a: float complex, b: double complex, n elements.                    */
/* ------------------------------------------------------------------ */
double ovo_diff_db(const float *a, const double *b, long n)
{
  double m1 = 0, m2 = 0;
  for (long i = 0; i < n; ++i) {
    if (isnan(a[2 * i]) || isnan(a[2 * i + 1]) || isnan(b[2 * i]) || isnan(b[2 * i + 1])) return 201.0;
    double v1 = (double)a[2 * i] * a[2 * i] + (double)a[2 * i + 1] * a[2 * i + 1];
    double v2 = b[2 * i] * b[2 * i] + b[2 * i + 1] * b[2 * i + 1];
    if (v1 > m1) m1 = v1;
    if (v2 > m2) m2 = v2;
  }
  double refmax = m1 > m2 ? m1 : m2;
  double maxsum = -HUGE_VAL;
  for (long i = 0; i < n; ++i) {
    double dr = a[2 * i] - b[2 * i], di = a[2 * i + 1] - b[2 * i + 1];
    double d2 = dr * dr + di * di;
    double s = d2 < 1e-20 ? -201.0 : 10.0 * log10(d2 / (2.0 * refmax));
    if (s > maxsum) maxsum = s;
  }
  return n ? maxsum : -201.0;
}
