/*
 * ovxhip.h -- C-ABI of the MI355X (gfx950) fast-convolution backend for
 * OpenVSIP.
 *
 * Every entry point returns an ovx_status (0 = success) and never throws.
 * The message of the last failure on the calling thread is available from
 * ovx_last_error().  Complex data is complex<float>, interleaved (re, im),
 * passed as void*.  Strides and sizes are in elements, as in the reference's
 * backend interface.  Unless a plan is created with OVX_FLAG_HOST_POINTERS,
 * data pointers are device pointers (HBM-resident), and work is enqueued on
 * the given hipStream_t (passed as void*; NULL = the default stream) without
 * blocking the host.  A plan is not safe for concurrent calls (as the
 * reference's Fft/Fftm objects, which own their plans and buffers); separate
 * plans are independent.  A device-pointer fastconv call that writes the
 * plan's scratch (the spectrum permuted for a hierarchical or four-step
 * kernel) on another stream than the plan's previous such call is refused
 * (OVX_ERR_STATE) while that stream still has work queued.
 *
 * Reference interfaces replaced (paths relative to the OpenVSIP tree):
 *   fft_fwd / fft_inv            src/ovxx/signal/fft/backend.hpp:19-22
 *   fftm_backend<c,c,A,D>        src/ovxx/signal/fft/backend.hpp:696-784
 *   fft_backend<1,c,c,S>         src/ovxx/signal/fft/backend.hpp:131-204
 *   Fftm/Fft construction        src/vsip/impl/signal/fft.hpp:75-161,
 *                                src/ovxx/fftw/fftw.cpp:904-1002
 *   fused fastconv (be::fc_expr) src/ovxx/dispatcher/tags.hpp:33-34,
 *                                src/ovxx/assign_fwd.hpp:58 (slot, empty
 *                                in the reference)
 *   Correlation                  src/ovxx/signal/corr.hpp:611-719
 *   Convolution                  src/ovxx/signal/conv.hpp:964-1134
 */
#ifndef OVXHIP_H
#define OVXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
typedef enum {
  OVX_OK = 0,
  OVX_ERR_INVALID = 1,     /* bad argument (null pointer, size mismatch)   */
  OVX_ERR_UNSUPPORTED = 2, /* valid request this backend does not handle:
                              the dispatcher should try the next backend
                              (rt_valid() == false / "unimplemented")     */
  OVX_ERR_NOMEM = 3,       /* device allocation failed (std::bad_alloc)    */
  OVX_ERR_DEVICE = 4,      /* HIP runtime error                            */
  OVX_ERR_BUDGET = 5,      /* admission refused: over the HBM budget       */
  OVX_ERR_TIMEOUT = 6,     /* serving: request not completed in time       */
  OVX_ERR_STATE = 7        /* call not valid in the object's current state */
} ovx_status;

/* Transform directions: the reference's values (backend.hpp:19-22). */
#define OVX_FFT_FWD (-2)
#define OVX_FFT_INV (-1)

/* Axis of an Fftm: OVX_ROW transforms each row, OVX_COL each column
   (vsip::row = 0, vsip::col = 1). */
#define OVX_ROW 0
#define OVX_COL 1

/* Plan flags. */
#define OVX_FLAG_HOST_POINTERS 0x1u /* data pointers are host memory; the call
                                       stages through device buffers and
                                       returns after the result is back    */
/* Kernel variant of a transform length (0 = the default or the autotuner's
   cached choice; v >= 1 forces compiled variant v-1, see
   ovx_length_variants). */
#define OVX_FLAG_VARIANT(v) (((unsigned)(v) & 0xffu) << 8)
/* Fastconv precision variant (no reference counterpart): the radix-16 DFT
   stages run on the MFMA cores as BF16 contractions with split (hi + lo)
   operands, hi.hi + hi.lo + lo.hi, accumulated in FP32.  n = 16384, device
   pointers, one spectrum vector.  Its error against FP64 is reported beside
   the FP32 path's; it does not meet the FP32 parity tolerance. */
#define OVX_FLAG_BF16X3 0x2u

/* Correlation / convolution enums (src/vsip/impl/signal/types.hpp). */
#define OVX_SUPPORT_FULL 0
#define OVX_SUPPORT_SAME 1
#define OVX_SUPPORT_MIN 2
#define OVX_BIASED 0
#define OVX_UNBIASED 1
#define OVX_NONSYM 0
#define OVX_SYM_EVEN_LEN_ODD 1
#define OVX_SYM_EVEN_LEN_EVEN 2

typedef struct ovx_plan_s *ovx_plan;

/* ---- library ---------------------------------------------------------- */
/* Select the HIP device for plans created afterwards on this thread.
   Replaces the per-thread library initialisation of vsipl
   (src/ovxx/library.cpp:51-106). */
int ovx_init(int device);
const char *ovx_last_error(void);
const char *ovx_version(void);
/* Largest transform length the backend accepts (rt_valid upper bound). */
long ovx_max_length(void);
/* 1 if `n` is a transform length this backend handles, else 0: 2^k, 3*2^k,
   5*2^k and 9*2^k up to 16384, and 2^k from 32768 to ovx_max_length(). */
int ovx_supported_length(long n);
/* Number of compiled kernel variants for transform length n. */
int ovx_length_variants(long n);
/* Autotuner (no reference counterpart; the reference's FFTW backend plans
   with FFTW_MEASURE, src/ovxx/fftw/fft.cpp:44-52): time every variant of
   length n on a rows x n fastconv, record the fastest under (n, rows rounded
   up to a power of two) in the cache file $OVX_TUNE_CACHE (default
   ~/.cache/ovxhip/tune.txt).  Plans created without an explicit variant use
   the cached choice. */
int ovx_autotune_fastconv(long n, long rows, int *best_variant, double *best_us);
/* Cached variant for (n, rows), or -1. */
int ovx_tuned_variant(long n, long rows);
/* Batch tiles (no reference counterpart): for a job of rows x n inside a
   batch or the server, time every tile body the persistent kernels hold for
   length n (the hierarchical body and the register-Stockham body with the
   batch's values per thread) as a one-job batch replayed from its graph, and
   record the fastest under (n, rows rounded up to a power of two) in the
   same cache.  ovx_batch_commit tunes the keys it lacks (unless the
   environment sets OVX_NO_BATCH_TUNE); the server uses the cached choices.
   A length with one tile body returns OVX_ERR_UNSUPPORTED. */
int ovx_autotune_batch(long n, long rows, int *best_variant, double *best_us);
/* Cached batch tile variant for (n, rows), or -1. */
int ovx_tuned_batch_variant(long n, long rows);

/* ---- Fftm (multiple 1-D transforms) ----------------------------------- */
/* Create a plan for a rows x cols Fftm along `axis` in `direction`, result
   multiplied by `scale` inside the kernel (supports_scale() == true, so the
   reference workspace skips its separate scale pass, workspace.hpp:474).
   Replaces Evaluator<op::fftm,...>::exec / fftm ctor
   (src/ovxx/fftw/fft.hpp:139-145, fftw.cpp:916-919). */
int ovx_fftm_plan_create(ovx_plan *plan, long rows, long cols, int axis,
                         int direction, float scale, unsigned flags);

/* fftm_backend::out_of_place (backend.hpp:760-770): strides s0 (between
   rows) and s1 (between columns); the stride along the transform axis must
   be 1 (the layout query_layout requests, backend.hpp:713-723). */
int ovx_fftm_out_of_place(ovx_plan plan, const void *in, long in_s0, long in_s1,
                          void *out, long out_s0, long out_s1, long rows,
                          long cols, void *stream);

/* fftm_backend::in_place (backend.hpp:735-745). */
int ovx_fftm_in_place(ovx_plan plan, void *inout, long s0, long s1, long rows,
                      long cols, void *stream);

/* ---- real Fftm (real <-> complex, multiple 1-D transforms) ------------- */
/* Plan for a rows x cols *real* domain, transforms of even length n = cols
   (OVX_ROW) or rows (OVX_COL); the complex side holds n/2 + 1 values per line.
   OVX_FFT_FWD: real -> complex, replaces fftm_backend<T, complex<T>, A,
   fft_fwd> (backend.hpp:600-646; FFTW r2c in fftw.cpp:778-843).
   OVX_FFT_INV: complex -> real, replaces fftm_backend<complex<T>, T, A,
   fft_inv> (backend.hpp:648-694; FFTW c2r in fftw.cpp:845-903): unnormalised
   like FFTW's c2r, the imaginary parts of frequencies 0 and n/2 ignored.
   The result is multiplied by `scale` in the kernel. */
int ovx_fftm_real_plan_create(ovx_plan *plan, long rows, long cols, int axis,
                              int direction, float scale, unsigned flags);
/* out_of_place of those backends: strides in elements of each side's type
   (float on the real side, complex<float> on the complex side), unit stride
   along the transform axis, even real line stride, 8-byte aligned real data;
   rows x cols is the real domain. */
int ovx_fftm_real_out_of_place(ovx_plan plan, const void *in, long in_s0,
                               long in_s1, void *out, long out_s0, long out_s1,
                               long rows, long cols, void *stream);

/* ---- Fft (one 1-D transform) ------------------------------------------- */
/* vsip::Fft<const_Vector,C,C,dir> construction (fft.hpp:75-103). */
int ovx_fft_plan_create(ovx_plan *plan, long n, int direction, float scale,
                        unsigned flags);
/* fft_backend<1,c,c,S>::out_of_place (backend.hpp:181-190); in_stride and
   out_stride must be 1. */
int ovx_fft_out_of_place(ovx_plan plan, const void *in, long in_stride,
                         void *out, long out_stride, long n, void *stream);
/* fft_backend<1,c,c,S>::in_place (backend.hpp:161-170). */
int ovx_fft_in_place(ovx_plan plan, void *inout, long stride, long n,
                     void *stream);

/* ---- fused fast convolution ------------------------------------------- */
/* Plan for Y = scale * inv_fftm( W (.) fwd_fftm(X) ) over rows x n, the
   single expression  out = inv_fftm(vmmul<0>(w, for_fftm(in)))
   (tests/fastconv.cpp:200-205, benchmarks/fastconv.cpp:797-801) and its
   coefficient-matrix form  out = inv_fftm(C * for_fftm(in))
   (tests/fastconv.cpp:215-245).  scale is the inverse Fftm's scale
   (normally 1/n). */
int ovx_fastconv_plan_create(ovx_plan *plan, long rows, long n, float scale,
                             unsigned flags);
/* x, y: rows x n at row strides ldx, ldy (unit column stride).  w: the
   replica spectrum (frequency domain), one vector when ldw == 0
   (vmmul<0>), else a coefficient matrix with row stride ldw.  In place
   (x == y) is allowed. */
int ovx_fastconv_exec(ovx_plan plan, const void *x, long ldx, const void *w,
                      long ldw, void *y, long ldy, long rows, void *stream);

/* ---- fused fast convolution without allocation -------------------------- */
/* No reference counterpart.  For callers that cannot allocate device memory
   -- code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc: ovx_fastconv_exec on device pointers (same arguments, same
   results bit for bit) with the call's scratch -- the spectrum that a
   hierarchical or four-step plan permutes into its kernel's order -- in a
   fixed device array of OVX_NOALLOC_SCRATCH_BYTES that the library holds on
   each GPU until the process ends, instead of a plan buffer made on first
   use.  The call allocates nothing (no hipMalloc, no hipFree; the plan's
   tables were made by ovx_fastconv_plan_create).  A call that would need
   more scratch (a four-step plan with a coefficient matrix of many rows, or
   a spectrum longer than 2^19 points) fails with OVX_ERR_UNSUPPORTED;
   host-pointer plans are refused.  Every call on one GPU shares that GPU's
   array and leaves its last contents there (each call writes all of the
   scratch it reads before reading it): calls that overlap on one GPU --
   other streams, other threads, other plans -- race, and must not be
   issued. */
#define OVX_NOALLOC_SCRATCH_BYTES (4u << 20)
int ovx_fastconv_exec_noalloc(ovx_plan plan, const void *x, long ldx,
                              const void *w, long ldw, void *y, long ldy,
                              long rows, void *stream);

/* ---- vector-matrix multiply -------------------------------------------- */
/* out = alpha * vmmul<dim>(v, m) over rows x cols, device pointers:
   dim 0: out(r, c) = alpha * v(c) * m(r, c);  dim 1: out(r, c) = alpha *
   v(r) * m(r, c).  Replaces Evaluator<op::assign<2>, be::op_expr,
   void(LHS &, expr::Vmmul<D, V, M> const &)> (src/vsip/impl/vmmul.hpp:50-106;
   element rule src/ovxx/expr/vmmul.hpp:53-59) in one pass.  v_complex /
   m_complex: 1 for complex<float> data, 0 for float; out is complex<float>
   when either is complex, else float (Promotion<T1, T2>).  Strides in
   elements of each operand's type; any layout (row- or column-major,
   strided subviews).  out may alias m (in place); complex products are
   computed as (ar*br - ai*bi, ar*bi + ai*br). */
int ovx_vmmul(int dim, int v_complex, const void *v, long v_stride,
              int m_complex, const void *m, long m_s0, long m_s1, void *out,
              long out_s0, long out_s1, long rows, long cols, float alpha,
              void *stream);

/* ---- replica spectra kept resident on the device ---------------------- */
/* Transform a time-domain replica of length m (m <= n), zero-padded to n,
   once with the forward Fft (tests/fastconv.cpp:20-26 t2f) and keep the
   spectrum on the device; when is_spectrum != 0, `replica` already holds the
   n-point spectrum.  `replica` is a host pointer.  Returns an id. */
int ovx_replica_create(long n, const void *replica, long m, int is_spectrum,
                       int *id);
int ovx_replica_destroy(int id);
/* Device pointer to the spectrum (n complex<float>), for ovx_fastconv_exec. */
int ovx_replica_spectrum(int id, const void **spectrum, long *n);
/* Fused fastconv against a resident replica: the spectrum was prepared once,
   at ovx_replica_create, in the layout the plan's kernel reads (permuted,
   with the 1/n scale folded in), so a call reads X, writes Y and reads the
   prepared spectrum -- nothing else.  Same result as ovx_fastconv_exec with
   the replica's spectrum.  Device pointers only. */
int ovx_fastconv_exec_replica(ovx_plan plan, const void *x, long ldx, int replica,
                              void *y, long ldy, long rows, void *stream);

/* ---- split-complex storage --------------------------------------------- */
/* The split overloads of fftm_backend / fft_backend (backend.hpp:741-783,
   131-204; the FFTW backend's split DFTs, fftw.cpp:945-1001): real and
   imaginary parts in two float arrays with the same strides (OpenVSIP's
   split_complex format), on the plans of the interleaved entry points.
   Results equal the interleaved call's on the same values bit for bit.  The
   kernels read and write split arrays themselves: the fastconv and complex
   transform kernels of every length (the four-step path in its first and
   last pass) and the real transforms' passes on their complex side. */
int ovx_fftm_split_out_of_place(ovx_plan plan, const void *in_re,
                                const void *in_im, long in_s0, long in_s1,
                                void *out_re, void *out_im, long out_s0,
                                long out_s1, long rows, long cols, void *stream);
int ovx_fftm_split_in_place(ovx_plan plan, void *re, void *im, long s0, long s1,
                            long rows, long cols, void *stream);
int ovx_fft_split_out_of_place(ovx_plan plan, const void *in_re, const void *in_im,
                               long in_stride, void *out_re, void *out_im,
                               long out_stride, long n, void *stream);
int ovx_fft_split_in_place(ovx_plan plan, void *re, void *im, long stride, long n,
                           void *stream);
/* Real Fftm with a split complex side (backend.hpp:600-694 split overloads;
   FFTW split r2c / c2r, fftw.cpp:778-903): r2c on a forward plan (real in,
   split out), c2r on an inverse plan (split in, real out); strides as in
   ovx_fftm_real_out_of_place, the complex side's in floats of each part. */
int ovx_fftm_r2c_split(ovx_plan plan, const void *in, long in_s0, long in_s1,
                       void *out_re, void *out_im, long out_s0, long out_s1,
                       long rows, long cols, void *stream);
int ovx_fftm_c2r_split(ovx_plan plan, const void *in_re, const void *in_im,
                       long in_s0, long in_s1, void *out, long out_s0,
                       long out_s1, long rows, long cols, void *stream);
/* ovx_fastconv_exec on split X, W and Y (W one vector when ldw == 0).  Not
   for OVX_FLAG_BF16X3 plans. */
int ovx_fastconv_split_exec(ovx_plan plan, const void *x_re, const void *x_im,
                            long ldx, const void *w_re, const void *w_im, long ldw,
                            void *y_re, void *y_im, long ldy, long rows,
                            void *stream);

/* ---- storage: a block's data in host memory and in HBM ----------------- */
/* The reference's storage_manager (src/ovxx/storage/manager.hpp:207-360)
   keeps a block's data in host memory plus, when configured, a device copy
   (src/ovxx/opencl/storage.hpp; commented out of the manager), each valid or
   stale.  ovx_storage is that manager with a HIP location: a block's data
   stays in HBM across calls and crosses the host link only when the other
   location asks for it.  Not safe for concurrent calls, like a plan. */
typedef struct ovx_storage_s *ovx_storage;
#define OVX_LOC_HOST 0
#define OVX_LOC_DEVICE 1
#define OVX_ACCESS_READ 1       /* const ptr(where): sync the location      */
#define OVX_ACCESS_WRITE 2      /* contents will be overwritten (dda::out):
                                   no copy; every other location goes stale */
#define OVX_ACCESS_READ_WRITE 3 /* ptr(where): sync, then the others go
                                   stale (manager.hpp:247-263)             */
/* `bytes` of storage.  With user_host != NULL the application's array is the
   host location and holds the data (user storage, admitted); otherwise the
   library allocates pinned host memory when the host first asks. */
int ovx_storage_create(ovx_storage *storage, size_t bytes, void *user_host);
int ovx_storage_destroy(ovx_storage storage);
/* A pointer valid in `where` for `access`.  A stale location is refreshed
   first with a copy queued on `stream`: the device pointer is ready for work
   queued on that stream afterwards; a host pointer is returned after its
   copy has landed.  Device work that wrote the data must have been queued
   on `stream` (or be complete) before a host access. */
int ovx_storage_ptr(ovx_storage storage, int where, int access, void **ptr,
                    void *stream);
/* User storage (manager.hpp:207-243): release brings the host array up to
   date and hands it to the application (pointer requests then fail with
   OVX_ERR_STATE); admit takes it back -- with update != 0 the application may
   have changed it, so the device copy is stale. */
int ovx_storage_release(ovx_storage storage, void *stream);
int ovx_storage_admit(ovx_storage storage, int update);
/* Which locations are valid (bit 1 << OVX_LOC_*) and the bytes copied each
   way so far. */
int ovx_storage_info(ovx_storage storage, unsigned *valid, uint64_t *h2d_bytes,
                     uint64_t *d2h_bytes);

/* ---- batches of uneven fastconv jobs ------------------------------------ */
/* No reference counterpart: the reference runs one Fftm object per shape
   (benchmarks/fastconv.cpp loops one size at a time).  A batch packs jobs of
   different n and row counts (each against a resident replica) into
   persistent-kernel launches over (instance, rows) tile queues (one of
   1024-thread workgroups for rows of 12288 and 16384 points, one of
   512-thread workgroups for shorter rows); the rows of jobs on the
   four-step path run in chunks of about 128 MiB, each chunk one launch per
   pass and line length (one line group per workgroup), so that a chunk
   stays in the Infinity Cache between its passes.  The launches are
   captured once into a hipGraph at commit and replayed by every exec.  Each job's result is bitwise the same as
   ovx_fastconv_exec_replica on the same job.  ovx_batch_launches returns
   the number of kernel launches per replay.  A committed batch holds the
   device addresses of its jobs' buffers and replica spectra: they must stay
   valid (and the replicas not destroyed) until the batch is destroyed. */
typedef struct ovx_batch_s *ovx_batch;
int ovx_batch_create(ovx_batch *batch);
int ovx_batch_add_fastconv(ovx_batch batch, long n, long rows, const void *x,
                           long ldx, int replica, void *y, long ldy);
int ovx_batch_commit(ovx_batch batch, void *stream);
int ovx_batch_exec(ovx_batch batch, void *stream);
/* The plain path it is compared with: one standalone launch per job. */
int ovx_batch_exec_plain(ovx_batch batch, void *stream);
long ovx_batch_tiles(ovx_batch batch);
int ovx_batch_launches(ovx_batch batch);
/* Kernel variant job `job` (in order of ovx_batch_add_fastconv) runs in the
   committed batch -- the same numbering as OVX_FLAG_VARIANT(v + 1), so a
   standalone plan made with that flag runs the same device body -- or -1. */
int ovx_batch_job_variant(ovx_batch batch, long job);
int ovx_batch_destroy(ovx_batch batch);

/* ---- continuous-batching server ----------------------------------------- */
/* No reference counterpart: serving pulse-compression jobs, standalone
   forward Fftms and correlations as they arrive.  A scheduler thread admits
   queued requests (in arrival order, while their device working sets fit the
   HBM budget) into a free slot -- a stream with a hipGraph of persistent
   tile-queue launches captured at creation -- and replays it; requests that
   arrive meanwhile go into the next free slot, so batches overlap on the
   device.  Data pointers are device pointers owned by the caller and must
   stay valid until the request completes, and the replicas and correlation
   references a request names until it completes.  The server's streams do
   not wait for work the caller queued on other streams: x must be written,
   and nothing else may still be writing y, when the request is submitted.
   plain != 0 runs every
   request as its own standalone launch instead (the per-request path).
   ovx_server_destroy waits for the batches already launched; requests still
   queued are dropped. */
typedef struct ovx_server_s *ovx_server;
typedef struct {
  int slots;          /* batches in flight at once (>= 1)                    */
  int max_jobs;       /* requests per batch                                  */
  long max_tiles;     /* tile-table capacity per queue (0 = 65536)           */
  long hbm_budget_mb; /* admission budget for request working sets           */
  long workspace_mb;  /* four-step Fftm workspace per slot                   */
  int plain;          /* 1: one standalone launch per request                */
} ovx_server_config;

#define OVX_REQ_FASTCONV 1 /* ref = replica id; y = fastconv of x            */
#define OVX_REQ_FFTM 2     /* forward Fftm by row, scale 1                    */
#define OVX_REQ_CORR 3     /* ref = correlation reference id; x has length n */
typedef struct {
  double t_arrival;   /* seconds from the start of a trace (run_trace only)  */
  int kind;
  long n, rows;
  const void *x;
  long ldx;
  void *y;
  long ldy;
  int ref;
  int bias;           /* correlation: OVX_BIASED / OVX_UNBIASED             */
} ovx_request;

int ovx_server_create(ovx_server *server, const ovx_server_config *cfg);
int ovx_server_submit(ovx_server server, const ovx_request *req, unsigned long *ticket);
/* Wait for a request; its latency (completion - submission) in seconds. */
int ovx_server_wait(ovx_server server, unsigned long ticket, double timeout_s,
                    double *latency_s);
/* Submit reqs[i] at t_arrival[i] (real time) and wait for all; per-request
   latencies and the time from the first arrival to the last completion. */
int ovx_server_run_trace(ovx_server server, const ovx_request *reqs, long count,
                         double *latency_s, double *makespan_s);
int ovx_server_destroy(ovx_server server);
/* A correlation reference prepared once for serving: reference of length m
   (host data) against inputs of length n with the given support. */
int ovx_corr_reference_create(int support, long m, long n, const void *ref, int *id);
int ovx_corr_reference_destroy(int id);

/* ---- Correlation (FFT-based, overlap-save) ----------------------------- */
/* vsip::Correlation<const_Vector, support, complex<float>> of a reference of
   length m against an input of length n (corr.hpp:616-628 sizes, output
   length from support/test/ref/corr.hpp:24-36). */
int ovx_corr_plan_create(ovx_plan *plan, int support, long m, long n,
                         unsigned flags);
long ovx_corr_output_size(int support, long m, long n);
/* Correlation::correlate (corr.hpp:663-719):
   out[i] = sum_k ref[k] * conj(in[i + k - shift]), unbiased divisors per
   position when bias == OVX_UNBIASED. */
int ovx_corr_exec(ovx_plan plan, int bias, const void *ref, const void *in,
                  void *out, void *stream);

/* ---- Convolution (FFT-based, overlap-save) ----------------------------- */
/* vsip::Convolution<const_Vector, symmetry, support, complex<float>> with
   coefficients of length mc (expanded by symmetry, conv.hpp:144-175),
   input length n and decimation d (output size conv.hpp:103-121). */
int ovx_conv_plan_create(ovx_plan *plan, int symmetry, int support,
                         const void *coeff, long mc, long n, long decimation,
                         unsigned flags);
long ovx_conv_output_size(int support, long m, long n, long decimation);
int ovx_conv_exec(ovx_plan plan, const void *in, void *out, void *stream);

/* ---- plans ------------------------------------------------------------- */
int ovx_plan_destroy(ovx_plan plan);
/* Name of the kernel path the plan runs (diagnostics, as the reference's
   dispatch diagnostics print the chosen evaluator, dispatcher/diagnostics.hpp). */
const char *ovx_plan_name(ovx_plan plan);
/* Bytes of device memory the plan holds (twiddles, workspaces). */
size_t ovx_plan_device_bytes(ovx_plan plan);

#ifdef __cplusplus
}
#endif

#endif /* OVXHIP_H */
