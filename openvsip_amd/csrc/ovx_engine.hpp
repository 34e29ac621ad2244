// ovx_engine.hpp -- register-resident Stockham FFT engine for one line of
// N complex<float> values handled by T = N/P threads (P values per thread).
//
// Layout contract ("slot form"): thread t holds in v[q] the element at
// position t + T*q of the line, q in [0, P).  Global loads and stores in slot
// form are coalesced (consecutive threads touch consecutive elements).
//
// Stage s has radix R_s and Ns = R_0*...*R_{s-1}.  Butterfly j (j < N/R) reads
// x[j + r*N/R] -- exactly slots q = m + r*(P/R) of thread t where
// j = t + m*T -- twiddles them by W_{Ns R}^{r (j mod Ns)}, runs a radix-R DFT
// and writes output r to position (j/Ns)*Ns*R + (j mod Ns) + r*Ns (Stockham
// autosort: natural order in, natural order out).  The last stage
// (Ns*R = N) writes back to the very positions it read, so its outputs stay in
// registers in slot form: a forward transform can feed an inverse transform
// (fast convolution) without an exchange, and the final stage stores straight
// to global memory.  Every other stage exchanges through LDS: one write of P
// values, a barrier, one read of P values in slot form.
//
// LDS layout of a line (LdsLayout): gfx950's LDS has 64 banks of 4 bytes and
// serves a 64-bit access for 32 lanes per cycle, so the 32 lanes of each
// half-wave should hit 32 distinct 8-byte slots.  Lines handled by at least
// 16 threads insert one pad element every 32 (slot = pos + pos/32), shorter
// ones one every 16; both keep the slot-form reads t + T q at constant
// offsets from pos(t), so the exchanges need no per-access address math.
#pragma once
#include "ovx_device.hpp"

namespace ovx {

// ---------------------------------------------------------------- plans
struct RadixPlan {
  int nst;
  int radix[16];
  int ns[16];
  int twoff[16]; // offset of stage s in the concatenated twiddle table
  int twtotal;
};

constexpr bool divides(int a, int b) { return b % a == 0; }

// Radix sequence for length N with P values per thread: radices must divide
// P; larger radices first (the first stage needs no twiddles, so the stage
// with the most twiddle multiplies goes there).
constexpr RadixPlan make_plan(int N, int P)
{
  RadixPlan p{};
  p.nst = 0;
  int rem = N;
  const int cand[6] = { 16, 8, 4, 5, 3, 2 };
  while (rem > 1) {
    int chosen = 0;
    for (int i = 0; i < 6; ++i) {
      int r = cand[i];
      if (divides(r, rem) && divides(r, P)) { chosen = r; break; }
    }
    if (chosen == 0) { p.nst = -1; return p; }
    p.radix[p.nst++] = chosen;
    rem /= chosen;
  }
  int ns = 1, off = 0;
  for (int s = 0; s < p.nst; ++s) {
    p.ns[s] = ns;
    p.twoff[s] = off;
    if (ns > 1) off += (p.radix[s] - 1) * ns;
    ns *= p.radix[s];
  }
  p.twtotal = off;
  return p;
}

constexpr int lds_row_elems(int N) { return N >= 16 ? N + N / 16 : N; }

template <int N, int P>
struct LdsLayout {
  static constexpr bool kPad32 = (N / P) >= 16 && N >= 64;
  static constexpr int elems = kPad32 ? N + (N + 31) / 32 : lds_row_elems(N);
  static constexpr int period = kPad32 ? 32 : 16; // one pad element per `period`
  __host__ __device__ static constexpr int idx(int pos) { return kPad32 ? pos + (pos >> 5) : pos + (pos >> 4); }
  // idx(pos + c) == idx(pos) + step(c) for every pos >= 0 when c is a
  // multiple of the period: the pad count of the sum splits exactly, so
  // accesses at such distances share one base address and differ by a
  // constant offset (the compiler does not find this by itself).
  __host__ __device__ static constexpr int step(int c) { return c + c / period; }
};

// Pointer into LDS with the address space spelled out (for code that is not
// inlined into the kernel owning the __shared__ array), and element access
// helpers that work for it and for a plain float2 pointer alike.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) f2v lds_f2v;
typedef lds_f2v *lds_ptr;
__device__ __forceinline__ lds_ptr as_lds(float2 *p) { return (lds_ptr)p; }
__device__ __forceinline__ void lds_st(float2 *p, int i, float2 v) { p[i] = v; }
__device__ __forceinline__ float2 lds_ld(const float2 *p, int i) { return p[i]; }
__device__ __forceinline__ void lds_st(lds_ptr p, int i, float2 v) { p[i] = f2v{ v.x, v.y }; }
__device__ __forceinline__ float2 lds_ld(lds_ptr p, int i)
{
  f2v t = p[i];
  return make_float2(t.x, t.y);
}
// Synchronisation between the LDS write and read halves of an exchange.
// Block scope: every wave of the workgroup takes part in the line.
// Wave scope: the line lives in one wavefront, whose LDS operations complete
// in order; waiting for them (lgkmcnt 0) is enough, no workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void exchange_sync()
{
  if constexpr (WAVE) {
    __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0), vmcnt/expcnt untouched
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// ---------------------------------------------------------------- engine
// Each stage's twiddles are loaded during the previous stage's LDS exchange
// (after its writes, before the synchronisation), into registers the
// exchanged values have just left, so that the stage finds them loaded.
// Loaded just before each multiply instead, they were left to the scheduler
// to batch, which it does not do in kernels whose register budget is tight:
// it then waited on every load in turn (the four-step passes and the batch
// tiles did, 15 to 80 times per transform).
template <int N, int P, bool WAVE = false>
struct Engine {
  static constexpr int T = N / P;
  static constexpr RadixPlan plan = make_plan(N, P);
  static constexpr int NST = plan.nst;
  using Layout = LdsLayout<N, P>;
  static constexpr int LROW = Layout::elems;
  static_assert(N % P == 0, "P must divide N");
  static_assert(plan.nst > 0 || N == 1, "no radix plan for N with this P");

  // Whether the slots a thread (0 <= t < T) reads, positions t + T q, sit at
  // idx(t) + step(T q): no pad boundary falls between them differently for
  // different t.  Then one base address and constant offsets serve all P
  // reads.  Checked exhaustively (N cases).
  static constexpr bool read_split()
  {
    for (int t = 0; t < T; ++t)
      for (int q = 0; q < P; ++q)
        if (Layout::idx(t + T * q) != Layout::idx(t) + Layout::step(T * q)) return false;
    return true;
  }
  // The same for the R slots stage S writes from base (j / Ns) Ns R + j % Ns,
  // j = 0 .. N / R - 1, at distances r Ns.
  template <int S>
  static constexpr bool write_split()
  {
    constexpr int R = plan.radix[S], Ns = plan.ns[S];
    for (int j = 0; j < N / R; ++j) {
      const int base = (j / Ns) * Ns * R + (j % Ns);
      for (int r = 0; r < R; ++r)
        if (Layout::idx(base + r * Ns) != Layout::idx(base) + Layout::step(r * Ns)) return false;
    }
    return true;
  }

  // The twiddles of stage S for this thread's M = P / R butterflies,
  // w[m (R-1) + r - 1] = tw[twoff_s + (r-1) Ns + (t + m T) % Ns].
  template <int S>
  __device__ __forceinline__ static void load_tw(int t, rsrc_t tw, float2 (&w)[P])
  {
    if constexpr (S < NST) {
      constexpr int R = plan.radix[S];
      constexpr int Ns = plan.ns[S];
      constexpr int M = P / R;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int k = (t + m * T) % Ns;
#pragma unroll
          for (int r = 1; r < R; ++r)
            w[m * (R - 1) + r - 1] = bload(tw, (uint32_t)k * 8u, (uint32_t)(plan.twoff[S] + (r - 1) * Ns) * 8u);
        }
      }
    }
  }

  // One stage.  tw: concatenated forward twiddle table (FP32 values of FP64
  // W_{Ns R}^{r k}, layout tw[twoff_s + (r-1)*Ns + k]); w: this stage's
  // twiddles (loaded by the previous stage, or by the caller for stage 0).
  template <int S, int DIR, typename LP>
  __device__ __forceinline__ static void stage(float2 (&v)[P], LP lds, int t, rsrc_t tw, float2 (&w)[P])
  {
    constexpr int R = plan.radix[S];
    constexpr int Ns = plan.ns[S];
    constexpr int M = P / R;
    constexpr bool last = (S == NST - 1);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      float2 a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[m + r * M];
      const int j = t + m * T;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = twmul<DIR>(a[r], w[m * (R - 1) + r - 1]);
      }
      Dft<R, DIR>::run(a);
      if constexpr (last) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[m + r * M] = a[r];
      } else {
        const int base = (j / Ns) * Ns * R + (j % Ns);
        if constexpr (write_split<S>()) {
          const int b = Layout::idx(base);
#pragma unroll
          for (int r = 0; r < R; ++r) lds_st(lds, b + Layout::step(r * Ns), a[r]);
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) lds_st(lds, Layout::idx(base + r * Ns), a[r]);
        }
      }
    }
    if constexpr (!last) {
      load_tw<S + 1>(t, tw, w);
      exchange_sync<WAVE>();
      if constexpr (read_split()) {
        const int b = Layout::idx(t);
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = lds_ld(lds, b + Layout::step(T * q));
      } else {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = lds_ld(lds, Layout::idx(t + T * q));
      }
      exchange_sync<WAVE>();
    }
  }

  // One stage on its own (a caller that runs a single stage).
  template <int S, int DIR, typename LP>
  __device__ __forceinline__ static void stage(float2 (&v)[P], LP lds, int t, rsrc_t tw)
  {
    float2 w[P];
    load_tw<S>(t, tw, w);
    stage<S, DIR>(v, lds, t, tw, w);
  }

  template <int S, int DIR, typename LP>
  __device__ __forceinline__ static void stages_from(float2 (&v)[P], LP lds, int t, rsrc_t tw, float2 (&w)[P])
  {
    if constexpr (S < NST) {
      stage<S, DIR>(v, lds, t, tw, w);
      stages_from<S + 1, DIR>(v, lds, t, tw, w);
    }
  }

  // Full transform of the line held in slot form; result in slot form.
  // LP: pointer to the line's LDS region (float2*, or an LDS address-space
  // pointer when the caller is not inlined into the kernel that owns it).
  template <int DIR, typename LP>
  __device__ __forceinline__ static void fft(float2 (&v)[P], LP lds, int t, rsrc_t tw)
  {
    float2 w[P];
    load_tw<0>(t, tw, w); // stage 0 has Ns = 1: nothing
    stages_from<0, DIR>(v, lds, t, tw, w);
  }
};

} // namespace ovx
