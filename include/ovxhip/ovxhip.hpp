// ovxhip.hpp -- C++ host-side mirror of the reference objects on the
// fast-convolution path, over the C-ABI of include/ovxhip.h.
//
// Error mapping follows the reference's conventions
// (src/ovxx/support.hpp:138-142 unimplemented, std::bad_alloc on plan
// failure src/ovxx/fftw/fftw.cpp:97-102, std::runtime_error otherwise):
//   OVX_ERR_UNSUPPORTED -> ovxhip::unimplemented (a std::runtime_error)
//   OVX_ERR_NOMEM       -> std::bad_alloc
//   OVX_ERR_INVALID     -> std::invalid_argument
//   anything else       -> ovxhip::error (a std::runtime_error)
//
// Header-only; link with libovxhip.so.  Data pointers are device pointers
// unless the object is created with OVX_FLAG_HOST_POINTERS.
#ifndef OVXHIP_HPP
#define OVXHIP_HPP

#include <complex>
#include <cstddef>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <utility>

#include "../ovxhip.h"

namespace ovxhip {

class error : public std::runtime_error {
public:
  error(int code, std::string const &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

private:
  int code_;
};

class unimplemented : public error {
public:
  explicit unimplemented(std::string const &msg) : error(OVX_ERR_UNSUPPORTED, msg) {}
};

inline void check(int status)
{
  if (status == OVX_OK) return;
  std::string msg = ovx_last_error();
  switch (status) {
  case OVX_ERR_UNSUPPORTED: throw unimplemented(msg);
  case OVX_ERR_NOMEM: throw std::bad_alloc();
  case OVX_ERR_INVALID: throw std::invalid_argument(msg);
  default: throw error(status, msg);
  }
}

typedef std::complex<float> cscalar;
// split-complex storage: (real parts, imaginary parts), as the reference's
// std::pair<T*, T*> overloads (src/ovxx/signal/fft/backend.hpp:741-783)
typedef std::pair<float *, float *> split_ptr;
typedef std::pair<float const *, float const *> const_split_ptr;

// Owns one plan handle.
class plan_holder {
public:
  plan_holder() = default;
  plan_holder(plan_holder const &) = delete;
  plan_holder &operator=(plan_holder const &) = delete;
  ~plan_holder()
  {
    if (h_) ovx_plan_destroy(h_);
  }
  ovx_plan get() const { return h_; }
  ovx_plan *out() { return &h_; }
  char const *name() const { return ovx_plan_name(h_); }

private:
  ovx_plan h_ = nullptr;
};

// vsip::Fftm<complex<float>, complex<float>, A, D> (fft.hpp:134-161) with the
// backend operations of fftm_backend (backend.hpp:696-784).  A: OVX_ROW or
// OVX_COL; D: OVX_FFT_FWD or OVX_FFT_INV.
template <int A, int D>
class Fftm {
public:
  Fftm(long rows, long cols, float scale, unsigned flags = 0) : rows_(rows), cols_(cols)
  {
    check(ovx_fftm_plan_create(p_.out(), rows, cols, A, D, scale, flags));
  }
  // the backend scales inside the kernel (workspace.hpp:474 skips its pass)
  static bool supports_scale() { return true; }
  void out_of_place(cscalar const *in, long in_s0, long in_s1, cscalar *out, long out_s0, long out_s1,
                    long rows, long cols, void *stream = nullptr)
  {
    check(ovx_fftm_out_of_place(p_.get(), in, in_s0, in_s1, out, out_s0, out_s1, rows, cols, stream));
  }
  void in_place(cscalar *inout, long s0, long s1, long rows, long cols, void *stream = nullptr)
  {
    check(ovx_fftm_in_place(p_.get(), inout, s0, s1, rows, cols, stream));
  }
  void out_of_place(const_split_ptr in, long in_s0, long in_s1, split_ptr out, long out_s0, long out_s1, long rows,
                    long cols, void *stream = nullptr)
  {
    check(ovx_fftm_split_out_of_place(p_.get(), in.first, in.second, in_s0, in_s1, out.first, out.second, out_s0,
                                      out_s1, rows, cols, stream));
  }
  void in_place(split_ptr inout, long s0, long s1, long rows, long cols, void *stream = nullptr)
  {
    check(ovx_fftm_split_in_place(p_.get(), inout.first, inout.second, s0, s1, rows, cols, stream));
  }
  // dense row-major (A = row) or column-major (A = col) convenience call
  void operator()(cscalar const *in, cscalar *out, void *stream = nullptr)
  {
    if (A == OVX_ROW) out_of_place(in, cols_, 1, out, cols_, 1, rows_, cols_, stream);
    else out_of_place(in, 1, rows_, out, 1, rows_, rows_, cols_, stream);
  }
  char const *name() const { return p_.name(); }

private:
  long rows_, cols_;
  plan_holder p_;
};

// Real Fftm: vsip::Fftm<float, complex<float>, A, fft_fwd> (r2c) and
// vsip::Fftm<complex<float>, float, A, fft_inv> (c2r) over a rows x cols real
// domain; the complex side holds n/2 + 1 values per line.
template <int A, int D>
class FftmReal {
public:
  FftmReal(long rows, long cols, float scale, unsigned flags = 0) : rows_(rows), cols_(cols)
  {
    check(ovx_fftm_real_plan_create(p_.out(), rows, cols, A, D, scale, flags));
  }
  static bool supports_scale() { return true; }
  // D == OVX_FFT_FWD: in real, out complex; OVX_FFT_INV: in complex, out real
  void out_of_place(void const *in, long in_s0, long in_s1, void *out, long out_s0, long out_s1, long rows,
                    long cols, void *stream = nullptr)
  {
    check(ovx_fftm_real_out_of_place(p_.get(), in, in_s0, in_s1, out, out_s0, out_s1, rows, cols, stream));
  }
  // split complex side: r2c (D == OVX_FFT_FWD) and c2r (OVX_FFT_INV)
  void out_of_place(float const *in, long in_s0, long in_s1, split_ptr out, long out_s0, long out_s1, long rows,
                    long cols, void *stream = nullptr)
  {
    check(ovx_fftm_r2c_split(p_.get(), in, in_s0, in_s1, out.first, out.second, out_s0, out_s1, rows, cols, stream));
  }
  void out_of_place(const_split_ptr in, long in_s0, long in_s1, float *out, long out_s0, long out_s1, long rows,
                    long cols, void *stream = nullptr)
  {
    check(ovx_fftm_c2r_split(p_.get(), in.first, in.second, in_s0, in_s1, out, out_s0, out_s1, rows, cols, stream));
  }
  char const *name() const { return p_.name(); }

private:
  long rows_, cols_;
  plan_holder p_;
};

// vsip::Fft<const_Vector, complex<float>, complex<float>, D> (fft.hpp:75-103).
template <int D>
class Fft {
public:
  Fft(long n, float scale, unsigned flags = 0) : n_(n) { check(ovx_fft_plan_create(p_.out(), n, D, scale, flags)); }
  static bool supports_scale() { return true; }
  void out_of_place(cscalar const *in, long in_s, cscalar *out, long out_s, void *stream = nullptr)
  {
    check(ovx_fft_out_of_place(p_.get(), in, in_s, out, out_s, n_, stream));
  }
  void in_place(cscalar *inout, long s, void *stream = nullptr)
  {
    check(ovx_fft_in_place(p_.get(), inout, s, n_, stream));
  }
  void out_of_place(const_split_ptr in, long in_s, split_ptr out, long out_s, void *stream = nullptr)
  {
    check(ovx_fft_split_out_of_place(p_.get(), in.first, in.second, in_s, out.first, out.second, out_s, n_, stream));
  }
  void in_place(split_ptr inout, long s, void *stream = nullptr)
  {
    check(ovx_fft_split_in_place(p_.get(), inout.first, inout.second, s, n_, stream));
  }
  char const *name() const { return p_.name(); }

private:
  long n_;
  plan_holder p_;
};

// out = inv_fftm(vmmul<0>(w, fwd_fftm(in))) with the inverse scale, fused
// (tests/fastconv.cpp:181-213, benchmarks/fastconv.cpp:740-849).
class Fastconv {
public:
  Fastconv(long rows, long n, float scale, unsigned flags = 0) : rows_(rows), n_(n)
  {
    check(ovx_fastconv_plan_create(p_.out(), rows, n, scale, flags));
  }
  // w: spectrum vector (ldw = 0) or coefficient matrix (row stride ldw)
  void operator()(cscalar const *x, long ldx, cscalar const *w, long ldw, cscalar *y, long ldy, long rows,
                  void *stream = nullptr)
  {
    check(ovx_fastconv_exec(p_.get(), x, ldx, w, ldw, y, ldy, rows, stream));
  }
  void operator()(cscalar const *x, cscalar const *w, cscalar *y, void *stream = nullptr)
  {
    (*this)(x, n_, w, 0, y, n_, rows_, stream);
  }
  void operator()(const_split_ptr x, long ldx, const_split_ptr w, long ldw, split_ptr y, long ldy, long rows,
                  void *stream = nullptr)
  {
    check(ovx_fastconv_split_exec(p_.get(), x.first, x.second, ldx, w.first, w.second, ldw, y.first, y.second, ldy,
                                  rows, stream));
  }
  char const *name() const { return p_.name(); }

private:
  long rows_, n_;
  plan_holder p_;
};

// The Fastconv plan for (rows, n, scale, flags) of the calling thread, made
// on first use, so that an evaluator that sees one expression shape per call
// -- the VSIPL++ fused fastconv -- does not rebuild twiddle tables and
// workspaces every time.  A plan is not safe for concurrent calls, so each
// thread has its own.  The cache is one process-wide table under a mutex
// (no thread-local storage: a header included after OpenVSIP's, whose
// ovxx/thread.hpp redefines the thread_local keyword, must not depend on
// it), keeps at most kMaxCachedPlans plans and drops the least recently used
// one beyond that.  The caller shares ownership, so a plan dropped from the
// cache lives until the call that holds it returns.
class fastconv_cache {
public:
  static std::size_t const kMaxCachedPlans = 16;
  static std::shared_ptr<Fastconv> get(long rows, long n, float scale, unsigned flags)
  {
    uint32_t bits;
    std::memcpy(&bits, &scale, sizeof bits);
    key_type key(std::this_thread::get_id(), rows, n, bits, flags);
    table &t = instance();
    {
      std::lock_guard<std::mutex> lock(t.mutex);
      for (auto it = t.entries.begin(); it != t.entries.end(); ++it)
        if (it->first == key) {
          t.entries.splice(t.entries.begin(), t.entries, it);
          return t.entries.front().second;
        }
    }
    // plan creation (device tables) runs outside the lock
    std::shared_ptr<Fastconv> plan(new Fastconv(rows, n, scale, flags));
    std::lock_guard<std::mutex> lock(t.mutex);
    t.entries.emplace_front(key, plan);
    while (t.entries.size() > kMaxCachedPlans) t.entries.pop_back();
    return plan;
  }
  static std::size_t size()
  {
    table &t = instance();
    std::lock_guard<std::mutex> lock(t.mutex);
    return t.entries.size();
  }
  static void clear()
  {
    table &t = instance();
    std::lock_guard<std::mutex> lock(t.mutex);
    t.entries.clear();
  }

private:
  typedef std::tuple<std::thread::id, long, long, uint32_t, unsigned> key_type;
  struct table {
    std::mutex mutex;
    std::list<std::pair<key_type, std::shared_ptr<Fastconv> > > entries; // most recent first
  };
  static table &instance()
  {
    static table t;
    return t;
  }
};

inline std::shared_ptr<Fastconv> cached_fastconv(long rows, long n, float scale, unsigned flags = 0)
{
  return fastconv_cache::get(rows, n, scale, flags);
}

// out = alpha * vmmul<dim>(v, m) on device data (ovx_vmmul): complex<float>
// or float operands, strides in elements of each operand's type.
inline void vmmul(int dim, cscalar const *v, long v_stride, cscalar const *m, long m_s0, long m_s1, cscalar *out,
                  long out_s0, long out_s1, long rows, long cols, float alpha = 1.0f, void *stream = nullptr)
{
  check(ovx_vmmul(dim, 1, v, v_stride, 1, m, m_s0, m_s1, out, out_s0, out_s1, rows, cols, alpha, stream));
}
inline void vmmul(int dim, float const *v, long v_stride, cscalar const *m, long m_s0, long m_s1, cscalar *out,
                  long out_s0, long out_s1, long rows, long cols, float alpha = 1.0f, void *stream = nullptr)
{
  check(ovx_vmmul(dim, 0, v, v_stride, 1, m, m_s0, m_s1, out, out_s0, out_s1, rows, cols, alpha, stream));
}
inline void vmmul(int dim, cscalar const *v, long v_stride, float const *m, long m_s0, long m_s1, cscalar *out,
                  long out_s0, long out_s1, long rows, long cols, float alpha = 1.0f, void *stream = nullptr)
{
  check(ovx_vmmul(dim, 1, v, v_stride, 0, m, m_s0, m_s1, out, out_s0, out_s1, rows, cols, alpha, stream));
}
inline void vmmul(int dim, float const *v, long v_stride, float const *m, long m_s0, long m_s1, float *out,
                  long out_s0, long out_s1, long rows, long cols, float alpha = 1.0f, void *stream = nullptr)
{
  check(ovx_vmmul(dim, 0, v, v_stride, 0, m, m_s0, m_s1, out, out_s0, out_s1, rows, cols, alpha, stream));
}

// vsip::Correlation<const_Vector, R, complex<float>> (corr.hpp:55-122).
class Correlation {
public:
  Correlation(int support, long ref_size, long input_size, unsigned flags = 0)
    : support_(support), m_(ref_size), n_(input_size)
  {
    check(ovx_corr_plan_create(p_.out(), support, ref_size, input_size, flags));
  }
  long reference_size() const { return m_; }
  long input_size() const { return n_; }
  long output_size() const { return ovx_corr_output_size(support_, m_, n_); }
  int support() const { return support_; }
  void operator()(int bias, cscalar const *ref, cscalar const *in, cscalar *out, void *stream = nullptr)
  {
    check(ovx_corr_exec(p_.get(), bias, ref, in, out, stream));
  }

private:
  int support_;
  long m_, n_;
  plan_holder p_;
};

// vsip::Convolution<const_Vector, S, R, complex<float>> (conv.hpp:55-108).
class Convolution {
public:
  Convolution(int symmetry, int support, cscalar const *coeff, long coeff_size, long input_size,
              long decimation = 1, unsigned flags = 0)
    : support_(support), n_(input_size), d_(decimation)
  {
    m_ = symmetry == OVX_SYM_EVEN_LEN_ODD ? 2 * coeff_size - 1
         : symmetry == OVX_SYM_EVEN_LEN_EVEN ? 2 * coeff_size : coeff_size;
    check(ovx_conv_plan_create(p_.out(), symmetry, support, coeff, coeff_size, input_size, decimation, flags));
  }
  long kernel_size() const { return m_; }
  long input_size() const { return n_; }
  long output_size() const { return ovx_conv_output_size(support_, m_, n_, d_); }
  long decimation() const { return d_; }
  void operator()(cscalar const *in, cscalar *out, void *stream = nullptr)
  {
    check(ovx_conv_exec(p_.get(), in, out, stream));
  }

private:
  int support_;
  long m_, n_, d_;
  plan_holder p_;
};

// A block's data kept in host memory and in HBM (storage_manager,
// src/ovxx/storage/manager.hpp:207-360, with a HIP device location).  T is
// the element type; `user` (optional) is the application's array, admitted.
template <typename T>
class Storage {
public:
  explicit Storage(long size, T *user = nullptr) : size_(size)
  {
    check(ovx_storage_create(&h_, (size_t)size * sizeof(T), user));
  }
  Storage(Storage const &) = delete;
  Storage &operator=(Storage const &) = delete;
  ~Storage() { ovx_storage_destroy(h_); }
  long size() const { return size_; }
  // ptr(where) and const ptr(where) of the manager, plus write-only access
  T *ptr(int where, int access = OVX_ACCESS_READ_WRITE, void *stream = nullptr)
  {
    void *p = nullptr;
    check(ovx_storage_ptr(h_, where, access, &p, stream));
    return static_cast<T *>(p);
  }
  T const *ptr(int where, void *stream = nullptr) const
  {
    void *p = nullptr;
    check(ovx_storage_ptr(h_, where, OVX_ACCESS_READ, &p, stream));
    return static_cast<T const *>(p);
  }
  T *device(int access = OVX_ACCESS_READ_WRITE, void *stream = nullptr) { return ptr(OVX_LOC_DEVICE, access, stream); }
  T *host(int access = OVX_ACCESS_READ_WRITE, void *stream = nullptr) { return ptr(OVX_LOC_HOST, access, stream); }
  void release(void *stream = nullptr) { check(ovx_storage_release(h_, stream)); }
  void admit(bool update = true) { check(ovx_storage_admit(h_, update ? 1 : 0)); }
  unsigned valid() const
  {
    unsigned v = 0;
    check(ovx_storage_info(h_, &v, nullptr, nullptr));
    return v;
  }

private:
  long size_;
  ovx_storage h_ = nullptr;
};

} // namespace ovxhip

#endif
