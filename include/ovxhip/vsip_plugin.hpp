// vsip_plugin.hpp -- registers the gfx950 backend with OpenVSIP's dispatcher.
//
// Include this header in a translation unit that includes <vsip/signal.hpp>
// (and therefore a configured OpenVSIP tree) and link libovxhip.so.  It adds
// Evaluator specialisations on be::user, which is first in every relevant
// dispatch list:
//   op::fft   list  src/ovxx/signal/fft.hpp:35-49
//   op::fftm  list  src/ovxx/signal/fft.hpp:51-64
//   assign<1>, assign<2> lists  src/ovxx/assign_fwd.hpp:49-62
//   op::corr  list  src/vsip/impl/signal/corr.hpp:26-38
//   op::conv  list  src/vsip/impl/signal/conv.hpp:27-38
// so existing vsip::Fft / vsip::Fftm / vmmul code selects the HIP kernels
// without source changes; rt_valid() returns false for a shape the backend
// does not handle, and the dispatcher falls through to the next backend
// (src/ovxx/dispatcher/dispatcher.hpp:96-130).
//
// Evaluators added:
//   * Fftm (complex, real -> complex, complex -> real) and 1-D Fft factories;
//   * the fused fast convolution in the three single-expression forms the
//     reference writes (tests/fastconv.cpp:186-245):
//       out = inv_fftm(vmmul<0>(w, fwd_fftm(in)))
//       out = inv_fftm(C * fwd_fftm(in)),  out = inv_fftm(fwd_fftm(in) * C)
//     and the per-row vector form  out.row(r) = inv_fft(w * fwd_fft(in.row(r)))
//     (tests/fastconv.cpp:160-179), each one fused launch;
//   * on Resident blocks (vsip_resident.hpp): vmmul<0>/<1> and the by-value
//     Fftm  out = fftm(in)  on HBM-resident data;
//   * Correlation and Convolution backends (complex<float>, 1-D).
//
// Data reach the backend through OpenVSIP's DDA as host pointers, so those
// plans are created with OVX_FLAG_HOST_POINTERS (the library stages the data
// through device buffers).  Blocks of type Resident<D, T> keep their data in
// HBM between calls: the evaluators above take device pointers from them and
// run plans made for device data, with no host copy.  (A by-reference
// Fftm call, fftm(in, out), reaches the backend through the reference's FFT
// workspace and DDA, i.e. through host memory, on any block type.)
//
// The backends ask DDA for the storage format of OpenVSIP's dense blocks on
// their complex sides (interleaved, or split when OpenVSIP is configured
// with split complex storage), so neither format is converted on the host;
// split data take the library's split entry points.  The fused evaluators
// request unit stride along rows (DDA copies a column-major or
// column-strided operand into that layout and back).
//
// tests/test_plugin_compile.py instantiates every evaluator here against the
// reference headers (compile only: it runs nothing).
#ifndef OVXHIP_VSIP_PLUGIN_HPP
#define OVXHIP_VSIP_PLUGIN_HPP

#include <memory>

#include <ovxx/dispatch.hpp>
#include <ovxx/dda.hpp>
#include <ovxx/expr/binary.hpp>
#include <ovxx/expr/operations.hpp>
#include <ovxx/expr/unary.hpp>
#include <ovxx/expr/vmmul.hpp>
#include <ovxx/signal/fft.hpp>
#include <ovxx/signal/fft/backend.hpp>
#include <ovxx/signal/fft/functor.hpp>
#include <ovxx/signal/conv.hpp>
#include <ovxx/signal/corr.hpp>

#include "ovxhip.hpp"
#include "vsip_resident.hpp"

namespace ovxx
{
namespace ovxhip_backend
{
using vsip::complex;
typedef complex<float> ctype;

// fftm_backend<c,c,A,D> on the gfx950 kernels (src/ovxx/signal/fft/backend.hpp:696-784).
template <int A, int D>
class fftm : public signal::fft::fftm_backend<ctype, ctype, A, D>
{
  typedef signal::fft::fftm_backend<ctype, ctype, A, D> base_type;
  static int const dir = D == fft_fwd ? OVX_FFT_FWD : OVX_FFT_INV;
  static int const axis = A == vsip::col ? OVX_COL : OVX_ROW;

public:
  fftm(vsip::Domain<2> const &dom, float scale)
    : plan_(dom[0].size(), dom[1].size(), scale, OVX_FLAG_HOST_POINTERS), scale_(scale), rows_(dom[0].size()),
      cols_(dom[1].size())
  {}
  virtual char const *name() { return "ovxhip::fftm"; }
  // the kernels scale: workspace.hpp:474 skips its separate pass
  virtual bool supports_scale() { return true; }
  // dense, unit stride along the transform axis (the default), in the
  // storage format OpenVSIP's dense blocks use -- interleaved or split, as
  // configured -- so that DDA converts nothing, as the FFTW backend asks
  // (src/ovxx/fftw/fft.cpp:24, fftw.cpp:298-307)
  virtual void query_layout(Rt_layout<2> &rtl_inout)
  {
    base_type::query_layout(rtl_inout);
    rtl_inout.storage_format = vsip::Dense<1, ctype>::storage_format;
  }
  virtual void query_layout(Rt_layout<2> &rtl_in, Rt_layout<2> &rtl_out)
  {
    base_type::query_layout(rtl_in, rtl_out);
    rtl_in.storage_format = rtl_out.storage_format = vsip::Dense<1, ctype>::storage_format;
  }
  virtual void in_place(ctype *inout, stride_type s0, stride_type s1, length_type rows, length_type cols)
  {
    plan_.in_place(reinterpret_cast<ovxhip::cscalar *>(inout), s0, s1, rows, cols);
  }
  virtual void in_place(std::pair<float *, float *> inout, stride_type s0, stride_type s1, length_type rows,
                        length_type cols)
  {
    plan_.in_place(inout, s0, s1, rows, cols);
  }
  virtual void out_of_place(ctype *in, stride_type in_s0, stride_type in_s1, ctype *out, stride_type out_s0,
                            stride_type out_s1, length_type rows, length_type cols)
  {
    plan_.out_of_place(reinterpret_cast<ovxhip::cscalar const *>(in), in_s0, in_s1,
                       reinterpret_cast<ovxhip::cscalar *>(out), out_s0, out_s1, rows, cols);
  }
  virtual void out_of_place(std::pair<float *, float *> in, stride_type in_s0, stride_type in_s1,
                            std::pair<float *, float *> out, stride_type out_s0, stride_type out_s1, length_type rows,
                            length_type cols)
  {
    plan_.out_of_place(ovxhip::const_split_ptr(in.first, in.second), in_s0, in_s1, out, out_s0, out_s1, rows, cols);
  }
  float scale() const { return scale_; }
  // the same transform on device pointers (Resident blocks), made on first use
  ovxhip::Fftm<axis, dir> &device_plan() const
  {
    std::lock_guard<std::mutex> lock(device_mutex_);
    if (!device_plan_) device_plan_.reset(new ovxhip::Fftm<axis, dir>(rows_, cols_, scale_));
    return *device_plan_;
  }

private:
  ovxhip::Fftm<axis, dir> plan_;
  float scale_;
  long rows_ = 0, cols_ = 0;
  mutable std::mutex device_mutex_;
  mutable std::unique_ptr<ovxhip::Fftm<axis, dir> > device_plan_;
};

// fft_backend<1,c,c,S> (src/ovxx/signal/fft/backend.hpp:131-204).
template <int S>
class fft : public signal::fft::fft_backend<1, ctype, ctype, S>
{
  typedef signal::fft::fft_backend<1, ctype, ctype, S> base_type;
  static int const dir = S == fft_fwd ? OVX_FFT_FWD : OVX_FFT_INV;

public:
  fft(vsip::Domain<1> const &dom, float scale) : plan_(dom.size(), scale, OVX_FLAG_HOST_POINTERS), scale_(scale) {}
  float scale() const { return scale_; }
  virtual char const *name() { return "ovxhip::fft"; }
  virtual bool supports_scale() { return true; }
  virtual void query_layout(Rt_layout<1> &rtl_inout)
  {
    base_type::query_layout(rtl_inout);
    rtl_inout.storage_format = vsip::Dense<1, ctype>::storage_format;
  }
  virtual void query_layout(Rt_layout<1> &rtl_in, Rt_layout<1> &rtl_out)
  {
    base_type::query_layout(rtl_in, rtl_out);
    rtl_in.storage_format = rtl_out.storage_format = vsip::Dense<1, ctype>::storage_format;
  }
  virtual void in_place(ctype *inout, stride_type s, length_type)
  {
    plan_.in_place(reinterpret_cast<ovxhip::cscalar *>(inout), s);
  }
  virtual void in_place(std::pair<float *, float *> inout, stride_type s, length_type)
  {
    plan_.in_place(inout, s);
  }
  virtual void out_of_place(ctype *in, stride_type in_s, ctype *out, stride_type out_s, length_type)
  {
    plan_.out_of_place(reinterpret_cast<ovxhip::cscalar const *>(in), in_s, reinterpret_cast<ovxhip::cscalar *>(out),
                       out_s);
  }
  virtual void out_of_place(std::pair<float *, float *> in, stride_type in_s, std::pair<float *, float *> out,
                            stride_type out_s, length_type)
  {
    plan_.out_of_place(ovxhip::const_split_ptr(in.first, in.second), in_s, out, out_s);
  }

private:
  ovxhip::Fft<dir> plan_;
  float scale_;
};

// fftm_backend<float, complex<float>, A, fft_fwd> (backend.hpp:600-646):
// real -> complex through the half-length complex kernels.
template <int A>
class fftm_r2c : public signal::fft::fftm_backend<float, ctype, A, fft_fwd>
{
  typedef signal::fft::fftm_backend<float, ctype, A, fft_fwd> base_type;
  static int const axis = A == vsip::col ? OVX_COL : OVX_ROW;

public:
  fftm_r2c(vsip::Domain<2> const &dom, float scale)
    : plan_(dom[0].size(), dom[1].size(), scale, OVX_FLAG_HOST_POINTERS)
  {}
  virtual char const *name() { return "ovxhip::fftm_r2c"; }
  virtual bool supports_scale() { return true; }
  // the complex side in the dense blocks' storage format (no DDA conversion)
  virtual void query_layout(Rt_layout<2> &rtl_in, Rt_layout<2> &rtl_out)
  {
    base_type::query_layout(rtl_in, rtl_out);
    rtl_out.storage_format = vsip::Dense<1, ctype>::storage_format;
  }
  virtual void out_of_place(float *in, stride_type in_s0, stride_type in_s1, ctype *out, stride_type out_s0,
                            stride_type out_s1, length_type rows, length_type cols)
  {
    plan_.out_of_place(in, in_s0, in_s1, out, out_s0, out_s1, rows, cols);
  }
  virtual void out_of_place(float *in, stride_type in_s0, stride_type in_s1, std::pair<float *, float *> out,
                            stride_type out_s0, stride_type out_s1, length_type rows, length_type cols)
  {
    plan_.out_of_place(static_cast<float const *>(in), in_s0, in_s1, out, out_s0, out_s1, rows, cols);
  }

private:
  ovxhip::FftmReal<axis, OVX_FFT_FWD> plan_;
};

// fftm_backend<complex<float>, float, A, fft_inv> (backend.hpp:648-694).
template <int A>
class fftm_c2r : public signal::fft::fftm_backend<ctype, float, A, fft_inv>
{
  typedef signal::fft::fftm_backend<ctype, float, A, fft_inv> base_type;
  static int const axis = A == vsip::col ? OVX_COL : OVX_ROW;

public:
  fftm_c2r(vsip::Domain<2> const &dom, float scale)
    : plan_(dom[0].size(), dom[1].size(), scale, OVX_FLAG_HOST_POINTERS)
  {}
  virtual char const *name() { return "ovxhip::fftm_c2r"; }
  virtual bool supports_scale() { return true; }
  virtual void query_layout(Rt_layout<2> &rtl_in, Rt_layout<2> &rtl_out)
  {
    base_type::query_layout(rtl_in, rtl_out);
    rtl_in.storage_format = vsip::Dense<1, ctype>::storage_format;
  }
  virtual void out_of_place(ctype *in, stride_type in_s0, stride_type in_s1, float *out, stride_type out_s0,
                            stride_type out_s1, length_type rows, length_type cols)
  {
    plan_.out_of_place(in, in_s0, in_s1, out, out_s0, out_s1, rows, cols);
  }
  virtual void out_of_place(std::pair<float *, float *> in, stride_type in_s0, stride_type in_s1, float *out,
                            stride_type out_s0, stride_type out_s1, length_type rows, length_type cols)
  {
    plan_.out_of_place(ovxhip::const_split_ptr(in.first, in.second), in_s0, in_s1, out, out_s0, out_s1, rows, cols);
  }

private:
  ovxhip::FftmReal<axis, OVX_FFT_INV> plan_;
};

inline bool length_ok(length_type n) { return ovx_supported_length((long)n) != 0; }
// a real transform of even length n runs on the complex kernels of n/2
inline bool real_length_ok(length_type n) { return n % 2 == 0 && n >= 4 && length_ok(n / 2); }

// Layouts the fused evaluators ask DDA for: unit stride along rows (a
// row-major view; DDA copies any other layout in and out), in the storage
// format of OpenVSIP's dense blocks.
typedef vsip::Layout<2, vsip::row2_type, vsip::unit_stride, vsip::Dense<2, ctype>::storage_format> rows_layout;
typedef vsip::Layout<1, vsip::row1_type, vsip::unit_stride, vsip::Dense<1, ctype>::storage_format> line_layout;

// The fused fastconv on what DDA hands over: interleaved complex pointers,
// or (real, imaginary) pairs when the blocks use split storage.
inline void fastconv_call(ovxhip::Fastconv &fc, ctype const *x, stride_type ldx, ctype const *w, stride_type ldw,
                          ctype *y, stride_type ldy, length_type rows)
{
  fc(reinterpret_cast<ovxhip::cscalar const *>(x), ldx, reinterpret_cast<ovxhip::cscalar const *>(w), ldw,
     reinterpret_cast<ovxhip::cscalar *>(y), ldy, rows);
}
template <typename PX, typename PW>
void fastconv_call(ovxhip::Fastconv &fc, std::pair<PX *, PX *> x, stride_type ldx, std::pair<PW *, PW *> w,
                   stride_type ldw, std::pair<float *, float *> y, stride_type ldy, length_type rows)
{
  fc(ovxhip::const_split_ptr(x.first, x.second), ldx, ovxhip::const_split_ptr(w.first, w.second), ldw, y, ldy,
     rows);
}

// out = scale * inv(w (.) fwd(in)) over rows x cols, the whole expression in
// one launch.  W is a vector (one spectrum for every row) or a rows x cols
// coefficient matrix.  Resident operands run on their HBM copies with a
// device plan; any other block goes through DDA and a host-pointer plan.
template <typename LHS, typename WB, typename MB>
void fused_fastconv(LHS &lhs, WB const &w, MB const &in, float scale, integral_constant<bool, true>)
{
  length_type rows = lhs.size(2, 0), cols = lhs.size(2, 1);
  ctype const *x = in.device();
  ctype const *wp = w.device();
  ctype *y = lhs.device(OVX_ACCESS_WRITE); // in place (lhs is in) is allowed
  std::shared_ptr<ovxhip::Fastconv> fc = ovxhip::cached_fastconv(rows, cols, scale, 0);
  (*fc)(x, (long)cols, wp, WB::dim == 2 ? (long)cols : 0l, y, (long)cols, (long)rows);
}
template <typename LHS, typename WB, typename MB>
void fused_fastconv(LHS &lhs, WB const &w, MB const &in, float scale, integral_constant<bool, false>)
{
  length_type rows = lhs.size(2, 0), cols = lhs.size(2, 1);
  typedef typename conditional<WB::dim == 2, rows_layout, line_layout>::type w_layout;
  dda::Data<WB, dda::in, w_layout> w_data(w);
  dda::Data<MB, dda::in, rows_layout> m_data(in);
  dda::Data<LHS, dda::out, rows_layout> out_data(lhs);
  std::shared_ptr<ovxhip::Fastconv> fc = ovxhip::cached_fastconv(rows, cols, scale, OVX_FLAG_HOST_POINTERS);
  fastconv_call(*fc, m_data.ptr(), m_data.stride(0), w_data.ptr(), WB::dim == 2 ? w_data.stride(0) : 0,
                out_data.ptr(), out_data.stride(0), rows);
}
template <typename LHS, typename WB, typename MB>
void fused_fastconv(LHS &lhs, WB const &w, MB const &in, float scale)
{
  static bool const resident =
    is_resident_block<LHS>::value && is_resident_block<WB>::value && is_resident_block<MB>::value;
  fused_fastconv(lhs, w, in, scale, integral_constant<bool, resident>());
}

// The two row-Fftm backends of a fused expression, or null when either
// transform runs on another backend.
typedef fftm<vsip::row, fft_fwd> fwd_fftm_type;
typedef fftm<vsip::row, fft_inv> inv_fftm_type;
template <typename FwdFunctor, typename InvFunctor>
bool fused_backends(FwdFunctor const &f, InvFunctor const &i, fwd_fftm_type const *&fwd, inv_fftm_type const *&inv)
{
  fwd = dynamic_cast<fwd_fftm_type const *>(&f.backend());
  inv = dynamic_cast<inv_fftm_type const *>(&i.backend());
  return fwd && inv;
}

// ---- Correlation and Convolution (src/vsip/impl/signal/corr.hpp:31-38,
// conv.hpp:33-38).  These dispatch at compile time: the first backend whose
// Evaluator is ct_valid becomes the base class of vsip::Correlation /
// vsip::Convolution (src/ovxx/dispatcher/dispatcher.hpp:89-93), so there is
// no run-time fallthrough; a shape the library declines is run by the
// reference's generic backend, held alongside.  Data go through DDA as dense
// interleaved host arrays (converted by DDA from split storage).
inline int support_of(vsip::support_region_type r)
{
  return r == vsip::support_full ? OVX_SUPPORT_FULL : r == vsip::support_same ? OVX_SUPPORT_SAME : OVX_SUPPORT_MIN;
}
inline int symmetry_of(vsip::symmetry_type s)
{
  return s == vsip::nonsym ? OVX_NONSYM : s == vsip::sym_even_len_odd ? OVX_SYM_EVEN_LEN_ODD : OVX_SYM_EVEN_LEN_EVEN;
}
// interleaved complex<float> arrays (the `array` format of complex data)
typedef vsip::Layout<1, vsip::row1_type, vsip::dense, vsip::array> dense_interleaved;

// signal::Correlation's interface (src/ovxx/signal/corr.hpp:611-661).
template <vsip::support_region_type R, unsigned N, vsip::alg_hint_type H>
class corr
{
  typedef signal::Correlation<1, R, ctype, N, H> generic_type;

public:
  static vsip::support_region_type const supprt = R;
  corr(vsip::Domain<1> const &ref_size, vsip::Domain<1> const &input_size)
    : ref_size_(ref_size.size()), input_size_(input_size.size()),
      output_size_(ovx_corr_output_size(support_of(R), ref_size.size(), input_size.size()))
  {
    try {
      plan_.reset(new ovxhip::Correlation(support_of(R), ref_size.size(), input_size.size(),
                                          OVX_FLAG_HOST_POINTERS));
    } catch (ovxhip::unimplemented const &) {
      generic_.reset(new generic_type(ref_size, input_size));
    }
  }
  vsip::Domain<1> const &reference_size() const { return ref_size_; }
  vsip::Domain<1> const &input_size() const { return input_size_; }
  vsip::Domain<1> const &output_size() const { return output_size_; }

  template <typename B1, typename B2, typename B3>
  void correlate(vsip::bias_type bias, vsip::const_Vector<ctype, B1> ref, vsip::const_Vector<ctype, B2> in,
                 vsip::Vector<ctype, B3> out)
  {
    if (generic_) {
      generic_->correlate(bias, ref, in, out);
      return;
    }
    dda::Data<B1, dda::in, dense_interleaved> r(ref.block());
    dda::Data<B2, dda::in, dense_interleaved> x(in.block());
    dda::Data<B3, dda::out, dense_interleaved> y(out.block());
    (*plan_)(bias == vsip::unbiased ? OVX_UNBIASED : OVX_BIASED,
             reinterpret_cast<ovxhip::cscalar const *>(r.ptr()), reinterpret_cast<ovxhip::cscalar const *>(x.ptr()),
             reinterpret_cast<ovxhip::cscalar *>(y.ptr()));
  }

private:
  vsip::Domain<1> ref_size_, input_size_, output_size_;
  std::unique_ptr<ovxhip::Correlation> plan_;
  std::unique_ptr<generic_type> generic_;
};

// signal::Convolution's interface (src/ovxx/signal/conv.hpp:964-1010).
template <vsip::symmetry_type S, vsip::support_region_type R, unsigned N, vsip::alg_hint_type H>
class conv
{
  // signal::Convolution::convolve is protected (its public entry is
  // vsip::Convolution::operator(), src/ovxx/signal/conv.hpp:1004-1008):
  // the fallback reaches it through a derived type
  struct generic_type : signal::Convolution<vsip::const_Vector, S, R, ctype, N, H>
  {
    typedef signal::Convolution<vsip::const_Vector, S, R, ctype, N, H> base_type;
    template <typename B>
    generic_type(vsip::const_Vector<ctype, B> coeffs, vsip::Domain<1> const &input_size, length_type decimation)
      : base_type(coeffs, input_size, decimation)
    {}
    using base_type::convolve;
  };

public:
  static vsip::symmetry_type const symmtry = S;
  static vsip::support_region_type const supprt = R;
  template <typename B>
  conv(vsip::const_Vector<ctype, B> coeffs, vsip::Domain<1> const &input_size, length_type decimation)
    : input_size_(input_size.size()), decimation_(decimation)
  {
    const length_type m = S == vsip::nonsym ? coeffs.size()
                          : S == vsip::sym_even_len_odd ? 2 * coeffs.size() - 1 : 2 * coeffs.size();
    kernel_size_ = vsip::Domain<1>(m);
    output_size_ = vsip::Domain<1>(ovx_conv_output_size(support_of(R), m, input_size.size(), decimation));
    dda::Data<B, dda::in, dense_interleaved> c(coeffs.block());
    try {
      plan_.reset(new ovxhip::Convolution(symmetry_of(S), support_of(R),
                                          reinterpret_cast<ovxhip::cscalar const *>(c.ptr()), coeffs.size(),
                                          input_size.size(), decimation, OVX_FLAG_HOST_POINTERS));
    } catch (ovxhip::unimplemented const &) {
      generic_.reset(new generic_type(coeffs, input_size, decimation));
    }
  }
  vsip::Domain<1> const &kernel_size() const { return kernel_size_; }
  vsip::Domain<1> const &filter_order() const { return kernel_size_; }
  vsip::Domain<1> const &input_size() const { return input_size_; }
  vsip::Domain<1> const &output_size() const { return output_size_; }
  vsip::symmetry_type symmetry() const { return S; }
  vsip::support_region_type support() const { return R; }
  length_type decimation() const { return decimation_; }

  template <typename B1, typename B2>
  void convolve(vsip::const_Vector<ctype, B1> in, vsip::Vector<ctype, B2> out)
  {
    if (generic_) {
      generic_->convolve(in, out);
      return;
    }
    dda::Data<B1, dda::in, dense_interleaved> x(in.block());
    dda::Data<B2, dda::out, dense_interleaved> y(out.block());
    (*plan_)(reinterpret_cast<ovxhip::cscalar const *>(x.ptr()), reinterpret_cast<ovxhip::cscalar *>(y.ptr()));
  }

private:
  vsip::Domain<1> kernel_size_, input_size_, output_size_;
  length_type decimation_;
  std::unique_ptr<ovxhip::Convolution> plan_;
  std::unique_ptr<generic_type> generic_;
};

} // namespace ovxx::ovxhip_backend

namespace dispatcher
{
// Fftm factory (pattern: share/examples/dispatch/fft.cpp:165-200).
template <int A, int D, vsip::return_mechanism_type R, unsigned N>
struct Evaluator<op::fftm<complex<float>, complex<float>, A, D, R, N>, be::user,
                 std::unique_ptr<signal::fft::fftm_backend<complex<float>, complex<float>, A, D> >(
                   vsip::Domain<2> const &, float)>
{
  static bool const ct_valid = true;
  static bool rt_valid(vsip::Domain<2> const &dom, float)
  {
    return ovxhip_backend::length_ok(A == vsip::col ? dom[0].size() : dom[1].size());
  }
  static std::unique_ptr<signal::fft::fftm_backend<complex<float>, complex<float>, A, D> >
  exec(vsip::Domain<2> const &dom, float scale)
  {
    return std::unique_ptr<signal::fft::fftm_backend<complex<float>, complex<float>, A, D> >(
      new ovxhip_backend::fftm<A, D>(dom, scale));
  }
};

// Real Fftm factories (the real-domain size is dom; the transform axis is A).
template <int A, vsip::return_mechanism_type R, unsigned N>
struct Evaluator<op::fftm<float, complex<float>, A, fft_fwd, R, N>, be::user,
                 std::unique_ptr<signal::fft::fftm_backend<float, complex<float>, A, fft_fwd> >(
                   vsip::Domain<2> const &, float)>
{
  static bool const ct_valid = true;
  static bool rt_valid(vsip::Domain<2> const &dom, float)
  {
    return ovxhip_backend::real_length_ok(A == vsip::col ? dom[0].size() : dom[1].size());
  }
  static std::unique_ptr<signal::fft::fftm_backend<float, complex<float>, A, fft_fwd> >
  exec(vsip::Domain<2> const &dom, float scale)
  {
    return std::unique_ptr<signal::fft::fftm_backend<float, complex<float>, A, fft_fwd> >(
      new ovxhip_backend::fftm_r2c<A>(dom, scale));
  }
};

template <int A, vsip::return_mechanism_type R, unsigned N>
struct Evaluator<op::fftm<complex<float>, float, A, fft_inv, R, N>, be::user,
                 std::unique_ptr<signal::fft::fftm_backend<complex<float>, float, A, fft_inv> >(
                   vsip::Domain<2> const &, float)>
{
  static bool const ct_valid = true;
  static bool rt_valid(vsip::Domain<2> const &dom, float)
  {
    return ovxhip_backend::real_length_ok(A == vsip::col ? dom[0].size() : dom[1].size());
  }
  static std::unique_ptr<signal::fft::fftm_backend<complex<float>, float, A, fft_inv> >
  exec(vsip::Domain<2> const &dom, float scale)
  {
    return std::unique_ptr<signal::fft::fftm_backend<complex<float>, float, A, fft_inv> >(
      new ovxhip_backend::fftm_c2r<A>(dom, scale));
  }
};

// Fft factory: one complex 1-D transform (e.g. the replica spectrum,
// tests/fastconv.cpp:20-26).
template <int S, vsip::return_mechanism_type R, unsigned N>
struct Evaluator<op::fft<1, complex<float>, complex<float>, S, R, N>, be::user,
                 std::unique_ptr<signal::fft::fft_backend<1, complex<float>, complex<float>, S> >(
                   vsip::Domain<1> const &, float)>
{
  static bool const ct_valid = true;
  static bool rt_valid(vsip::Domain<1> const &dom, float) { return ovxhip_backend::length_ok(dom.size()); }
  static std::unique_ptr<signal::fft::fft_backend<1, complex<float>, complex<float>, S> >
  exec(vsip::Domain<1> const &dom, float scale)
  {
    return std::unique_ptr<signal::fft::fft_backend<1, complex<float>, complex<float>, S> >(
      new ovxhip_backend::fft<S>(dom, scale));
  }
};

// The fused fast convolution, LHS = inv_fftm(vmmul<0>(w, fwd_fftm(M))): the
// by-value expression of tests/fastconv.cpp:200-205 and
// benchmarks/fastconv.cpp:797-801, which the reference evaluates as four
// passes through be::rbo_expr.  This evaluator makes one fused call.
// Pattern: share/examples/eval/fft_expression.cpp:48-74; guard:
// is_fftm_functor (src/ovxx/signal/fft/functor.hpp:130-139).  Both
// transforms must be this backend's row Fftms; their scales multiply (the
// expression is linear), so a forward scale other than 1 is kept.
template <typename LHS, template <typename> class FI, typename VB, template <typename> class FF, typename MB>
struct Evaluator<op::assign<2>, be::user,
                 void(LHS &, expr::Unary<FI, expr::Vmmul<0, VB, expr::Unary<FF, MB> const> const> const &),
                 typename enable_if<expr::is_fftm_functor<FI<expr::Vmmul<0, VB, expr::Unary<FF, MB> const> const> >::value &&
                                    expr::is_fftm_functor<FF<MB> >::value>::type>
{
  static char const *name() { return "be::user ovxhip fused fastconv"; }
  typedef expr::Unary<FF, MB> fwd_block_type;
  typedef expr::Vmmul<0, VB, fwd_block_type const> vmmul_block_type;
  typedef expr::Unary<FI, vmmul_block_type const> RHS;

  static bool const ct_valid = is_same<typename LHS::value_type, complex<float> >::value &&
                               is_same<typename VB::value_type, complex<float> >::value &&
                               is_same<typename MB::value_type, complex<float> >::value;

  static bool rt_valid(LHS &lhs, RHS const &rhs)
  {
    ovxhip_backend::fwd_fftm_type const *fwd;
    ovxhip_backend::inv_fftm_type const *inv;
    return ovxhip_backend::fused_backends(rhs.arg().get_mblk().operation(), rhs.operation(), fwd, inv) &&
           ovxhip_backend::length_ok(lhs.size(2, 1));
  }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    ovxhip_backend::fwd_fftm_type const *fwd;
    ovxhip_backend::inv_fftm_type const *inv;
    ovxhip_backend::fused_backends(rhs.arg().get_mblk().operation(), rhs.operation(), fwd, inv);
    ovxhip_backend::fused_fastconv(lhs, rhs.arg().get_vblk(), rhs.arg().get_mblk().arg(),
                                   fwd->scale() * inv->scale());
  }
};

// The coefficient-matrix forms, LHS = inv_fftm(C * fwd_fftm(M)) and
// LHS = inv_fftm(fwd_fftm(M) * C) (tests/fastconv.cpp:215-245, order 0 and 1;
// benchmarks/fastconv.cpp:857-969): one fused call reading C row by row.
template <typename LHS, template <typename> class FI, typename CB, template <typename> class FF, typename MB>
struct Evaluator<
  op::assign<2>, be::user,
  void(LHS &, expr::Unary<FI, expr::Binary<expr::op::Mult, CB, expr::Unary<FF, MB> const, true> const> const &),
  typename enable_if<
    expr::is_fftm_functor<FI<expr::Binary<expr::op::Mult, CB, expr::Unary<FF, MB> const, true> const> >::value &&
    expr::is_fftm_functor<FF<MB> >::value>::type>
{
  static char const *name() { return "be::user ovxhip fused fastconv C * fftm(M)"; }
  typedef expr::Binary<expr::op::Mult, CB, expr::Unary<FF, MB> const, true> mult_block_type;
  typedef expr::Unary<FI, mult_block_type const> RHS;

  static bool const ct_valid = CB::dim == 2 && is_same<typename LHS::value_type, complex<float> >::value &&
                               is_same<typename CB::value_type, complex<float> >::value &&
                               is_same<typename MB::value_type, complex<float> >::value;

  static bool rt_valid(LHS &lhs, RHS const &rhs)
  {
    ovxhip_backend::fwd_fftm_type const *fwd;
    ovxhip_backend::inv_fftm_type const *inv;
    return ovxhip_backend::fused_backends(rhs.arg().arg2().operation(), rhs.operation(), fwd, inv) &&
           ovxhip_backend::length_ok(lhs.size(2, 1));
  }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    ovxhip_backend::fwd_fftm_type const *fwd;
    ovxhip_backend::inv_fftm_type const *inv;
    ovxhip_backend::fused_backends(rhs.arg().arg2().operation(), rhs.operation(), fwd, inv);
    ovxhip_backend::fused_fastconv(lhs, rhs.arg().arg1(), rhs.arg().arg2().arg(), fwd->scale() * inv->scale());
  }
};

template <typename LHS, template <typename> class FI, template <typename> class FF, typename MB, typename CB>
struct Evaluator<
  op::assign<2>, be::user,
  void(LHS &, expr::Unary<FI, expr::Binary<expr::op::Mult, expr::Unary<FF, MB> const, CB, true> const> const &),
  typename enable_if<
    expr::is_fftm_functor<FI<expr::Binary<expr::op::Mult, expr::Unary<FF, MB> const, CB, true> const> >::value &&
    expr::is_fftm_functor<FF<MB> >::value>::type>
{
  static char const *name() { return "be::user ovxhip fused fastconv fftm(M) * C"; }
  typedef expr::Binary<expr::op::Mult, expr::Unary<FF, MB> const, CB, true> mult_block_type;
  typedef expr::Unary<FI, mult_block_type const> RHS;

  static bool const ct_valid = CB::dim == 2 && is_same<typename LHS::value_type, complex<float> >::value &&
                               is_same<typename CB::value_type, complex<float> >::value &&
                               is_same<typename MB::value_type, complex<float> >::value;

  static bool rt_valid(LHS &lhs, RHS const &rhs)
  {
    ovxhip_backend::fwd_fftm_type const *fwd;
    ovxhip_backend::inv_fftm_type const *inv;
    return ovxhip_backend::fused_backends(rhs.arg().arg1().operation(), rhs.operation(), fwd, inv) &&
           ovxhip_backend::length_ok(lhs.size(2, 1));
  }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    ovxhip_backend::fwd_fftm_type const *fwd;
    ovxhip_backend::inv_fftm_type const *inv;
    ovxhip_backend::fused_backends(rhs.arg().arg1().operation(), rhs.operation(), fwd, inv);
    ovxhip_backend::fused_fastconv(lhs, rhs.arg().arg2(), rhs.arg().arg1().arg(), fwd->scale() * inv->scale());
  }
};

// The per-vector form, out = inv_fft(w * fwd_fft(in)) on 1-D views
// (tests/fastconv.cpp:160-179: one row at a time): one fused call of one
// row.  Both transforms must be this backend's 1-D Ffts.
template <typename LHS, template <typename> class FI, typename VB, template <typename> class FF, typename B>
struct Evaluator<
  op::assign<1>, be::user,
  void(LHS &, expr::Unary<FI, expr::Binary<expr::op::Mult, VB, expr::Unary<FF, B> const, true> const> const &),
  typename enable_if<
    expr::is_fft_functor<FI<expr::Binary<expr::op::Mult, VB, expr::Unary<FF, B> const, true> const> >::value &&
    expr::is_fft_functor<FF<B> >::value>::type>
{
  static char const *name() { return "be::user ovxhip fused fastconv (vector)"; }
  typedef expr::Binary<expr::op::Mult, VB, expr::Unary<FF, B> const, true> mult_block_type;
  typedef expr::Unary<FI, mult_block_type const> RHS;
  typedef ovxhip_backend::fft<fft_fwd> fwd_type;
  typedef ovxhip_backend::fft<fft_inv> inv_type;

  static bool const ct_valid = VB::dim == 1 && is_same<typename LHS::value_type, complex<float> >::value &&
                               is_same<typename VB::value_type, complex<float> >::value &&
                               is_same<typename B::value_type, complex<float> >::value;

  static bool rt_valid(LHS &lhs, RHS const &rhs)
  {
    return dynamic_cast<fwd_type const *>(&rhs.arg().arg2().operation().backend()) &&
           dynamic_cast<inv_type const *>(&rhs.operation().backend()) && ovxhip_backend::length_ok(lhs.size());
  }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    fwd_type const *fwd = dynamic_cast<fwd_type const *>(&rhs.arg().arg2().operation().backend());
    inv_type const *inv = dynamic_cast<inv_type const *>(&rhs.operation().backend());
    length_type n = lhs.size();
    typedef ovxhip_backend::line_layout L;
    dda::Data<VB, dda::in, L> w_data(rhs.arg().arg1());
    dda::Data<B, dda::in, L> x_data(rhs.arg().arg2().arg());
    dda::Data<LHS, dda::out, L> y_data(lhs);
    std::shared_ptr<ovxhip::Fastconv> fc =
      ovxhip::cached_fastconv(1, n, fwd->scale() * inv->scale(), OVX_FLAG_HOST_POINTERS);
    ovxhip_backend::fastconv_call(*fc, x_data.ptr(), (stride_type)n, w_data.ptr(), 0, y_data.ptr(), (stride_type)n,
                                  1);
  }
};

// vmmul<D>(v, m) on Resident blocks (src/vsip/impl/vmmul.hpp:50-106): one
// pass over HBM-resident data, so that the phased form of fastconv
// (tests/fastconv.cpp:133-150: Fftm, vmmul, Fftm) keeps its intermediate on
// the device.  Other blocks keep the reference's evaluators (a host-memory
// loop is cheaper than two trips over the host link).
template <typename LHS, typename VB, typename MB, dimension_type D>
struct Evaluator<op::assign<2>, be::user, void(LHS &, expr::Vmmul<D, VB, MB> const &)>
{
  static char const *name() { return "be::user ovxhip vmmul"; }
  typedef expr::Vmmul<D, VB, MB> RHS;
  typedef typename LHS::value_type lt;
  typedef typename VB::value_type vt;
  typedef typename MB::value_type mt;
  static bool const v_ok = is_same<vt, float>::value || is_same<vt, complex<float> >::value;
  static bool const m_ok = is_same<mt, float>::value || is_same<mt, complex<float> >::value;
  static bool const l_ok = is_same<lt, typename vsip::Promotion<vt, mt>::type>::value;

  static bool const ct_valid = ovxhip_backend::is_resident_block<LHS>::value &&
                               ovxhip_backend::is_resident_block<VB>::value &&
                               ovxhip_backend::is_resident_block<MB>::value && v_ok && m_ok && l_ok;

  static bool rt_valid(LHS &, RHS const &) { return true; }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    length_type rows = lhs.size(2, 0), cols = lhs.size(2, 1);
    void const *v = rhs.get_vblk().device();
    void const *m = rhs.get_mblk().device();
    void *out = lhs.device(OVX_ACCESS_WRITE); // may be m: in place is allowed
    ovxhip::check(ovx_vmmul(D, is_complex<vt>::value, v, 1, is_complex<mt>::value, m, (long)cols, 1, out,
                            (long)cols, 1, (long)rows, (long)cols, 1.0f, nullptr));
  }
};

// The by-value Fftm, out = fftm(in), on Resident blocks: the backend's
// device plan on the HBM copies (other blocks take the reference's
// return-block path, src/ovxx/signal/fft.hpp:324-345, into the backend).
template <typename LHS, template <typename> class F>
struct Evaluator<op::assign<2>, be::user,
                 void(LHS &, expr::Unary<F, ovxhip_backend::Resident<2, complex<float> > > const &),
                 typename enable_if<expr::is_fftm_functor<F<ovxhip_backend::Resident<2, complex<float> > > >::value>::type>
{
  static char const *name() { return "be::user ovxhip resident fftm"; }
  typedef ovxhip_backend::Resident<2, complex<float> > MB;
  typedef expr::Unary<F, MB> RHS;
  typedef typename F<MB>::backend_type backend_type;
  // fftm_backend<I, O, A, D> derives from backend<2, T, 1 - A, exponent>
  // (src/ovxx/signal/fft/backend.hpp:37-44, 697-699)
  typedef ovxhip_backend::fftm<1 - backend_type::axis, backend_type::exponent == -1 ? fft_fwd : fft_inv> our_type;

  static bool const ct_valid = ovxhip_backend::is_resident_block<LHS>::value &&
                               ovxhip_backend::is_resident_block<MB>::value &&
                               is_same<typename LHS::value_type, complex<float> >::value &&
                               is_same<typename MB::value_type, complex<float> >::value;

  static bool rt_valid(LHS &, RHS const &rhs)
  {
    return dynamic_cast<our_type const *>(&rhs.operation().backend()) != 0;
  }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    our_type const *be = dynamic_cast<our_type const *>(&rhs.operation().backend());
    length_type rows = lhs.size(2, 0), cols = lhs.size(2, 1);
    ovxhip::cscalar const *x = rhs.arg().device();
    ovxhip::cscalar *y = lhs.device(OVX_ACCESS_WRITE);
    be->device_plan().out_of_place(x, (long)cols, 1, y, (long)cols, 1, (long)rows, (long)cols);
  }
};

// complex<float> 1-D correlation and convolution on the gfx950 kernels
template <vsip::support_region_type R, unsigned N, vsip::alg_hint_type H>
struct Evaluator<op::corr<1, R, complex<float>, N, H>, be::user>
{
  static bool const ct_valid = true;
  typedef ovxhip_backend::corr<R, N, H> backend_type;
};

template <vsip::symmetry_type S, vsip::support_region_type R, unsigned N, vsip::alg_hint_type H>
struct Evaluator<op::conv<1, S, R, complex<float>, N, H>, be::user>
{
  static bool const ct_valid = true;
  typedef ovxhip_backend::conv<S, R, N, H> backend_type;
};

} // namespace ovxx::dispatcher
} // namespace ovxx

#endif
