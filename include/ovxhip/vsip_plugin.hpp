// vsip_plugin.hpp -- registers the gfx950 backend with OpenVSIP's dispatcher.
//
// Include this header in a translation unit that includes <vsip/signal.hpp>
// (and therefore a configured OpenVSIP tree) and link libovxhip.so.  It adds
// Evaluator specialisations on be::user, which is first in every relevant
// dispatch list:
//   op::fft   list  src/ovxx/signal/fft.hpp:35-49
//   op::fftm  list  src/ovxx/signal/fft.hpp:51-64
//   assign<2> list  src/ovxx/assign_fwd.hpp:49-62
//   op::corr  list  src/vsip/impl/signal/corr.hpp:26-38
//   op::conv  list  src/vsip/impl/signal/conv.hpp:27-38
// so existing vsip::Fft / vsip::Fftm / vmmul code selects the HIP kernels
// without source changes; rt_valid() returns false for a shape the backend
// does not handle, and the dispatcher falls through to the next backend
// (src/ovxx/dispatcher/dispatcher.hpp:96-130).
//
// Data reach the backend through OpenVSIP's DDA as host pointers, so the
// plans are created with OVX_FLAG_HOST_POINTERS (the library stages the data
// through device buffers); an application that keeps its data on the device
// calls include/ovxhip.h directly.
//
// The backends ask DDA for the storage format of OpenVSIP's dense blocks on
// their complex sides (interleaved, or split when OpenVSIP is configured
// with split complex storage), so neither format is converted on the host;
// split data take the library's split entry points.
//
// This header compiles only against an OpenVSIP tree; it is not built by this
// repository's tests.
#ifndef OVXHIP_VSIP_PLUGIN_HPP
#define OVXHIP_VSIP_PLUGIN_HPP

#include <memory>

#include <ovxx/dispatch.hpp>
#include <ovxx/dda.hpp>
#include <ovxx/expr/unary.hpp>
#include <ovxx/expr/vmmul.hpp>
#include <ovxx/signal/fft.hpp>
#include <ovxx/signal/fft/backend.hpp>
#include <ovxx/signal/fft/functor.hpp>
#include <ovxx/signal/conv.hpp>
#include <ovxx/signal/corr.hpp>

#include "ovxhip.hpp"

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
    : plan_(dom[0].size(), dom[1].size(), scale, OVX_FLAG_HOST_POINTERS), scale_(scale)
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

private:
  ovxhip::Fftm<axis, dir> plan_;
  float scale_;
};

// fft_backend<1,c,c,S> (src/ovxx/signal/fft/backend.hpp:131-204).
template <int S>
class fft : public signal::fft::fft_backend<1, ctype, ctype, S>
{
  typedef signal::fft::fft_backend<1, ctype, ctype, S> base_type;
  static int const dir = S == fft_fwd ? OVX_FFT_FWD : OVX_FFT_INV;

public:
  fft(vsip::Domain<1> const &dom, float scale) : plan_(dom.size(), scale, OVX_FLAG_HOST_POINTERS) {}
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

// The fused fastconv on what DDA hands over: interleaved complex pointers,
// or (real, imaginary) pairs when the blocks use split storage.
inline void fastconv_call(ovxhip::Fastconv &fc, ctype const *x, stride_type ldx, ctype const *w, ctype *y,
                          stride_type ldy, length_type rows)
{
  fc(reinterpret_cast<ovxhip::cscalar const *>(x), ldx, reinterpret_cast<ovxhip::cscalar const *>(w), 0,
     reinterpret_cast<ovxhip::cscalar *>(y), ldy, rows);
}
template <typename PX, typename PW>
void fastconv_call(ovxhip::Fastconv &fc, std::pair<PX *, PX *> x, stride_type ldx, std::pair<PW *, PW *> w,
                   std::pair<float *, float *> y, stride_type ldy, length_type rows)
{
  fc(ovxhip::const_split_ptr(x.first, x.second), ldx, ovxhip::const_split_ptr(w.first, w.second), 0, y, ldy, rows);
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
typedef vsip::Layout<1, vsip::row1_type, vsip::dense, vsip::interleaved_complex> dense_interleaved;

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
  typedef signal::Convolution<vsip::const_Vector, S, R, ctype, N, H> generic_type;

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
// passes through be::rbo_expr.  This evaluator reads w and M through DDA and
// makes one fused call.  Pattern: share/examples/eval/fft_expression.cpp:48-74;
// guard: is_fftm_functor (src/ovxx/signal/fft/functor.hpp:130-139).
template <typename LHS, template <typename> class FI, typename VB, template <typename> class FF, typename MB>
struct Evaluator<op::assign<2>, be::user,
                 void(LHS &, expr::Unary<FI, expr::Vmmul<0, VB, expr::Unary<FF, MB> const> const> const &),
                 typename enable_if<expr::is_fftm_functor<FI<expr::Vmmul<0, VB, expr::Unary<FF, MB> const> > >::value &&
                                    expr::is_fftm_functor<FF<MB> >::value>::type>
{
  static char const *name() { return "be::user ovxhip fused fastconv"; }
  typedef expr::Unary<FF, MB> fwd_block_type;
  typedef expr::Vmmul<0, VB, fwd_block_type const> vmmul_block_type;
  typedef expr::Unary<FI, vmmul_block_type const> RHS;
  typedef ovxhip_backend::fftm<vsip::row, fft_inv> inv_backend_type;
  typedef ovxhip_backend::fftm<vsip::row, fft_fwd> fwd_backend_type;

  static bool const ct_valid = is_same<typename LHS::value_type, complex<float> >::value &&
                               is_same<typename VB::value_type, complex<float> >::value &&
                               is_same<typename MB::value_type, complex<float> >::value;

  // both transforms must be this backend's row Fftms (their plans hold the
  // inverse scale) and the rows must fit a supported length
  static bool rt_valid(LHS &lhs, RHS const &rhs)
  {
    inv_backend_type const *inv = dynamic_cast<inv_backend_type const *>(&rhs.operation().backend());
    fwd_backend_type const *fwd =
      dynamic_cast<fwd_backend_type const *>(&rhs.arg().get_mblk().operation().backend());
    return inv && fwd && ovxhip_backend::length_ok(lhs.size(2, 1));
  }

  static void exec(LHS &lhs, RHS const &rhs)
  {
    inv_backend_type const *inv = dynamic_cast<inv_backend_type const *>(&rhs.operation().backend());
    VB const &w = rhs.arg().get_vblk();
    MB const &m = rhs.arg().get_mblk().arg();
    length_type rows = lhs.size(2, 0), cols = lhs.size(2, 1);
    dda::Data<VB, dda::in> w_data(w);
    dda::Data<MB, dda::in> m_data(m);
    dda::Data<LHS, dda::out> out_data(lhs);
    ovxhip::Fastconv &fc = ovxhip::cached_fastconv(rows, cols, inv->scale(), OVX_FLAG_HOST_POINTERS);
    ovxhip_backend::fastconv_call(fc, m_data.ptr(), m_data.stride(0), w_data.ptr(), out_data.ptr(),
                                  out_data.stride(0), rows);
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
