// Instantiates every evaluator of include/ovxhip/vsip_plugin.hpp against the
// OpenVSIP headers (compiled with -fsyntax-only by
// tests/test_plugin_compile.py; nothing here runs).  The static_asserts check
// that be::user -- first in the assign list, src/ovxx/assign_fwd.hpp:49-62 --
// takes the reference's own expression types of tests/fastconv.cpp:160-245
// and tests/vmmul.cpp at compile time (rt_valid then decides at run time).
#include <vsip/initfin.hpp>
#include <vsip/matrix.hpp>
#include <vsip/signal.hpp>
#include <vsip/vector.hpp>

#include <ovxhip/vsip_plugin.hpp>

#include <cstring>

using namespace vsip;
typedef complex<float> C;
typedef ovxx::ovxhip_backend::Resident<2> R2;
typedef ovxx::ovxhip_backend::Resident<1> R1;
typedef ovxx::ovxhip_backend::Resident<2, float> R2f;
typedef ovxx::ovxhip_backend::Resident<1, float> R1f;
namespace d = ovxx::dispatcher;

// be::user's evaluator for `lhs = rhs` is one of this plugin's
template <dimension_type D, typename LV, typename RV>
void expect_user(LV const &, RV const &, char const *name)
{
  typedef typename LV::block_type L;
  typedef typename RV::block_type R;
  typedef d::Evaluator<d::op::assign<D>, d::be::user, void(L &, R const &)> E;
  static_assert(E::ct_valid, "be::user does not take this expression");
  if (std::strcmp(E::name(), name) != 0) throw std::logic_error(E::name());
}

// ...and is not consulted for one it must leave to the reference
template <dimension_type D, typename LV, typename RV>
void expect_not_user(LV const &, RV const &)
{
  typedef typename LV::block_type L;
  typedef typename RV::block_type R;
  static_assert(!d::Evaluator<d::op::assign<D>, d::be::user, void(L &, R const &)>::ct_valid,
                "be::user claims an expression it does not handle");
}

int main(int argc, char **argv)
{
  vsipl init(argc, argv);
  length_type const rows = 64, n = 2048;
  typedef Fftm<C, C, row, fft_fwd, by_value> fwd_t;
  typedef Fftm<C, C, row, fft_inv, by_value> inv_t;
  typedef Fftm<C, C, row, fft_fwd, by_reference> fwd_ref_t;
  typedef Fftm<C, C, row, fft_inv, by_reference> inv_ref_t;
  typedef Fftm<C, C, col, fft_fwd, by_value> colfwd_t;
  fwd_t f(Domain<2>(rows, n), 1.f);
  inv_t i(Domain<2>(rows, n), 1.f / n);
  fwd_ref_t fr(Domain<2>(rows, n), 1.f);
  inv_ref_t ir(Domain<2>(rows, n), 1.f / n);
  colfwd_t fc(Domain<2>(rows, n), 1.f);

  // fused fastconv on Dense blocks (host pointers through DDA)
  Matrix<C> in(rows, n), out(rows, n), cm(rows, n), tmp(rows, n);
  Vector<C> w(n);
  expect_user<2>(out, i(vmmul<0>(w, f(in))), "be::user ovxhip fused fastconv");
  expect_user<2>(out, i(cm * f(in)), "be::user ovxhip fused fastconv C * fftm(M)");
  expect_user<2>(out, i(f(in) * cm), "be::user ovxhip fused fastconv fftm(M) * C");
  out = i(vmmul<0>(w, f(in)));
  out = i(cm * f(in));
  out = i(f(in) * cm);
  in = i(vmmul<0>(w, f(in))); // in place

  // the per-row vector form (Fft by value on rows)
  Fft<const_Vector, C, C, fft_fwd, by_value> vf(Domain<1>(n), 1.f);
  Fft<const_Vector, C, C, fft_inv, by_value> vi(Domain<1>(n), 1.f / n);
  expect_user<1>(out.row(0), vi(w * vf(in.row(0))), "be::user ovxhip fused fastconv (vector)");
  for (index_type r = 0; r < rows; ++r) out.row(r) = vi(w * vf(in.row(r)));

  // the phased form: Fftm by reference, vmmul, Fftm by reference; vmmul of
  // Dense blocks stays with the reference
  fr(in, tmp);
  expect_not_user<2>(tmp, vmmul<0>(w, tmp));
  tmp = vmmul<0>(w, tmp);
  ir(tmp, out);

  // Resident blocks: fused forms, vmmul and the by-value Fftm on HBM data
  Matrix<C, R2> rin(rows, n), rout(rows, n), rcm(rows, n), rtmp(rows, n);
  Vector<C, R1> rw(n), rv(rows);
  Matrix<float, R2f> rfm(rows, n);
  Vector<float, R1f> rfv(n);
  expect_user<2>(rout, i(vmmul<0>(rw, f(rin))), "be::user ovxhip fused fastconv");
  expect_user<2>(rout, i(rcm * f(rin)), "be::user ovxhip fused fastconv C * fftm(M)");
  expect_user<2>(rout, vmmul<0>(rw, rin), "be::user ovxhip vmmul");
  expect_user<2>(rout, vmmul<1>(rv, rin), "be::user ovxhip vmmul");
  expect_user<2>(rout, vmmul<0>(rfv, rin), "be::user ovxhip vmmul");
  expect_user<2>(rout, vmmul<0>(rw, rfm), "be::user ovxhip vmmul");
  expect_user<2>(rfm, vmmul<0>(rfv, rfm), "be::user ovxhip vmmul");
  expect_user<2>(rtmp, f(rin), "be::user ovxhip resident fftm");
  expect_user<2>(rtmp, fc(rin), "be::user ovxhip resident fftm");
  rout = i(vmmul<0>(rw, f(rin)));
  rout = i(rcm * f(rin));
  rtmp = f(rin);
  rtmp = vmmul<0>(rw, rtmp);
  rout = i(rtmp);
  rfm = vmmul<1>(Vector<float, R1f>(rows), rfm);
  rin(3, 4) = C(1, 2);
  C v = rin.get(3, 4);
  (void)v;

  // Correlation and Convolution (the generic fallback included).  With
  // split complex storage the reference's own generic Convolution does not
  // compile (src/ovxx/signal/conv.hpp:1058-1068 pass split pointers to the
  // interleaved conv_full/conv_same/conv_min), so that configuration leaves
  // Convolution out.
  Vector<C> x(4096), y(4096 + n - 1), ref(16), c(64 + 16 - 1);
  Correlation<const_Vector, support_full, C> corr(Domain<1>(16), Domain<1>(64));
  corr(biased, ref, x(Domain<1>(64)), c);
  Correlation<const_Vector, support_same, C> corr_same(Domain<1>(16), Domain<1>(64));
  Vector<C> cs(64);
  corr_same(unbiased, ref, x(Domain<1>(64)), cs);
#if !OVXX_DEFAULT_COMPLEX_STORAGE_SPLIT
  Convolution<const_Vector, nonsym, support_full, C> conv(w, Domain<1>(4096));
  conv(x, y);
  Convolution<const_Vector, sym_even_len_odd, support_min, C> conv_sym(w(Domain<1>(8)), Domain<1>(4096), 2);
  Vector<C> ym(conv_sym.output_size().size());
  conv_sym(x, ym);
#endif
  return 0;
}
