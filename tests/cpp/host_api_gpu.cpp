// Runs include/ovxhip/ovxhip.hpp on the GPU the way the VSIPL++ plugin does:
// host pointers (OVX_FLAG_HOST_POINTERS), interleaved and split storage, the
// per-thread fastconv plan cache, and a Storage-resident call on device
// pointers.  Results are checked against an FP64 DFT written here, and the
// split and resident calls against the interleaved host-pointer call, bit
// for bit.  Prints OK on success.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "ovxhip/ovxhip.hpp"

typedef std::complex<double> zd;
typedef ovxhip::cscalar zf;

static int fails = 0;
#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      std::printf("FAILED line %d: %s\n", __LINE__, #c);      \
      ++fails;                                                \
    }                                                         \
  } while (0)

// X[k] = sum_j x[j] exp(sign 2 pi i j k / n), in FP64
static std::vector<zd> dft(std::vector<zd> const &x, int sign)
{
  const size_t n = x.size();
  std::vector<zd> y(n);
  const double w = sign * 2.0 * M_PI / (double)n;
  for (size_t k = 0; k < n; ++k) {
    zd acc = 0;
    for (size_t j = 0; j < n; ++j) acc += x[j] * std::polar(1.0, w * (double)((j * k) % n));
    y[k] = acc;
  }
  return y;
}

static double rel_l2(zf const *got, std::vector<zd> const &ref)
{
  double num = 0, den = 0;
  for (size_t i = 0; i < ref.size(); ++i) {
    zd d = zd(got[i].real(), got[i].imag()) - ref[i];
    num += std::norm(d);
    den += std::norm(ref[i]);
  }
  return std::sqrt(num / den);
}

int main()
{
  const long rows = 4, n = 1024;
  std::mt19937 gen(7);
  std::normal_distribution<float> g(0.f, 0.7f);
  std::vector<zf> x(rows * n), w(n), y(rows * n);
  for (auto &v : x) v = zf(g(gen), g(gen));
  for (auto &v : w) v = zf(g(gen), g(gen));

  // Fftm forward on host pointers, against the FP64 DFT
  ovxhip::Fftm<OVX_ROW, OVX_FFT_FWD> fftm(rows, n, 1.0f, OVX_FLAG_HOST_POINTERS);
  fftm(x.data(), y.data());
  for (long r = 0; r < rows; ++r) {
    std::vector<zd> xr(x.begin() + r * n, x.begin() + (r + 1) * n);
    CHECK(rel_l2(&y[r * n], dft(xr, -1)) < 1e-5);
  }

  // the same transform on split storage, in place: the same floats
  std::vector<float> re(rows * n), im(rows * n);
  for (long i = 0; i < rows * n; ++i) re[i] = x[i].real(), im[i] = x[i].imag();
  fftm.in_place(ovxhip::split_ptr(re.data(), im.data()), n, 1, rows, n);
  bool same = true;
  for (long i = 0; i < rows * n; ++i) same = same && re[i] == y[i].real() && im[i] == y[i].imag();
  CHECK(same);

  // fused fastconv through the plan cache: y = IDFT(w . DFT(x)) / n
  std::shared_ptr<ovxhip::Fastconv> fc = ovxhip::cached_fastconv(rows, n, 1.0f / n, OVX_FLAG_HOST_POINTERS);
  CHECK(fc == ovxhip::cached_fastconv(rows, n, 1.0f / n, OVX_FLAG_HOST_POINTERS));
  CHECK(fc != ovxhip::cached_fastconv(rows, n, 0.5f / n, OVX_FLAG_HOST_POINTERS));
  // another thread gets its own plan (a plan is not safe for concurrent calls)
  std::shared_ptr<ovxhip::Fastconv> other;
  std::thread([&] { other = ovxhip::cached_fastconv(rows, n, 1.0f / n, OVX_FLAG_HOST_POINTERS); }).join();
  CHECK(other && other != fc);
  // the cache is bounded; a plan dropped from it stays valid for its holder
  for (int i = 0; i < (int)ovxhip::fastconv_cache::kMaxCachedPlans + 4; ++i)
    ovxhip::cached_fastconv(1, 64, 1.0f + i, OVX_FLAG_HOST_POINTERS);
  CHECK(ovxhip::fastconv_cache::size() == ovxhip::fastconv_cache::kMaxCachedPlans);
  CHECK(fc != ovxhip::cached_fastconv(rows, n, 1.0f / n, OVX_FLAG_HOST_POINTERS));
  std::vector<zf> yc(rows * n);
  (*fc)(x.data(), w.data(), yc.data());
  for (long r = 0; r < rows; ++r) {
    std::vector<zd> xr(x.begin() + r * n, x.begin() + (r + 1) * n);
    std::vector<zd> f = dft(xr, -1);
    for (long k = 0; k < n; ++k) f[k] *= zd(w[k].real(), w[k].imag()) / (double)n;
    CHECK(rel_l2(&yc[r * n], dft(f, +1)) < 1e-5);
  }

  // the same fastconv with X and W kept in Storage, on device pointers
  ovxhip::Fastconv dev_fc(rows, n, 1.0f / n);
  ovxhip::Storage<zf> sx(rows * n, x.data()), sw(n, w.data()), sy(rows * n);
  dev_fc(sx.device(OVX_ACCESS_READ), sw.device(OVX_ACCESS_READ), sy.device(OVX_ACCESS_WRITE));
  zf const *yh = sy.host(OVX_ACCESS_READ);
  CHECK(std::memcmp(yh, yc.data(), yc.size() * sizeof(zf)) == 0);
  CHECK(sx.valid() == ((1u << OVX_LOC_HOST) | (1u << OVX_LOC_DEVICE)));

  std::printf("%s\n", fails ? "FAIL" : "OK");
  return fails;
}
