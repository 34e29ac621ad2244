// Compiles include/ovxhip/ovxhip.hpp against the C-ABI and runs the error
// paths that need no device: each C-ABI status must surface as the exception
// the header maps it to (unimplemented, std::invalid_argument), the way the
// reference's backends report a declined shape or a bad argument.
#include <cstdio>
#include <stdexcept>

#include "ovxhip/ovxhip.hpp"

static int fails = 0;

// compiled, not run (plans need a device): the split-storage overloads
void split_overloads(ovxhip::Fftm<OVX_ROW, OVX_FFT_FWD> &f, ovxhip::Fft<OVX_FFT_INV> &g, ovxhip::Fastconv &c,
                     float *re, float *im)
{
  ovxhip::split_ptr z(re, im);
  ovxhip::const_split_ptr cz(re, im);
  f.out_of_place(cz, 16, 1, z, 16, 1, 4, 16);
  f.in_place(z, 16, 1, 4, 16);
  g.out_of_place(cz, 1, z, 1);
  g.in_place(z, 1);
  c(cz, 16, cz, 0, z, 16, 4);
}

void real_split_overloads(ovxhip::FftmReal<OVX_ROW, OVX_FFT_FWD> &r, ovxhip::FftmReal<OVX_ROW, OVX_FFT_INV> &c,
                          float *x, float *re, float *im)
{
  r.out_of_place(x, 16, 1, ovxhip::split_ptr(re, im), 9, 1, 4, 16);
  c.out_of_place(ovxhip::const_split_ptr(re, im), 9, 1, x, 16, 1, 4, 16);
}

#define EXPECT_THROW(stmt, type)                                  \
  do {                                                            \
    try {                                                         \
      stmt;                                                       \
      std::printf("no exception: %s\n", #stmt);                   \
      ++fails;                                                    \
    } catch (type const &) {                                      \
    } catch (std::exception const &e) {                           \
      std::printf("wrong exception for %s: %s\n", #stmt, e.what()); \
      ++fails;                                                    \
    }                                                             \
  } while (0)

int main()
{
  // a bad axis is an invalid argument
  EXPECT_THROW((ovxhip::Fftm<5, OVX_FFT_FWD>(4, 16, 1.0f)), std::invalid_argument);
  // a prime length is declined: the dispatcher's next backend takes it
  EXPECT_THROW((ovxhip::Fftm<OVX_ROW, OVX_FFT_FWD>(4, 19, 1.0f)), ovxhip::unimplemented);
  EXPECT_THROW((ovxhip::Fft<OVX_FFT_INV>(68, 1.0f)), ovxhip::unimplemented);
  // real Fftm needs an even length
  EXPECT_THROW((ovxhip::FftmReal<OVX_ROW, OVX_FFT_FWD>(4, 15, 1.0f)), ovxhip::unimplemented);
  // a reference longer than the input
  EXPECT_THROW((ovxhip::Correlation(0, 10, 5)), std::invalid_argument);
  try {
    ovxhip::check(OVX_ERR_UNSUPPORTED);
    ++fails;
  } catch (ovxhip::error const &e) {
    if (e.code() != OVX_ERR_UNSUPPORTED) ++fails;
  }
  if (ovx_plan_destroy(nullptr) != OVX_ERR_INVALID) ++fails;
  // a plan the cache could not make is not cached: the error repeats
  EXPECT_THROW(ovxhip::cached_fastconv(4, 4096, 1.0f, OVX_FLAG_BF16X3), ovxhip::unimplemented);
  EXPECT_THROW(ovxhip::cached_fastconv(4, 4096, 1.0f, OVX_FLAG_BF16X3), ovxhip::unimplemented);
  // user storage on the host needs no device: the application's array is
  // the host location, valid, and released / admitted without a copy
  {
    ovxhip::cscalar a[8] = {};
    ovxhip::Storage<ovxhip::cscalar> s(8, a);
    if (s.host(OVX_ACCESS_READ) != a || s.valid() != 1u << OVX_LOC_HOST) ++fails;
    s.release();
    EXPECT_THROW(s.host(), ovxhip::error);
    s.admit(true);
    if (s.host(OVX_ACCESS_WRITE) != a) ++fails;
    EXPECT_THROW(s.ptr(7), std::invalid_argument);
  }
  std::printf("%s\n", fails ? "FAIL" : "OK");
  return fails;
}
