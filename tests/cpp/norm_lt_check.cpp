// Host check of norm_lt (cms_device.h): the squared-distance prefilter must
// decide norm_seq(a) < lim exactly as the square root does.  Inputs: random
// vectors and limits, plus limits placed within a few ulps of |a| in both
// directions (where the prefilter's band is exercised).  Prints the number
// of cases and of mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "cms_device.h"

using namespace cms;

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> u(-12.0, 12.0), lr(0.0, 25.0);
  int64_t n = 0, bad = 0, near = 0;
  for (int it = 0; it < 2000000; ++it) {
    const V3 a = v3(u(rng), u(rng), u(rng));
    const double nrm = norm_seq(a);
    double lims[8];
    lims[0] = lr(rng);
    lims[1] = nrm;
    lims[2] = std::nextafter(nrm, 0.0);
    lims[3] = std::nextafter(nrm, 1e300);
    lims[4] = std::nextafter(std::nextafter(nrm, 1e300), 1e300);
    lims[5] = nrm * (1.0 + 0x1p-41);
    lims[6] = nrm * (1.0 - 0x1p-41);
    lims[7] = nrm * (1.0 + (u(rng) * 1e-12));
    for (double lim : lims) {
      const bool want = std::sqrt(dot_seq(a, a)) < lim;
      const bool got = norm_lt(a, lim);
      ++n;
      bad += want != got;
      const double q = dot_seq(a, a), l2 = lim * lim;
      near += !(q < l2 * (1.0 - 0x1p-40)) && !(q > l2 * (1.0 + 0x1p-40));
    }
  }
  // special values
  const double specials[] = {0.0, 1e-300, 1e300, INFINITY, NAN};
  for (double x : specials)
    for (double lim : specials) {
      const V3 a = v3(x, 0.0, 0.0);
      ++n;
      bad += (std::sqrt(dot_seq(a, a)) < lim) != norm_lt(a, lim);
    }
  std::printf("cases %lld near-band %lld mismatches %lld\n", (long long)n, (long long)near, (long long)bad);
  return bad != 0;
}
