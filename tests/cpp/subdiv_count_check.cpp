// Host check of subdiv_count (cms_device.h): the closed-form estimate with
// exact correction steps must equal the defining loop -- the number of
// i < min(MAX_SUBDIV, ceil(angle/delta) + 1) with delta*(i + 1/2) <= angle,
// stopping at the first failure (vec_sub_div, _core.py:2429-2484; a_i is
// non-decreasing, so this equals its mask count).  Inputs: random deltas
// and angles, angles placed exactly on and one ulp around the points
// delta*(i + 1/2), and special values.  Prints cases and mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "cms_device.h"

using namespace cms;

static int by_loop(double delta, double angle) {
  const double raw = angle / delta;
  if (!(raw > 0.0) || !(raw < 1e300)) return 0;
  int n = 0;
  const int lim = (int)std::fmin((double)MAX_SUBDIV, std::ceil(raw) + 1.0);
  for (int i = 0; i < lim; ++i) {
    const double a = delta * ((double)i + 0.5);
    if (a <= angle) n = i + 1; else break;
  }
  return n;
}

int main() {
  std::mt19937_64 rng(2024);
  std::uniform_real_distribution<double> dens(0.5, 40.0), rad(0.05, 4.0), ang(0.0, 7.0);
  std::uniform_int_distribution<int> pick(0, 120);
  int64_t n = 0, bad = 0, exact_hits = 0;
  for (int it = 0; it < 3000000; ++it) {
    const double delta = 1.0 / (std::sqrt(dens(rng)) * rad(rng));
    double angles[5];
    angles[0] = ang(rng);
    const double on = delta * ((double)pick(rng) + 0.5);
    angles[1] = on;
    angles[2] = std::nextafter(on, 0.0);
    angles[3] = std::nextafter(on, 1e300);
    angles[4] = 2.0 * M_PI;
    for (double a : angles) {
      const int want = by_loop(delta, a), got = subdiv_count(delta, a);
      ++n;
      bad += want != got;
      exact_hits += a == on;
    }
  }
  const double specials[] = {0.0, -1.0, 1e-300, 1e300, INFINITY, -INFINITY, NAN, 2.0 * M_PI, 1e-3};
  for (double d : specials)
    for (double a : specials) {
      ++n;
      bad += by_loop(d, a) != subdiv_count(d, a);
    }
  std::printf("cases %lld on-point %lld mismatches %lld\n", (long long)n, (long long)exact_hits, (long long)bad);
  return bad != 0;
}
