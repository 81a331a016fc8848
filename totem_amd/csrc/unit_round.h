// Host helper: FP64 unit-modulus value -> FP32 pair with the squared modulus
// nearest 1.
//
// Rounding the two components independently leaves |w|^2 - 1 of a few 1e-8 with
// a fixed sign per value.  Twiddles and propagator entries are applied hundreds
// of times to every probe, so such a bias adds up coherently; among the
// neighbouring floats of each component this picks the pair closest to the unit
// circle (ties: closest value), which is within one ulp of the FP64 value.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace totem {

inline float2 unit_float2(double c, double s) {
  const float c0 = (float)c, s0 = (float)s;
  float best_c = c0, best_s = s0;
  double best = 1e300;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      float cc = c0, ss = s0;
      if (i) cc = std::nextafter(c0, i < 0 ? -2.0f : 2.0f);
      if (j) ss = std::nextafter(s0, j < 0 ? -2.0f : 2.0f);
      const double dn = std::fabs((double)cc * cc + (double)ss * ss - 1.0);
      const double dv = std::fabs(cc - c) + std::fabs(ss - s);
      const double score = dn * 1e6 + dv;
      if (score < best) {
        best = score;
        best_c = cc;
        best_s = ss;
      }
    }
  return make_float2(best_c, best_s);
}

}  // namespace totem
