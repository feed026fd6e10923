// oic_tables.h -- host-side DP tables of prefixEditDistance, built once per context and uploaded.
//
//   Error_Bound[i]      = ceil(i * maxErate)                      (prefixEditDistance.C:58-61)
//   Edit_Match_Limit[e] = smallest n with P[>= e-1 errors in n trials at maxErate*alignNoise] > 1e-4,
//                         minus one; exact binomial sum for n <= 35, normal shortcut (z <= 3.62) or
//                         Poisson sum above; entries from 2000 on follow the fitted line with slope
//                         0.982064188397525/e + 0.067835741959926 for 21-bit read lengths
//                         (Binomial_Bound.C:35-195).
// maxErate is the float the command line parses (strtof) widened to double.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace oic {

constexpr uint32_t kMaxReadLenBits = 21;
constexpr uint32_t kMaxReadLen = (1u << kMaxReadLenBits) - 1u;

inline int binomial_bound(int e, double p, int start) {
  const double q = 1.0 - p;
  if (start < e) start = e;
  for (int n = start; n < (int)kMaxReadLen; n++) {
    if (n <= 35) {
      double sum = 0.0, p_pow = 1.0, q_pow = std::pow(q, n);
      int32_t coeff = 1, ct = 0;
      for (int k = 0; k < e && 1.0 - sum > 1e-4; k++) {
        const double x = coeff * p_pow * q_pow;
        sum += x;
        coeff = (int32_t)((uint32_t)coeff * (uint32_t)(n - ct));   // int arithmetic of the reference, wrapping
        coeff /= ++ct;
        p_pow *= p;
        q_pow /= q;
      }
      if (1.0 - sum > 1e-4) return n;
    } else {
      const double z = (e - 0.5 - n * p) / std::sqrt(n * p * q);
      if (z <= 3.62) return n;
      double sum = 0.0, mu_pow = 1.0, fact = 1.0;
      const double pois = std::exp(-n * p);
      for (int k = 0; k < e; k++) {
        sum += mu_pow * pois / fact;
        mu_pow *= n * p;
        fact *= k + 1;
      }
      if (1.0 - sum > 1e-4) return n;
    }
  }
  return (int)kMaxReadLen;
}

inline void match_limit(std::vector<int32_t> &ml, double rate, int32_t max_errors) {
  ml.assign((size_t)max_errors + 1, 0);
  int32_t e = 0, s = 1;
  const int32_t l = max_errors < 2000 ? max_errors : 2000;
  while (e <= 1) ml[e++] = 0;
  while (e < l) {
    s = binomial_bound(e - 1, rate, s);
    ml[e] = s - 1;
    e++;
  }
  const double sl = 0.982064188397525 / rate + 0.067835741959926;
  double vl = ml[e - 1] + sl;
  while (e < max_errors) {
    ml[e] = (int32_t)std::ceil(vl);
    vl += sl;
    e++;
  }
}

struct Tables {
  double maxErate = 0, alignNoise = 1;
  uint32_t max_errors = 0;
  double bmv = 0, slope_min = 0;
  std::vector<int32_t> error_bound;   // [kMaxReadLen + 1]
  std::vector<int32_t> match_limit;   // [max_errors + 1]

  void build(double cli_erate, double cli_noise) {
    maxErate = (double)(float)cli_erate;
    alignNoise = (double)(float)cli_noise;
    max_errors = 1 + (uint32_t)(int)std::ceil(maxErate * kMaxReadLen);
    slope_min = (maxErate > 0.06) ? 1.0 : 0.20;
    bmv = maxErate / (1 + maxErate);
    error_bound.resize((size_t)kMaxReadLen + 1);
    for (uint32_t i = 0; i <= kMaxReadLen; i++) error_bound[i] = (int32_t)std::ceil(i * maxErate);
    oic::match_limit(match_limit, maxErate * alignNoise, (int32_t)max_errors);
  }
};

}  // namespace oic
