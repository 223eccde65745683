// root_ref.hpp - oracle's pow(x, 1./D) (TEST INFRASTRUCTURE ONLY).
//
// The reference evaluates pow(vol, 1./DIM) with glibc (S3:mpm.cpp:1089,
// 1399; S3:particle_flow.cpp:473; S3:particle.cpp:643).  Both the oracle and
// the HIP library replace it by the same deterministic IEEE sequence so that
// split/merge decisions agree bit for bit; tests/test_oracle_math.py checks
// this sequence against glibc pow(x, 1./3.) and sqrt on random inputs.
//
// D = 2: sqrt(x) (correctly rounded; pow(x, 0.5) of glibc rounds the same
//        value except in rare near-midpoint cases).
// D = 3: y = x^(1/3 - eps) with 1./3. = 1/3 - eps, eps = 2^-54/3:
//        Newton for the cube root, one exact-residual correction with fma,
//        and the first-order factor (1 - eps ln x).
#pragma once
#include <cmath>

namespace oracle {

inline double cbrt_pow13(double x) {
  if (!(x > 0) || std::isinf(x)) {
    if (x == 0) return 0.0;
    if (std::isinf(x) && x > 0) return x;
    return std::nan("");
  }
  int e;
  double m = std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  int q = e / 3;
  int r = e - 3 * q;
  if (r < 0) {
    r += 3;
    q -= 1;
  }
  double y0 = std::ldexp(m, r);  // in [0.5, 4)
  // initial guess: linear fit of cbrt on [0.5, 4), then Newton
  double t = 0.6 + 0.4 * y0 - 0.02 * y0 * y0;
  for (int it = 0; it < 6; it++) t = t - (t * t * t - y0) / (3.0 * t * t);
  // exact residual y0 - t^3 with error-free products
  double p = t * t;
  double pe = std::fma(t, t, -p);
  double c3 = p * t;
  double ce = std::fma(p, t, -c3);
  double res = ((y0 - c3) - ce) - pe * t;
  double delta = res / (3.0 * p);
  // ln x = e ln2 + ln m, ln m = 2 atanh((m-1)/(m+1)) (|s| <= 1/3, five
  // terms, |err| < 1e-6: the eps term then moves y by < 1e-7 ulp)
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double lnm = 2.0 * s * (1.0 + s2 * (1.0 / 3.0 + s2 * (0.2 + s2 * (1.0 / 7.0 + s2 * (1.0 / 9.0)))));
  double lnx = (double)e * 0.6931471805599453 + lnm;
  const double eps = 1.850371707708594e-17;  // 1/3 - 1./3.
  double corr = delta - eps * lnx * t;
  double y = t + corr;
  return std::ldexp(y, q);
}

inline double root_dim(int dim, double x) { return dim == 2 ? std::sqrt(x) : cbrt_pow13(x); }

}  // namespace oracle
