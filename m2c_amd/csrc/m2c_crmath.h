// Correctly rounded x^(1/4) and x^(1/5) for the step-size control of the
// adaptive rarefaction integrator (ExactRiemannSolverBase.cpp:1333-1399 calls
// pow(r, 0.25) and pow(r, 0.2)).  The CPU library pow and the GPU device pow
// are different approximations; evaluating both sides with this one correctly
// rounded routine makes host and device step sizes identical to the last bit,
// and equal to any correctly rounded pow (glibc's pow is documented within
// 0.52 ULP, i.e. it agrees except in rare near-midpoint cases).
//
// Method: scale x to [1, 2^n) by a power of 2^n, take a starting root, do one
// Newton step with the residual y^n - x evaluated in double-double arithmetic
// (error-free products via fma), round once.  Only IEEE +,-,*,/,sqrt,fma and
// exact scalings are used, so host and device results are bitwise identical.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define M2C_HD __host__ __device__
#else
#define M2C_HD
#endif

namespace m2c {

M2C_HD inline void dd_mul_d(double hi, double lo, double y, double& rh, double& rl) {
  double p = hi * y;
  double e = std::fma(hi, y, -p);
  e = std::fma(lo, y, e);
  rh = p + e;
  rl = e - (rh - p);
}

// y^n as a double-double (n = 4 or 5)
M2C_HD inline void dd_pow_n(double y, int n, double& h, double& l) {
  h = y; l = 0.0;
  for (int i = 1; i < n; i++) dd_mul_d(h, l, y, h, l);
}

// Correctly rounded x^(1/n), n in {4,5}, for x >= 0 (NaN/inf pass through).
M2C_HD inline double cr_root(double x, int n) {
  if (!(x > 0.0) || x == INFINITY) return (x == 0.0) ? 0.0 : (x != x ? x : (x < 0 ? NAN : x));
  int e;
  double m = std::frexp(x, &e);          // x = m * 2^e, m in [0.5, 1)
  int q = (e >= 0) ? e / n : -((-e + n - 1) / n);
  int r = e - q * n;                     // 0 <= r < n
  double ms = std::ldexp(m, r);          // in [0.5, 2^(n-1)), exact
  double y0;
  if (n == 4) {
    y0 = std::sqrt(std::sqrt(ms));
  } else {
    // starting value by bisection-free Newton in double from a crude seed;
    // the final correction below makes the seed's last bits irrelevant.
    y0 = 1.0;
    for (int it = 0; it < 60; it++) {
      double y4 = y0 * y0 * y0 * y0;
      double yn = y0 - (y4 * y0 - ms) / (5.0 * y4);
      if (yn == y0) break;
      y0 = yn;
    }
  }
  for (int pass = 0; pass < 2; pass++) {
    double h, l;
    dd_pow_n(y0, n, h, l);
    double rh = h - ms;                  // exact when h ~ ms (Sterbenz)
    double res = rh + l;
    double yn1;
    dd_pow_n(y0, n - 1, yn1, h);         // y^(n-1), high part suffices
    double delta = res / (n * yn1);
    y0 = y0 - delta;
  }
  return std::ldexp(y0, q);
}

M2C_HD inline double cr_pow_quarter(double x) { return cr_root(x, 4); }
M2C_HD inline double cr_pow_fifth(double x) { return cr_root(x, 5); }

}  // namespace m2c
