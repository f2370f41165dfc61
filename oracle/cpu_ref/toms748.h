// Oracle (test infrastructure only): restatement of the bracketing root
// finder the reference calls at ExactRiemannSolverBase.cpp:1540-1542,
// boost::math::tools::toms748_solve (Boost.Math >= 1.72, the version floor of
// CMakeLists.txt:54; Boost is not vendored in the reference).  The algorithm
// is Alefeld, Potra & Shi, "Algorithm 748: enclosing zeros of continuous
// functions", ACM TOMS 21(3), 1995, in the step sequence Boost implements:
// secant step, quadratic step (2 Newton iterations), then repeated
// {cubic-or-quadratic(2), cubic-or-quadratic(3), double-length secant,
// optional bisection} with the bracket() safeguard.  Iterate-level agreement
// with Boost's object code cannot be checked here (Boost is absent): parity of
// this function is pinned only through the shock states of the reference's
// Riemann fixtures and the closed-form SG Hugoniot.
#pragma once
#include <cmath>
#include <cfloat>
#include <cstdint>
#include <utility>

namespace orc {

namespace t748 {

inline int sgn(double x) { return (x > 0) - (x < 0); }

inline double safe_div(double num, double denom, double r) {
  if (std::fabs(denom) < 1) {
    if (std::fabs(denom * DBL_MAX) <= std::fabs(num)) return r;
  }
  return num / denom;
}

inline double secant_interpolate(double a, double b, double fa, double fb) {
  double tol = DBL_EPSILON * 5;
  double c = a - (fa / (fb - fa)) * (b - a);
  if ((c <= a + std::fabs(a) * tol) || (c >= b - std::fabs(b) * tol)) return (a + b) / 2;
  return c;
}

inline double quadratic_interpolate(double a, double b, double d, double fa, double fb, double fd, unsigned count) {
  double B = safe_div(fb - fa, b - a, DBL_MAX);
  double A = safe_div(fd - fb, d - b, DBL_MAX);
  A = safe_div(A - B, d - a, 0.0);
  if (A == 0) return secant_interpolate(a, b, fa, fb);
  double c = (sgn(A) * sgn(fa) > 0) ? a : b;
  for (unsigned i = 1; i <= count; ++i)
    c -= safe_div(fa + (B + A * (c - b)) * (c - a), B + A * (2 * c - a - b), 1 + c - a);
  if ((c <= a) || (c >= b)) c = secant_interpolate(a, b, fa, fb);
  return c;
}

inline double cubic_interpolate(double a, double b, double d, double e, double fa, double fb, double fd, double fe) {
  double q11 = (d - e) * fd / (fe - fd);
  double q21 = (b - d) * fb / (fd - fb);
  double q31 = (a - b) * fa / (fb - fa);
  double d21 = (b - d) * fd / (fd - fb);
  double d31 = (a - b) * fb / (fb - fa);
  double q22 = (d21 - q11) * fb / (fe - fb);
  double q32 = (d31 - q21) * fa / (fd - fa);
  double d32 = (d31 - q21) * fd / (fd - fa);
  double q33 = (d32 - q22) * fa / (fe - fa);
  double c = q31 + q32 + q33 + a;
  if ((c <= a) || (c >= b)) c = quadratic_interpolate(a, b, d, fa, fb, fd, 3);
  return c;
}

template <class F>
void bracket(F& f, double& a, double& b, double c, double& fa, double& fb, double& d, double& fd) {
  double tol = DBL_EPSILON * 2;
  if ((b - a) < 2 * tol * a)
    c = a + (b - a) / 2;
  else if (c <= a + std::fabs(a) * tol)
    c = a + std::fabs(a) * tol;
  else if (c >= b - std::fabs(b) * tol)
    c = b - std::fabs(b) * tol;
  double fc = f(c);
  if (fc == 0) {
    a = c; fa = 0; d = 0; fd = 0;
    return;
  }
  if (sgn(fa) * sgn(fc) < 0) {
    d = b; fd = fb; b = c; fb = fc;
  } else {
    d = a; fd = fa; a = c; fa = fc;
  }
}

inline bool prof_check(double fa, double fb, double fd, double fe) {
  const double min_diff = DBL_MIN * 32;
  return (std::fabs(fa - fb) < min_diff) || (std::fabs(fa - fd) < min_diff) || (std::fabs(fa - fe) < min_diff) ||
         (std::fabs(fb - fd) < min_diff) || (std::fabs(fb - fe) < min_diff) || (std::fabs(fd - fe) < min_diff);
}

}  // namespace t748

// Returns the final bracket [a,b]; `max_iter` in: budget, out: used.
// `ok` is false where Boost would raise a domain error (a >= b, or f(a) and
// f(b) of equal sign).
template <class F, class Tol>
std::pair<double, double> toms748_solve(F& f, double ax, double bx, double fax, double fbx, Tol tol,
                                        std::uintmax_t& max_iter, bool& ok) {
  using namespace t748;
  ok = true;
  std::uintmax_t count = max_iter;
  double a = ax, b = bx, fa = fax, fb = fbx, c, u, fu, a0, b0, d = 0, fd, e, fe;
  const double mu = 0.5;
  if (a >= b) { ok = false; return {a, b}; }
  if (tol(a, b) || (fa == 0) || (fb == 0)) {
    max_iter = 0;
    if (fa == 0) b = a;
    else if (fb == 0) a = b;
    return {a, b};
  }
  if (sgn(fa) * sgn(fb) > 0) { ok = false; return {a, b}; }
  fe = e = fd = 1e5;
  if (fa != 0) {
    c = secant_interpolate(a, b, fa, fb);
    bracket(f, a, b, c, fa, fb, d, fd);
    --count;
    if (count && (fa != 0) && !tol(a, b)) {
      c = quadratic_interpolate(a, b, d, fa, fb, fd, 2);
      e = d; fe = fd;
      bracket(f, a, b, c, fa, fb, d, fd);
      --count;
    }
  }
  while (count && (fa != 0) && !tol(a, b)) {
    a0 = a; b0 = b;
    if (prof_check(fa, fb, fd, fe)) c = quadratic_interpolate(a, b, d, fa, fb, fd, 2);
    else c = cubic_interpolate(a, b, d, e, fa, fb, fd, fe);
    e = d; fe = fd;
    bracket(f, a, b, c, fa, fb, d, fd);
    if ((0 == --count) || (fa == 0) || tol(a, b)) break;
    if (prof_check(fa, fb, fd, fe)) c = quadratic_interpolate(a, b, d, fa, fb, fd, 3);
    else c = cubic_interpolate(a, b, d, e, fa, fb, fd, fe);
    bracket(f, a, b, c, fa, fb, d, fd);
    if ((0 == --count) || (fa == 0) || tol(a, b)) break;
    if (std::fabs(fa) < std::fabs(fb)) { u = a; fu = fa; }
    else { u = b; fu = fb; }
    c = u - 2 * (fu / (fb - fa)) * (b - a);
    if (std::fabs(c - u) > (b - a) / 2) c = a + (b - a) / 2;
    e = d; fe = fd;
    bracket(f, a, b, c, fa, fb, d, fd);
    if ((0 == --count) || (fa == 0) || tol(a, b)) break;
    if ((b - a) < mu * (b0 - a0)) continue;
    e = d; fe = fd;
    bracket(f, a, b, a + (b - a) / 2, fa, fb, d, fd);
    --count;
  }
  max_iter -= count;
  if (fa == 0) b = a;
  else if (fb == 0) a = b;
  return {a, b};
}

}  // namespace orc
