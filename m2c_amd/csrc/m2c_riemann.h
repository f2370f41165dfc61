// Device exact two-material Riemann solver (one problem per lane), restating
// ExactRiemannSolverBase (ExactRiemannSolverBase.cpp:59-1896) and its
// non-adaptive fallback (:2243-2989) for gfx950.
//
// GPU shape: the solve is branchy and iterative, so it runs in its own kernel
// over a compacted list of interface faces (never inside the stencil
// kernel).  Each lane keeps the two cached rarefaction integration paths
// (integrationPath1/3, ExactRiemannSolverBase.h:29-30) in a private slab of
// HBM scratch, interleaved by lane so that lanes at the same path position
// read one cache line.  The fallback solver is a second pass of the same loop
// (no recursion).  pow(x, 0.25) / pow(x, 0.2) of the step control are the
// correctly rounded roots of m2c_crmath.h.
#pragma once
#include "m2c_physics.h"
#include "m2c_crmath.h"

namespace m2c {

struct RiemannCfg {
  int maxIts_main, maxIts_bracket, maxIts_shock, numSteps;
  double tol_main, tol_shock, tol_rare, min_pressure, failure_threshold, pressure_at_failure;
};

// Per-lane integration path: (p, rho, u) triplets, point-major, lane-minor.
struct Path {
  double* base;   // this lane's column: element (pt, comp) at base[(pt*3+comp)*stride]
  long stride;
  int n, cap;
  int hint;       // index the last search returned (a starting guess only)
  bool overflow;
  __device__ __forceinline__ double get(int pt, int c) const { return base[((long)pt * 3 + c) * stride]; }
  __device__ __forceinline__ void push(double p, double r, double u) {
    if (n >= cap) { overflow = true; return; }
    base[((long)n * 3 + 0) * stride] = p;
    base[((long)n * 3 + 1) * stride] = r;
    base[((long)n * 3 + 2) * stride] = u;
    n++;
  }
  __device__ __forceinline__ void reset(double p, double r, double u) {
    n = 0;
    hint = 0;
    push(p, r, u);
  }
};

struct RStats {
  int it_main, fevals;
  bool trivial, fallback, fatal;
};

// EOS evaluations of the solver: e(rho, p) with c^2, e alone, and the
// Hugoniot function.  Inlined: as out-of-line calls every evaluation saved
// and restored live registers through scratch memory, which dominated the
// latency of the solver's serial chain.
struct EC2 {
  double e, c2;
};
__device__ __forceinline__ EC2 ool_ec2(const DevEos* E, double rho, double p) {
  EC2 r;
  r.e = eos_e(*E, rho, p);
  r.c2 = eos_c2(*E, rho, r.e);
  return r;
}
__device__ __forceinline__ double ool_e(const DevEos* E, double rho, double p) { return eos_e(*E, rho, p); }
// HugoniotEquation (ExactRiemannSolverBase.h:66-83)
__device__ __forceinline__ double ool_hugo(const DevEos* E, double r, double ps, double he, double pavg, double oor) {
  return eos_e(*E, r, ps) - he + pavg * (1.0 / r - oor);
}

struct DevRiemann {
  const DevMats& M;
  const RiemannCfg& P;
  Path path1, path3;
  bool nonadaptive;
  RStats st;

  __device__ DevRiemann(const DevMats& m, const RiemannCfg& c, double* slab, long stride, int cap)
      : M(m), P(c), nonadaptive(false) {
    path1.base = slab;
    path1.stride = stride;
    path1.cap = cap;
    path1.n = 0;
    path1.hint = 0;
    path1.overflow = false;
    path3.base = slab + (long)cap * 3 * stride;
    path3.stride = stride;
    path3.cap = cap;
    path3.n = 0;
    path3.hint = 0;
    path3.overflow = false;
    st = RStats{0, 0, false, false, false};
  }

// bracket() of the TOMS748 driver, as a macro over the shock() locals
#define BRACKET(cin)                                                          \
  do {                                                                        \
    double cc_ = (cin);                                                       \
    const double tb_ = DBL_EPSILON * 2;                                       \
    if ((b - a) < 2 * tb_ * a) cc_ = a + (b - a) / 2;                         \
    else if (cc_ <= a + fabs(a) * tb_) cc_ = a + fabs(a) * tb_;              \
    else if (cc_ >= b - fabs(b) * tb_) cc_ = b - fabs(b) * tb_;              \
    double fc_ = ool_hugo(&E, cc_, ps, he, pavg, oor);                       \
    if (fc_ == 0) { a = cc_; fa = 0; d = 0; fd = 0; }                         \
    else if (sgn(fa) * sgn(fc_) < 0) { d = b; fd = fb; b = cc_; fb = fc_; }   \
    else { d = a; fd = fa; a = cc_; fa = fc_; }                               \
  } while (0)

  // ------------------------------------------------------------------ shock
  // HugoniotEquation (ExactRiemannSolverBase.h:66-83) and the bracketing +
  // TOMS748 of :1422-1621
  __device__ bool shock(const DevEos& E, double rho, double u, double p, double ps, double rhos0, double rhos1,
                        int wave, double& rhos, double& us) {
    const double he = ool_e(&E, rho, p);
    const double pavg = 0.5 * (p + ps);
    const double oor = 1.0 / rho;
#define HUGO(r) ool_hugo(&E, (r), ps, he, pavg, oor)
    double f0, f1;
    double drho = fmax_std(fabs(rhos0 - rhos1), 0.001 * rhos0);
    bool found0 = false, found1 = false;
    if (fmin_std(rhos0, rhos1) >= rho) {
      f0 = HUGO(rhos0);
      f1 = HUGO(rhos1);
      if (f0 * f1 <= 0) {
        if (rhos0 > rhos1) { double t = rhos0; rhos0 = rhos1; rhos1 = t; t = f0; f0 = f1; f1 = t; }
        found0 = found1 = true;
      } else {
        rhos0 = rhos1;
        f0 = f1;
      }
    } else {
      if (rhos1 > rhos0) { rhos0 = rhos1; f0 = HUGO(rhos0); }
      if (rhos0 < rho) { rhos0 = rho; f0 = HUGO(rhos0); found0 = true; }
      else f0 = HUGO(rhos0);
    }
    if (!found0 || !found1) {
      int i = 0;
      double factor = 1.5, tmp, ftmp;
      rhos1 = rhos0;
      f1 = f0;
      while (!found0) {
        if (++i >= P.maxIts_shock) return false;
        tmp = rhos1; ftmp = f1;
        rhos1 = rhos0; f1 = f0;
        rhos0 = rhos1 - factor * drho;
        if (rhos0 <= rho) { rhos0 = rho; found0 = true; }
        f0 = HUGO(rhos0);
        if (f0 * f1 <= 0) {
          found0 = found1 = true;
        } else {
          rhos1 = tmp; f1 = ftmp;
          tmp = rhos0; ftmp = f0;
          rhos0 = rhos1; f0 = f1;
          rhos1 = rhos0 + factor * drho;
          f1 = HUGO(rhos1);
          if (f0 * f1 <= 0) found0 = found1 = true;
          else { rhos0 = tmp; f0 = ftmp; drho = rhos1 - rhos0; }
        }
      }
      if (!found1) {
        i = 0;
        factor = 2.5;
        while (!found1) {
          if (++i >= P.maxIts_shock) return false;
          rhos0 = rhos1; f0 = f1;
          rhos1 = rhos0 + factor * drho;
          f1 = HUGO(rhos1);
          if (f0 * f1 <= 0) found1 = true;
          else drho = rhos1 - rhos0;
        }
      }
    }
    double sa, sb;
    if (f0 == 0.0) { sa = sb = rhos0; }
    else if (f1 == 0.0) { sa = sb = rhos1; }
    else {
      // boost::math::tools::toms748_solve with tol(a,b) = b-a < min(tol_shock*min(r0,r1), 0.001*(r1-r0))
      const double ltol = P.tol_shock * fmin_std(rhos0, rhos1);
      const double ttol = fmin_std(ltol, 0.001 * (rhos1 - rhos0));
      double a = rhos0, b = rhos1, fa = f0, fb = f1;
      if (a >= b) { st.fatal = true; return false; }
      long count = P.maxIts_shock;
      bool done = (b - a < ttol) || (fa == 0) || (fb == 0);
      if (done) {
        if (fa == 0) b = a;
        else if (fb == 0) a = b;
      } else {
        if (sgn(fa) * sgn(fb) > 0) { st.fatal = true; return false; }
        double c, d = 0, fd, e, fe, u_, fu, a0, b0;
        fe = e = fd = 1e5;
        c = secant_i(a, b, fa, fb);
        BRACKET(c);
        --count;
        if (count && (fa != 0) && !(b - a < ttol)) {
          c = quad_i(a, b, d, fa, fb, fd, 2);
          e = d; fe = fd;
          BRACKET(c);
          --count;
        }
        while (count && (fa != 0) && !(b - a < ttol)) {
          a0 = a; b0 = b;
          if (prof(fa, fb, fd, fe)) c = quad_i(a, b, d, fa, fb, fd, 2);
          else c = cubic_i(a, b, d, e, fa, fb, fd, fe);
          e = d; fe = fd;
          BRACKET(c);
          if ((0 == --count) || (fa == 0) || (b - a < ttol)) break;
          if (prof(fa, fb, fd, fe)) c = quad_i(a, b, d, fa, fb, fd, 3);
          else c = cubic_i(a, b, d, e, fa, fb, fd, fe);
          BRACKET(c);
          if ((0 == --count) || (fa == 0) || (b - a < ttol)) break;
          if (fabs(fa) < fabs(fb)) { u_ = a; fu = fa; }
          else { u_ = b; fu = fb; }
          c = u_ - 2 * (fu / (fb - fa)) * (b - a);
          if (fabs(c - u_) > (b - a) / 2) c = a + (b - a) / 2;
          e = d; fe = fd;
          BRACKET(c);
          if ((0 == --count) || (fa == 0) || (b - a < ttol)) break;
          if ((b - a) < 0.5 * (b0 - a0)) continue;
          e = d; fe = fd;
          BRACKET(a + (b - a) / 2);
          --count;
        }
        if (fa == 0) b = a;
        else if (fb == 0) a = b;
      }
      sa = a;
      sb = b;
    }
    rhos = 0.5 * (sa + sb);
    const double du = -(ps - p) * (1.0 / rhos - 1.0 / rho);
    if (du < 0) return false;
    if (eos_check(E, rhos, ps)) return false;
    us = (wave == 1) ? u - sqrt(du) : u + sqrt(du);
    return true;
#undef HUGO
  }

  // -------------------------------------------------- TOMS748 building blocks
  __device__ static __forceinline__ int sgn(double x) { return (x > 0) - (x < 0); }
  __device__ static __forceinline__ double fmax_std(double a, double b) { return (a < b) ? b : a; }
  __device__ static __forceinline__ double fmin_std(double a, double b) { return (b < a) ? b : a; }
  __device__ static __forceinline__ double safe_div(double num, double den, double r) {
    if (fabs(den) < 1) {
      if (fabs(den * DBL_MAX) <= fabs(num)) return r;
    }
    return num / den;
  }
  __device__ static __forceinline__ double secant_i(double a, double b, double fa, double fb) {
    const double tol = DBL_EPSILON * 5;
    double c = a - (fa / (fb - fa)) * (b - a);
    if ((c <= a + fabs(a) * tol) || (c >= b - fabs(b) * tol)) return (a + b) / 2;
    return c;
  }
  __device__ static double quad_i(double a, double b, double d, double fa, double fb, double fd, unsigned count) {
    double B = safe_div(fb - fa, b - a, DBL_MAX);
    double A = safe_div(fd - fb, d - b, DBL_MAX);
    A = safe_div(A - B, d - a, 0.0);
    if (A == 0) return secant_i(a, b, fa, fb);
    double c = (sgn(A) * sgn(fa) > 0) ? a : b;
    for (unsigned i = 1; i <= count; ++i) c -= safe_div(fa + (B + A * (c - b)) * (c - a), B + A * (2 * c - a - b), 1 + c - a);
    if ((c <= a) || (c >= b)) c = secant_i(a, b, fa, fb);
    return c;
  }
  __device__ static double cubic_i(double a, double b, double d, double e, double fa, double fb, double fd,
                                   double fe) {
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
    if ((c <= a) || (c >= b)) c = quad_i(a, b, d, fa, fb, fd, 3);
    return c;
  }
  __device__ static __forceinline__ bool prof(double fa, double fb, double fd, double fe) {
    const double md = DBL_MIN * 32;
    return (fabs(fa - fb) < md) || (fabs(fa - fd) < md) || (fabs(fa - fe) < md) || (fabs(fb - fd) < md) ||
           (fabs(fb - fe) < md) || (fabs(fd - fe) < md);
  }

  // ------------------------------------------------------------ rarefaction
  // Cash-Karp step (:1773-1896); outputs are written before the last check
  __device__ bool rk45(const DevEos& E, int wave, double rho_0, double u_0, double p_0, double dp, double& rho,
                       double& u, double& p, double& xi, double& uErr, double& rhoErr) {
    dp = -dp;
    const EC2 q_0 = ool_ec2(&E, rho_0, p_0);
    double c_0_square = q_0.c2;
    if (rho_0 <= 0 || c_0_square < 0) return false;
    double c_0 = sqrt(c_0_square);
    double rho_1 = rho_0 + 0.2 * dp / c_0_square;
    double p_1 = p_0 + 0.2 * dp;
    const EC2 q_1 = ool_ec2(&E, rho_1, p_1);
    double c_1_square = q_1.c2;
    if (rho_1 <= 0 || c_1_square < 0) return false;
    double rho_2 = rho_0 + 3. / 40. * dp / c_0_square + 9. / 40. * dp / c_1_square;
    double p_2 = p_1 + 3. / 10. * dp;  // the reference's abscissa (p0 + dp/2), :1814
    const EC2 q_2 = ool_ec2(&E, rho_2, p_2);
    double c_2_square = q_2.c2;
    if (rho_2 <= 0 || c_2_square < 0) return false;
    double c_2 = sqrt(c_2_square);
    double rho_3 = rho_0 + 3. / 10. * dp / c_0_square - 9. / 10. * dp / c_1_square + 6. / 5. * dp / c_2_square;
    double p_3 = p_0 + 3. / 5. * dp;
    const EC2 q_3 = ool_ec2(&E, rho_3, p_3);
    double c_3_square = q_3.c2;
    if (rho_3 <= 0 || c_3_square < 0) return false;
    double c_3 = sqrt(c_3_square);
    double rho_4 = rho_0 - 11. / 54. * dp / c_0_square + 2.5 * dp / c_1_square - 70. / 27. * dp / c_2_square +
                   35. / 27. * dp / c_3_square;
    double p_4 = p_0 + dp;
    const EC2 q_4 = ool_ec2(&E, rho_4, p_4);
    double c_4_square = q_4.c2;
    if (rho_4 <= 0 || c_4_square < 0) return false;
    double c_4 = sqrt(c_4_square);
    double rho_5 = rho_0 + 1631. / 55296. * dp / c_0_square + 175. / 512. * dp / c_1_square +
                   575. / 13824. * dp / c_2_square + 44275. / 110592. * dp / c_3_square +
                   253. / 4096. * dp / c_4_square;
    double p_5 = p_0 + 7. / 8. * dp;
    const EC2 q_5 = ool_ec2(&E, rho_5, p_5);
    double c_5_square = q_5.c2;
    if (rho_5 <= 0 || c_5_square < 0) return false;
    double c_5 = sqrt(c_5_square);
    double drho = (37. / 378. / c_0_square + 250. / 621. / c_2_square + 125. / 594. / c_3_square +
                   512. / 1771. / c_5_square) * dp;
    double drho_err = (2825. / 27648. / c_0_square + 18575. / 48384. / c_2_square + 13525. / 55296. / c_3_square +
                       277. / 14336. / c_4_square + 0.25 / c_5_square) * dp;
    double du = (37. / 378. / c_0 / rho_0 + 250. / 621. / c_2 / rho_2 + 125. / 594. / c_3 / rho_3 +
                 512. / 1771. / c_5 / rho_5) * dp;
    double du_err = (2825. / 27648. / c_0 / rho_0 + 18575. / 48384. / c_2 / rho_2 +
                     13525. / 55296. / c_3 / rho_3 + 277. / 14336. / c_4 / rho_4 + 0.25 / c_5 / rho_5) * dp;
    if (isnan(du) || isnan(du_err)) return false;
    rhoErr = fabs(drho_err - drho);
    uErr = fabs(du_err - du);
    rho = rho_0 + drho;
    p = p_0 + dp;
    u = (wave == 1) ? u_0 - du : u_0 + du;
    const EC2 qe = ool_ec2(&E, rho, p);
    double c = qe.c2;
    if (rho <= 0 || c < 0) return false;
    c = sqrt(c);
    xi = (wave == 1) ? u - c : u + c;
    return true;
  }

  // classical RK4 of the non-adaptive solver (:2897-2986)
  __device__ bool rk4(const DevEos& E, int wave, double rho_0, double u_0, double p_0, double dp, double& rho,
                      double& u, double& p, double& xi) {
    dp = -dp;
    const EC2 q_0 = ool_ec2(&E, rho_0, p_0);
    double c_0_square = q_0.c2;
    if (rho_0 <= 0 || c_0_square < 0) return false;
    double c_0 = sqrt(c_0_square);
    double rho_1 = rho_0 + 0.5 * dp / c_0_square;
    double p_1 = p_0 + 0.5 * dp;
    const EC2 q_1 = ool_ec2(&E, rho_1, p_1);
    double c_1_square = q_1.c2;
    if (rho_1 <= 0 || c_1_square < 0) return false;
    double c_1 = sqrt(c_1_square);
    double rho_2 = rho_0 + 0.5 * dp / c_1_square;
    double p_2 = p_1;
    const EC2 q_2 = ool_ec2(&E, rho_2, p_2);
    double c_2_square = q_2.c2;
    if (rho_2 <= 0 || c_2_square < 0) return false;
    double c_2 = sqrt(c_2_square);
    double rho_3 = rho_0 + dp / c_2_square;
    double p_3 = p_0 + dp;
    const EC2 q_3 = ool_ec2(&E, rho_3, p_3);
    double c_3_square = q_3.c2;
    if (rho_3 <= 0 || c_3_square < 0) return false;
    double c_3 = sqrt(c_3_square);
    double drho = 1.0 / 6.0 * dp * (1. / c_0_square + 2.0 * (1. / c_1_square + 1. / c_2_square) + 1. / c_3_square);
    double du = 1.0 / 6.0 * dp * (1. / c_0 / rho_0 + 2.0 * (1. / c_1 / rho_1 + 1. / c_2 / rho_2) + 1. / c_3 / rho_3);
    if (isnan(du)) return false;
    rho = rho_0 + drho;
    p = p_0 + dp;
    u = (wave == 1) ? u_0 - du : u_0 + du;
    const EC2 qe = ool_ec2(&E, rho, p);
    double c = qe.c2;
    if (rho <= 0 || c < 0) return false;
    c = sqrt(c);
    xi = (wave == 1) ? u - c : u + c;
    return true;
  }

  // ComputeRhoUStar (:1213-1420 adaptive, :2523-2690 non-adaptive)
  __device__ __noinline__ bool rhoustar(int wave, Path& path, double rho, double u, double p, double ps, int id, double rhos0,
                           double rhos1, double& rhos, double& us, bool* trans_rare = nullptr,
                           double* Vrare = nullptr, bool count = true) {
    if (count && wave == 1) st.fevals++;
    rhos = rho;
    us = u;
    // the material's parameters copied into registers once per call: read
    // through the table, every EOS evaluation of the serial chain below
    // waited on its own memory round trip (stores to the integration path
    // may alias the table for all the compiler knows)
    const DevEos E = M.m[id];
    if (!(p > ps)) return shock(E, rho, u, p, ps, rhos0, rhos1, wave, rhos, us);
    const int nsteps = P.numSteps;
    double rhos_0 = rho, us_0 = u, ps_0 = p, xi_0;
    double rhos_1 = rho, us_1 = u, ps_1 = p, xi_1 = 0.0;
    double dp, dp_min_adaption = 0.0, dp_target = 0.0;
    const EC2 q0 = ool_ec2(&E, rho, p);
    const double e = q0.e;
    if (!nonadaptive) {
      dp = (p - ps) / nsteps;
      dp_min_adaption = dp / 2.5;
    } else {
      dp_target = (p - ps) / nsteps;
      double dpdrho = eos_dpdrho(E, rho, e);
      double drho_op1 = (p - ps) / dpdrho / nsteps;
      double drho_op2 = rho / (2.5 * nsteps);
      double drho_target = fmin_std(drho_op1, drho_op2);
      dp = fmin_std(dp_target, drho_target * dpdrho);
    }
    const double pressure_endpoint_tol = P.tol_rare * fmax_std(1.0, fabs(p));
    double c = q0.c2;
    if (rho <= 0 || c < 0) return false;
    c = sqrt(c);
    if (path.n > 1) {
      // The reference scans the cached path backwards for the last point
      // with pressure above ps (:1243-1251).  Points are only appended below
      // the last one (:1355), so pressures strictly decrease along the path and the
      // points above ps form a prefix: a binary search over the prefix gives
      // the same index with log2(n) dependent loads instead of up to n.
      // Successive solves of one problem ask for nearby pressures, so the
      // index the last search returned is checked first: it is the answer
      // when its point lies above ps and the next one does not.  The three
      // loads are independent (one memory round trip instead of log2(n)).
      int index0 = 0;
      const int hn = path.n;
      const int h = path.hint < hn ? path.hint : hn - 1;
      const int h1 = (h + 1 < hn) ? h + 1 : h;
      const double p_first = path.get(0, 0), p_h = path.get(h, 0), p_h1 = path.get(h1, 0);
      if (p_first > ps) {
        if (p_h > ps && (h1 == h || !(p_h1 > ps))) {
          index0 = h;
        } else {
          int lo = 0, hi = hn - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (path.get(mid, 0) > ps) lo = mid;
            else hi = mid - 1;
          }
          index0 = lo;
        }
        path.hint = index0;
      }
      ps_0 = path.get(index0, 0);
      rhos_0 = path.get(index0, 1);
      us_0 = path.get(index0, 2);
      if (!nonadaptive) {
        dp_min_adaption = (ps_0 - ps) / nsteps / 2.5;
        if (index0 != path.n - 1) dp = ps_0 - ps;
        else dp = fmin_std(path.get(index0 - 1, 0) - path.get(index0, 0), ps_0 - ps);
      } else {
        dp = fmin_std((p - ps) / nsteps, ps_0 - ps);
      }
    }
    double xi = (wave == 1) ? u - c : u + c;
    xi_0 = xi;
    bool done = false;
    double uErr = 0., rhoErr = 0.;
    const int maxsteps = nsteps * (nonadaptive ? 5 : 3);
    for (int i = 0; i < maxsteps; i++) {
      if (fabs(ps_1 - ps) <= pressure_endpoint_tol) {
        rhos = rhos_1;
        us = us_1;
        if (eos_check(E, rhos, ps)) return false;
        done = true;
        break;
      }
      bool ok = nonadaptive ? rk4(E, wave, rhos_0, us_0, ps_0, dp, rhos_1, us_1, ps_1, xi_1)
                            : rk45(E, wave, rhos_0, us_0, ps_0, dp, rhos_1, us_1, ps_1, xi_1, uErr, rhoErr);
      if (!ok) {
        dp = dp / 2.0;
        if (nonadaptive && dp == 0.0) break;
        continue;
      }
      if (ps_1 - ps < -pressure_endpoint_tol) { dp = ps_0 - ps; continue; }
      double rhoErrScaled = 0, uErrScaled = 0, errBar = 0;
      const double safety = 0.9;
      if (!nonadaptive) {
        const double tiny = 1.e-30;
        errBar = fmax_std(1.0e-8, fmin_std(1.0e-5, 100.0 * P.tol_rare));
        uErrScaled = uErr / c + tiny;
        rhoErrScaled = rhoErr / rho + tiny;
        if (isnan(rhoErrScaled) || isnan(uErrScaled)) { st.fatal = true; return false; }
        double dpTemp;
        if (rhoErrScaled >= errBar && dp > dp_min_adaption) {
          dpTemp = safety * dp * cr_pow_quarter(fabs(errBar / rhoErrScaled));
          dpTemp = fmax_std(fmax_std(dpTemp, 0.2 * dp), dp_min_adaption);
          dp = fmin_std(dpTemp, ps_0 - ps);
          continue;
        }
        if (uErrScaled >= errBar && dp > dp_min_adaption) {
          dpTemp = safety * dp * cr_pow_quarter(fabs(errBar / uErrScaled));
          dpTemp = fmax_std(fmax_std(dpTemp, 0.2 * dp), dp_min_adaption);
          dp = fmin_std(dpTemp, ps_0 - ps);
          continue;
        }
      }
      if (ps_1 < path.get(path.n - 1, 0)) path.push(ps_1, rhos_1, us_1);
      if (trans_rare && Vrare && xi_0 * xi_1 <= 0) {
        *trans_rare = true;
        double w0 = fabs(xi_1), w1 = fabs(xi_0);
        double ww = w0 + w1;
        w0 /= ww;
        w1 /= ww;
        Vrare[0] = w0 * rhos_0 + w1 * rhos_1;
        Vrare[1] = w0 * us_0 + w1 * us_1;
        Vrare[2] = w0 * ps_0 + w1 * ps_1;
      }
      if (!nonadaptive) {
        double dpTemp_rho = safety * dp * cr_pow_fifth(fabs(errBar / rhoErrScaled));
        double dpTemp_u = safety * dp * cr_pow_fifth(fabs(errBar / uErrScaled));
        double dpTemp = fmin_std(dpTemp_rho, dpTemp_u);
        dpTemp = fmin_std(dpTemp, 10 * dp);
        dpTemp = fmax_std(dpTemp, dp_min_adaption);
        dp = fmin_std(dpTemp, ps_1 - ps);
      } else {
        dp = fmin_std(fmin_std(dp_target, ps_1 - ps), 4.0 * dp);
      }
      rhos_0 = rhos_1; us_0 = us_1; ps_0 = ps_1; xi_0 = xi_1;
    }
    return done;
  }

  // ----------------------------------------------------------- bracketing
  // wave strengths between which the learned p* guess replaces the acoustic one
  static constexpr double LEARNED_MIN_STRENGTH = 0.3, LEARNED_MAX_STRENGTH = 10.0;

  // FindInitialFeasiblePointsByAcousticTheory (:1107-1158); `p0g` replaces the
  // acoustic p0 when positive (learned guess), see m2cx_scheme.riem_guess.
  __device__ int acoustic(double rhol, double ul, double pl, double cl, int idl, double rhor, double ur, double pr,
                          double cr, int idr, double& p0, double& rhol0, double& rhor0, double& ul0, double& ur0,
                          double& p1, double& rhol1, double& rhor1, double& ul1, double& ur1, double p0g) {
    int found = 0;
    double Cl = rhol * cl, Cr = rhor * cr;
    p0 = (Cr * pl + Cl * pr + Cl * Cr * (ul - ur)) / (Cl + Cr);
    // learned guess: acoustic theory is exact to second order in the wave
    // strength s = |p0 - p| / (rho c^2) and beats the network where both
    // waves are weak; the prediction replaces it where the acoustic pressure
    // is not positive and for strong waves up to the strength the network
    // was trained on
    if (p0g > 0) {
      const double sl = fabs(p0 - pl) / (Cl * cl), sr = fabs(p0 - pr) / (Cr * cr);
      const double sw = fmax_std(sl, sr);
      if (!(p0 > 0) || (sw >= LEARNED_MIN_STRENGTH && sw < LEARNED_MAX_STRENGTH)) p0 = p0g;
    }
    if (!rhoustar(1, path1, rhol, ul, pl, p0, idl, rhol, (p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0, ul0)) return found;
    if (!rhoustar(3, path3, rhor, ur, pr, p0, idr, rhor, (p0 > pr) ? rhor * 1.1 : rhor * 0.9, rhor0, ur0)) return found;
    found = 1;
    double Clbar = (ul0 == ul) ? Cl : fabs(p0 - pl) / fabs(ul0 - ul);
    double Crbar = (ur0 == ur) ? Cr : fabs(p0 - pr) / fabs(ur0 - ur);
    p1 = (Crbar * pl + Clbar * pr + Clbar * Crbar * (ul - ur)) / (Clbar + Crbar);
    double tmp = fmax_std(fabs(p0), fabs(p1));
    if (fabs(p1 - p0) / tmp < 1.0e-8) p1 = p0 + 1.0e-8 * tmp;
    if (!rhoustar(1, path1, rhol, ul, pl, p1, idl, rhol, rhol0, rhol1, ul1)) return found;
    if (!rhoustar(3, path3, rhor, ur, pr, p1, idr, rhor, rhor0, rhor1, ur1)) return found;
    return 2;
  }

  // FindInitialFeasiblePoints (:1003-1103)
  __device__ bool feasible(double rhol, double ul, double pl, double cl, int idl, double rhor, double ur, double pr,
                           double cr, int idr, double& p0, double& rhol0, double& rhor0, double& ul0, double& ur0,
                           double& p1, double& rhol1, double& rhor1, double& ul1, double& ur1, double p0g) {
    bool success = true;
    int found = acoustic(rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, p0, rhol0, rhor0, ul0, ur0, p1, rhol1, rhor1,
                         ul1, ur1, p0g);
    if (found == 2) return true;
    double dp;
    if (found != 1) {
      dp = (pl != pr) ? fabs(pl - pr) : 0.5 * pl;
      const double lo = fmin_std(pl, pr), sc = fmin_std(dp, fmin_std(fabs(pl), fabs(pr)));
      for (int i = 0; i < P.maxIts_bracket; i++) {
        p0 = lo + 0.01 * (i + 1) * (i + 1) * sc;
        if (p0 < P.min_pressure) p0 = P.pressure_at_failure;
        success = rhoustar(1, path1, rhol, ul, pl, p0, idl, rhol, (p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0, ul0);
        success = success && rhoustar(3, path3, rhor, ur, pr, p0, idr, rhor, (p0 > pr) ? rhor * 1.1 : rhor * 0.9, rhor0, ur0);
        if (success) break;
      }
      if (!success) {
        for (int i = 0; i < P.maxIts_bracket; i++) {
          p0 = lo - 0.01 * (i + 1) * (i + 1) * sc;
          if (p0 < P.min_pressure || i == (int)(P.maxIts_bracket / 2)) p0 = P.pressure_at_failure;
          success = rhoustar(1, path1, rhol, ul, pl, p0, idl, rhol, (p0 > pl) ? rhol * 1.1 : rhol * 0.9, rhol0, ul0);
          success = success && rhoustar(3, path3, rhor, ur, pr, p0, idr, rhor, (p0 > pr) ? rhor * 1.1 : rhor * 0.9, rhor0, ur0);
          if (success) break;
        }
      }
      if (!success) return false;
    }
    dp = fmin_std(fabs(p0 - pl), fabs(p0 - pr));
    for (int i = 0; i < P.maxIts_bracket; i++) {
      p1 = p0 + 0.01 * (i + 1) * (i + 1) * dp;
      success = rhoustar(1, path1, rhol, ul, pl, p1, idl, rhol, rhol0, rhol1, ul1);
      success = success && rhoustar(3, path3, rhor, ur, pr, p1, idr, rhor, rhor0, rhor1, ur1);
      if (success) break;
    }
    if (!success)
      for (int i = 0; i < P.maxIts_bracket; i++) {
        p1 = p0 - 0.01 * (i + 1) * (i + 1) * dp;
        if (p1 < P.min_pressure) p1 = P.pressure_at_failure * 1000.0;
        success = rhoustar(1, path1, rhol, ul, pl, p1, idl, rhol, rhol0, rhol1, ul1);
        success = success && rhoustar(3, path3, rhor, ur, pr, p1, idr, rhor, rhor0, rhor1, ur1);
        if (success) break;
      }
    if (!success) return false;
    if (p0 > p1) {
      double t;
      t = p0; p0 = p1; p1 = t;
      t = rhol0; rhol0 = rhol1; rhol1 = t;
      t = rhor0; rhor0 = rhor1; rhor1 = t;
      t = ul0; ul0 = ul1; ul1 = t;
      t = ur0; ur0 = ur1; ur1 = t;
    }
    return true;
  }

  // FindInitialInterval (:675-839)
  __device__ bool interval(double rhol, double ul, double pl, double cl, int idl, double rhor, double ur, double pr,
                           double cr, int idr, double& p0, double& rhol0, double& rhor0, double& ul0, double& ur0,
                           double& p1, double& rhol1, double& rhor1, double& ul1, double& ur1, double p0g) {
    bool success = feasible(rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, p0, rhol0, rhor0, ul0, ur0, p1, rhol1,
                            rhor1, ul1, ur1, p0g);
    if (!success) {
      p0 = p1 = P.pressure_at_failure;
      return false;
    }
    double p2 = 0, rhol2 = 0, rhor2 = 0, ul2 = 0, ur2 = 0, f0, f1;
    int i;
    double fmin, p_fmin, rhol_fmin, rhor_fmin, ul_fmin, ur_fmin;
    if (fabs(ul0 - ur0) < fabs(ul1 - ur1)) {
      fmin = fabs(ul0 - ur0); p_fmin = p0; rhol_fmin = rhol0; rhor_fmin = rhor0; ul_fmin = ul0; ur_fmin = ur0;
    } else {
      fmin = fabs(ul1 - ur1); p_fmin = p1; rhol_fmin = rhol1; rhor_fmin = rhor1; ul_fmin = ul1; ur_fmin = ur1;
    }
    for (i = 0; i < P.maxIts_bracket; i++) {
      f0 = ul0 - ur0;
      f1 = ul1 - ur1;
      if (f0 * f1 <= 0.0) return true;
      if (fabs(f0 - f1) > 1e-9) {
        p2 = p1 - f1 * (p1 - p0) / (f1 - f0);
        if (p2 < p0) p2 -= 0.1 * (p1 - p0);
        else p2 += 0.1 * (p1 - p0);
      } else {
        p2 = 1.1 * p1;
      }
      if (p2 < P.min_pressure || i == int(P.maxIts_bracket / 2)) p2 = 1.0e-8;
      success = rhoustar(1, path1, rhol, ul, pl, p2, idl, rhol0, rhol1, rhol2, ul2);
      success = success && rhoustar(3, path3, rhor, ur, pr, p2, idr, rhor0, rhor1, rhor2, ur2);
      if (!success) {
        for (int j = 0; j < P.maxIts_bracket / 2; j++) {
          if (p2 < p0) p2 = p0 - 0.5 * (p0 - p2);
          else p2 = p1 + 0.5 * (p2 - p1);
          success = rhoustar(1, path1, rhol, ul, pl, p2, idl, rhol0, rhol1, rhol2, ul2);
          success = success && rhoustar(3, path3, rhor, ur, pr, p2, idr, rhor0, rhor1, rhor2, ur2);
          if (success) break;
        }
      }
      if (!success) break;
      if (fabs(ul2 - ur2) < fmin) {
        fmin = fabs(ul2 - ur2); p_fmin = p2; rhol_fmin = rhol2; rhor_fmin = rhor2; ul_fmin = ul2; ur_fmin = ur2;
      }
      if (p2 < p0) {
        p1 = p0; rhol1 = rhol0; rhor1 = rhor0; ul1 = ul0; ur1 = ur0;
        p0 = p2; rhol0 = rhol2; rhor0 = rhor2; ul0 = ul2; ur0 = ur2;
      } else {
        p0 = p1; rhol0 = rhol1; rhor0 = rhor1; ul0 = ul1; ur0 = ur1;
        p1 = p2; rhol1 = rhol2; rhor1 = rhor2; ul1 = ul2; ur1 = ur2;
      }
    }
    if (!success || i == P.maxIts_bracket) {
      if (fmin < P.failure_threshold * fabs(ul - ur)) {
        p0 = p1 = p_fmin; rhol0 = rhol1 = rhol_fmin; rhor0 = rhor1 = rhor_fmin; ul0 = ul1 = ul_fmin;
        ur0 = ur1 = ur_fmin;
      } else {
        p2 = P.pressure_at_failure;
        success = rhoustar(1, path1, rhol, ul, pl, p2, idl, rhol0, rhol1, rhol2, ul2);
        success = success && rhoustar(3, path3, rhor, ur, pr, p2, idr, rhor0, rhor1, rhor2, ur2);
        if (success) {
          p0 = p1 = p2; rhol0 = rhol1 = rhol2; rhor0 = rhor1 = rhor2; ul0 = ul1 = ul2; ur0 = ur1 = ur2;
        } else {
          p0 = p1 = p_fmin; rhol0 = rhol1 = rhol_fmin; rhor0 = rhor1 = rhor_fmin; ul0 = ul1 = ul_fmin;
          ur0 = ur1 = ur_fmin;
        }
      }
      return false;
    }
    return true;
  }

  // FinalizeSolution (:352-558), normal build
  __device__ void finalize(const double* dir, const double* Vm, const double* Vp, double rhol, double ul, double pl,
                           int idl, double rhor, double ur, double pr, int idr, double rhol2, double rhor2,
                           double u2, double p2, bool trans_rare, const double* Vrare, double* Vs, int& id,
                           double* Vsm, double* Vsp) {
    double utanl[3] = {Vm[1] - ul * dir[0], Vm[2] - ul * dir[1], Vm[3] - ul * dir[2]};
    double utanr[3] = {Vp[1] - ur * dir[0], Vp[2] - ur * dir[1], Vp[3] - ur * dir[2]};
    id = (u2 >= 0) ? idl : idr;
    for (int i = 0; i < 5; i++) Vs[i] = 0.0;
    if (trans_rare) {
      Vs[0] = Vrare[0];
      Vs[1] = Vrare[1] * dir[0];
      Vs[2] = Vrare[1] * dir[1];
      Vs[3] = Vrare[1] * dir[2];
      Vs[4] = Vrare[2];
    } else if (u2 >= 0) {
      bool star = false;
      if (pl >= p2) {
        const DevEos& E = M.m[idl];
        double cl2 = ool_ec2(&E, rhol2, p2).c2;
        if (rhol2 <= 0 || cl2 < 0) st.fatal = true;
        cl2 = sqrt(cl2);
        if (u2 - cl2 <= 0) star = true;
      } else {
        double us = (rhol2 * u2 - rhol * ul) / (rhol2 - rhol);
        if (us <= 0) star = true;
      }
      if (star) { Vs[0] = rhol2; Vs[1] = u2 * dir[0]; Vs[2] = u2 * dir[1]; Vs[3] = u2 * dir[2]; Vs[4] = p2; }
      else { Vs[0] = rhol; Vs[1] = ul * dir[0]; Vs[2] = ul * dir[1]; Vs[3] = ul * dir[2]; Vs[4] = pl; }
    } else {
      bool star = false;
      if (pr >= p2) {
        const DevEos& E = M.m[idr];
        double cr2 = ool_ec2(&E, rhor2, p2).c2;
        if (rhor2 <= 0 || cr2 < 0) st.fatal = true;
        cr2 = sqrt(cr2);
        if (u2 - cr2 >= 0) star = true;
      } else {
        double us = (rhor2 * u2 - rhor * ur) / (rhor2 - rhor);
        if (us >= 0) star = true;
      }
      if (star) { Vs[0] = rhor2; Vs[1] = u2 * dir[0]; Vs[2] = u2 * dir[1]; Vs[3] = u2 * dir[2]; Vs[4] = p2; }
      else { Vs[0] = rhor; Vs[1] = ur * dir[0]; Vs[2] = ur * dir[1]; Vs[3] = ur * dir[2]; Vs[4] = pr; }
    }
    if (u2 > 0) {
      for (int i = 1; i <= 3; i++) Vs[i] += utanl[i - 1];
    } else if (u2 < 0) {
      for (int i = 1; i <= 3; i++) Vs[i] += utanr[i - 1];
    } else {
      for (int i = 1; i <= 3; i++) Vs[i] += 0.5 * (utanl[i - 1] + utanr[i - 1]);
    }
    Vsm[0] = rhol2;
    Vsm[1] = utanl[0] + u2 * dir[0];
    Vsm[2] = utanl[1] + u2 * dir[1];
    Vsm[3] = utanl[2] + u2 * dir[2];
    Vsm[4] = p2;
    Vsp[0] = rhor2;
    Vsp[1] = utanr[0] + u2 * dir[0];
    Vsp[2] = utanr[1] + u2 * dir[1];
    Vsp[3] = utanr[2] + u2 * dir[2];
    Vsp[4] = p2;
  }

  // One pass of ComputeRiemannSolution.  Returns 0/1 as the reference's error
  // code; `again` asks for the non-adaptive pass.
  __device__ int solve_pass(const double* dir, const double* Vm, int idl, const double* Vp, int idr, double* Vs,
                            int& id, double* Vsm, double* Vsp, double p0g, bool& again) {
    again = false;
    double rhol = Vm[0];
    double ul = Vm[1] * dir[0] + Vm[2] * dir[1] + Vm[3] * dir[2];
    double pl = Vm[4];
    double rhor = Vp[0];
    double ur = Vp[1] * dir[0] + Vp[2] * dir[1] + Vp[3] * dir[2];
    double pr = Vp[4];
    path1.reset(pl, rhol, ul);
    path3.reset(pr, rhor, ur);
    double cl = ool_ec2(&M.m[idl], rhol, pl).c2;
    if (rhol <= 0 || cl < 0) { st.fatal = true; return 1; }
    cl = sqrt(cl);
    double cr = ool_ec2(&M.m[idr], rhor, pr).c2;
    if (rhor <= 0 || cr < 0) { st.fatal = true; return 1; }
    cr = sqrt(cr);
    double p0(DBL_MIN), ul0(0.0), ur0(0.0), rhol0(DBL_MIN), rhor0(DBL_MIN);
    double pp1(DBL_MIN), ul1(0.0), ur1(0.0), rhol1(DBL_MIN), rhor1(DBL_MIN);
    double p2(DBL_MIN), ul2(0.0), ur2(0.0), rhol2(DBL_MIN), rhor2(DBL_MIN);
    double f0(0.0), f1(0.0), f2(0.0);
    bool trans_rare = false;
    double Vrare[3] = {0, 0, 0};
    if (ul == ur && pl == pr) {
      st.trivial = true;
      finalize(dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol, rhor, ul, pl, trans_rare, Vrare, Vs, id, Vsm,
               Vsp);
      return 0;
    }
    bool success = interval(rhol, ul, pl, cl, idl, rhor, ur, pr, cr, idr, p0, rhol0, rhor0, ul0, ur0, pp1, rhol1,
                            rhor1, ul1, ur1, nonadaptive ? -1.0 : p0g);
    if (!success) {
      success = rhoustar(1, path1, rhol, ul, pl, pp1, idl, rhol0, rhol0 * 1.1, rhol2, ul2, &trans_rare, Vrare);
      success = success && rhoustar(3, path3, rhor, ur, pr, pp1, idr, rhor0, rhol0 * 1.1, rhor2, ur2, &trans_rare, Vrare);
      if (!success) {
        for (int i = 0; i < 5; i++) { Vsm[i] = Vm[i]; Vsp[i] = Vp[i]; }
        double u_avg = 0.5 * (ul + ur);
        if (u_avg >= 0) { id = idl; for (int i = 0; i < 5; i++) Vs[i] = Vm[i]; }
        else { id = idr; for (int i = 0; i < 5; i++) Vs[i] = Vp[i]; }
        again = !nonadaptive;
        return 1;
      }
      finalize(dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol2, rhor2, 0.5 * (ul2 + ur2), pp1, trans_rare,
               Vrare, Vs, id, Vsm, Vsp);
      again = !nonadaptive;
      return 1;
    }
    f0 = ul0 - ur0;
    f1 = ul1 - ur1;
    double denom = 0;
    int iter = 0;
    double err_p = 1.0, err_u = 1.0;
    p2 = pp1;
    f2 = f1;
    for (iter = 0; iter < P.maxIts_main; iter++) {
      st.it_main++;
      denom = f1 - f0;
      if (denom == 0) {
        p2 = p0 + 0.5 * (pp1 - p0);
      } else {
        p2 = p2 - f2 * (pp1 - p0) / denom;
        if (p2 <= p0 || p2 >= pp1) p2 = p0 + 0.5 * (pp1 - p0);
      }
      bool ok;
      while (true) {
        ok = rhoustar(1, path1, rhol, ul, pl, p2, idl, rhol0, rhol1, rhol2, ul2, &trans_rare, Vrare);
        if (!ok) {
          p2 = 0.5 * (p2 + p0);
          iter++;
          if (iter < P.maxIts_main) { st.it_main++; continue; }
          break;
        }
        ok = rhoustar(3, path3, rhor, ur, pr, p2, idr, rhor0, rhor1, rhor2, ur2, &trans_rare, Vrare);
        if (!ok) {
          p2 = 0.5 * (p2 + pp1);
          iter++;
          if (iter < P.maxIts_main) { st.it_main++; continue; }
          break;
        }
        break;
      }
      if (!ok) break;
      f2 = ul2 - ur2;
      if (f0 * f2 < 0.0) { pp1 = p2; f1 = f2; rhol1 = rhol2; rhor1 = rhor2; }
      else { p0 = p2; f0 = f2; rhol0 = rhol2; rhor0 = rhor2; }
      err_p = fabs(pp1 - p0) / fmax_std(fabs(pl + 0.5 * rhol * ul * ul), fabs(pr + 0.5 * rhor * ur * ur));
      err_u = fabs(f2) / fmax_std(cl, cr);
      if ((err_p < P.tol_main && err_u < P.tol_main) || (err_p < P.tol_main * 1e-3) || (err_u < P.tol_main * 1e-3))
        break;
      trans_rare = false;
    }
    double u2 = 0.5 * (ul2 + ur2);
    finalize(dir, Vm, Vp, rhol, ul, pl, idl, rhor, ur, pr, idr, rhol2, rhor2, u2, p2, trans_rare, Vrare, Vs, id, Vsm,
             Vsp);
    if (iter == P.maxIts_main) {
      again = !nonadaptive;
      return 1;
    }
    return 0;
  }

  __device__ __noinline__ int solve(const double* dir, const double* Vm, int idl, const double* Vp, int idr, double* Vs, int& id,
                       double* Vsm, double* Vsp, double p0g = -1.0) {
    nonadaptive = false;
    bool again;
    int r = solve_pass(dir, Vm, idl, Vp, idr, Vs, id, Vsm, Vsp, p0g, again);
    if (again) {
      st.fallback = true;
      nonadaptive = true;
      r = solve_pass(dir, Vm, idl, Vp, idr, Vs, id, Vsm, Vsp, p0g, again);
      nonadaptive = false;
    }
    return r;
  }
};

}  // namespace m2c
