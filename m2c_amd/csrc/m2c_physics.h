// Device-side point physics of the explicit step: equations of state, numerical
// fluxes and slope limiters, for gfx950.
//
// Every expression keeps the reference's operation order (the library is
// compiled with -ffp-contract=off), so results match the reference built
// without FMA contraction to the last bit:
//   EOS            VarFcnBase.h:120-298, VarFcnSG.h:64-108, VarFcnNASG.h:55-86,
//                  VarFcnMG.h:70-99, VarFcnDummy.h:18-58
//   fluxes         FluxFcnBase.h:83-123,287-303, FluxFcnHLLC.h:58-245,
//                  FluxFcnGenRoe.h:64-221, FluxFcnLLF.h:47-185
//   limiters       Reconstructor.h:96-111
//   characteristic FluxFcnBase.h:309-517
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include "../../include/m2c_amd.h"

namespace m2c {

#define M2C_MAX_MAT 8

struct DevEos {
  int type;
  int pad;
  double gam, pc, gam1, invgam1, gam_pc, b, q;
  double rho0, c0, Gamma0, s, e0, rho0_c0_c0, Gamma0_over_2, Gamma0_rho0;
  double rhomin, pmin, rhomax, pmax, failsafe;
  double d_rho, d_p, d_e;
};

struct DevMats {
  DevEos m[M2C_MAX_MAT + 1];  // index nmat is the inactive (dummy) material
  int nmat;
};

__host__ inline DevEos make_eos(const m2cx_material& mm) {
  DevEos e{};
  e.type = mm.eos;
  e.rhomin = mm.rhomin; e.pmin = mm.pmin; e.rhomax = mm.rhomax; e.pmax = mm.pmax;
  e.failsafe = mm.failsafe_density;
  e.gam = mm.gamma; e.pc = mm.pc; e.gam1 = mm.gamma - 1.0; e.invgam1 = 1.0 / e.gam1;
  e.gam_pc = mm.gamma * mm.pc; e.b = mm.b; e.q = mm.q;
  e.rho0 = mm.rho0; e.c0 = mm.c0; e.Gamma0 = mm.Gamma0; e.s = mm.s; e.e0 = mm.e0;
  e.rho0_c0_c0 = mm.rho0 * mm.c0 * mm.c0;
  e.Gamma0_over_2 = 0.5 * mm.Gamma0;
  e.Gamma0_rho0 = mm.Gamma0 * mm.rho0;
  e.d_rho = mm.dummy_rho; e.d_p = mm.dummy_p; e.d_e = mm.dummy_e;
  return e;
}

// ---------------------------------------------------------------- EOS
__device__ __forceinline__ double eos_p(const DevEos& E, double rho, double e) {
  if (E.type == M2CX_EOS_SG) return E.gam1 * rho * e - E.gam * E.pc;
  if (E.type == M2CX_EOS_NASG) return E.gam1 * (e - E.q) / (1.0 / rho - E.b) - E.gam_pc;
  if (E.type == M2CX_EOS_MG) {
    double eta = 1.0 - E.rho0 / rho;
    return E.rho0_c0_c0 * eta * (1.0 - E.Gamma0_over_2 * eta) / ((1.0 - E.s * eta) * (1.0 - E.s * eta)) +
           E.Gamma0_rho0 * (e - E.e0);
  }
  return E.d_p;
}
__device__ __forceinline__ double eos_e(const DevEos& E, double rho, double p) {
  if (E.type == M2CX_EOS_SG) return (p + E.gam * E.pc) / (E.gam1 * rho);
  if (E.type == M2CX_EOS_NASG) return E.invgam1 * (p + E.gam_pc) * (1.0 / rho - E.b) + E.q;
  if (E.type == M2CX_EOS_MG) {
    double eta = 1.0 - E.rho0 / rho;
    return (p - E.rho0_c0_c0 * eta * (1.0 - E.Gamma0_over_2 * eta) / ((1.0 - E.s * eta) * (1.0 - E.s * eta))) /
               E.Gamma0_rho0 + E.e0;
  }
  return E.d_e;
}
__device__ __forceinline__ double eos_dpdrho(const DevEos& E, double rho, double e) {
  if (E.type == M2CX_EOS_SG) return E.gam1 * e;
  if (E.type == M2CX_EOS_NASG) {
    double V = 1.0 / rho;
    return E.gam1 * V * V * (e - E.q) / ((V - E.b) * (V - E.b));
  }
  if (E.type == M2CX_EOS_MG) {
    double eta = 1.0 - E.rho0 / rho;
    double S = 1.0 - E.s * eta;
    return E.rho0_c0_c0 * (1.0 + (E.s - E.Gamma0) * eta) / (S * S * S) * E.rho0 / (rho * rho);
  }
  return 0.0;
}
__device__ __forceinline__ double eos_Gamma(const DevEos& E, double rho) {
  if (E.type == M2CX_EOS_SG) return E.gam1;
  if (E.type == M2CX_EOS_NASG) return E.gam1 / (1.0 - E.b * rho);
  if (E.type == M2CX_EOS_MG) return E.Gamma0_rho0 / rho;
  return 0.0;
}
__device__ __forceinline__ double eos_c2(const DevEos& E, double rho, double e) {
  if (E.type == M2CX_EOS_DUMMY) return DBL_MIN;
  return eos_dpdrho(E, rho, e) + eos_p(E, rho, e) / rho * eos_Gamma(E, rho);
}
__device__ __forceinline__ double eos_H(const DevEos& E, const double* V) {
  if (E.type == M2CX_EOS_DUMMY) return E.d_e;
  double e = eos_e(E, V[0], V[4]);
  return e + 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) + V[4] / V[0];
}
// true = state fails (not hyperbolic / non-finite)
__device__ __forceinline__ bool eos_check(const DevEos& E, double rho, double p) {
  if (E.type == M2CX_EOS_DUMMY) return false;
  if (!isfinite(rho) || !isfinite(p)) return true;
  if (E.type == M2CX_EOS_SG || E.type == M2CX_EOS_NASG) return rho <= 0.0 || p + E.pc <= 0.0;
  if (rho <= 0.0) return true;
  double e = eos_e(E, rho, p);
  double c2 = eos_dpdrho(E, rho, e) + p / rho * eos_Gamma(E, rho);
  return c2 <= 0;
}
__device__ __forceinline__ bool eos_check5(const DevEos& E, const double* V) {
  if (E.type == M2CX_EOS_DUMMY) return false;
  if (!isfinite(V[0]) || !isfinite(V[1]) || !isfinite(V[2]) || !isfinite(V[3]) || !isfinite(V[4])) return true;
  return eos_check(E, V[0], V[4]);
}
__device__ __forceinline__ void eos_c2p(const DevEos& E, const double* U, double* V) {
  if (E.type == M2CX_EOS_DUMMY) { V[0] = E.d_rho; V[1] = V[2] = V[3] = 0.0; V[4] = E.d_p; return; }
  V[0] = U[0];
  double invRho = 1.0 / U[0];
  V[1] = U[1] * invRho;
  V[2] = U[2] * invRho;
  V[3] = U[3] * invRho;
  double e = (U[4] - 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3])) * invRho;
  V[4] = eos_p(E, V[0], e);
}
__device__ __forceinline__ void eos_p2c(const DevEos& E, const double* V, double* U) {
  if (E.type == M2CX_EOS_DUMMY) { U[0] = E.d_rho; U[1] = U[2] = U[3] = 0.0; U[4] = E.d_rho * E.d_e; return; }
  U[0] = V[0];
  U[1] = V[0] * V[1];
  U[2] = V[0] * V[2];
  U[3] = V[0] * V[3];
  double e = eos_e(E, V[0], V[4]);
  U[4] = V[0] * (e + 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]));
}
// returns true if clipped (V modified); dummy resets V and reports false
__device__ __forceinline__ bool eos_clip(const DevEos& E, double* V) {
  if (E.type == M2CX_EOS_DUMMY) { V[0] = E.d_rho; V[1] = V[2] = V[3] = 0.0; V[4] = E.d_p; return false; }
  bool c = false;
  if (V[0] < E.rhomin) { V[0] = E.rhomin; c = true; }
  if (V[4] < E.pmin) { V[4] = E.pmin; c = true; }
  if (V[0] > E.rhomax) { V[0] = E.rhomax; c = true; }
  if (V[4] > E.pmax) { V[4] = E.pmax; c = true; }
  return c;
}
// would ClipDensityAndPressure modify V?  (no side effects)
__device__ __forceinline__ bool eos_would_clip(const DevEos& E, const double* V) {
  if (E.type == M2CX_EOS_DUMMY) return false;
  return V[0] < E.rhomin || V[4] < E.pmin || V[0] > E.rhomax || V[4] > E.pmax;
}

// ---------------------------------------------------------------- fluxes
// Velocity component normal to axis d.  The empty asm pins the three
// candidates in registers, so that the compiler cannot turn the select into
// an indexed load from a private array (which would live in scratch memory).
__device__ __forceinline__ double vsel(const double* V, int d) {
  double a = V[1], b = V[2], c = V[3];
  asm("" : "+v"(a), "+v"(b), "+v"(c));
  return d == 0 ? a : (d == 1 ? b : c);
}

__device__ __forceinline__ void flux_phys(const DevEos& E, int d, const double* V, double* F) {
  double H = eos_H(E, V);
  double m = V[0] * vsel(V, d);
  F[0] = m;
  F[1] = (d == 0) ? m * V[1] + V[4] : m * V[1];
  F[2] = (d == 1) ? m * V[2] + V[4] : m * V[2];
  F[3] = (d == 2) ? m * V[3] + V[4] : m * V[3];
  F[4] = m * H;
}

__device__ __forceinline__ double wavg(double w1, double f1, double w2, double f2) {
  return (w1 * f1 + w2 * f2) / (w1 + w2);
}

// Generalized Roe-averaged sound speed of Hu et al. (FluxFcnHLLC.h:67-128)
struct RoeHat {
  double rho_hat, srm, srp, drho, dp, e_hat, c_hat, c2_hat, dpdrho_hat, Gamma_hat;
};
__device__ __forceinline__ void roe_hat(const DevEos& E, const double* Vm, const double* Vp, double dvel,
                                        RoeHat& h) {
  const double eps = 1e-10;
  h.rho_hat = sqrt(Vm[0] * Vp[0]);
  h.srm = sqrt(Vm[0]);
  h.srp = sqrt(Vp[0]);
  h.drho = Vp[0] - Vm[0];
  h.dp = Vp[4] - Vm[4];
  double diff = dvel / (h.srm + h.srp);
  double pr = wavg(h.srm, Vm[4] / Vm[0], h.srp, Vp[4] / Vp[0]) + 0.5 * diff * diff;
  double em = eos_e(E, Vm[0], Vm[4]);
  double ep = eos_e(E, Vp[0], Vp[4]);
  double de = ep - em;
  h.e_hat = wavg(h.srm, em, h.srp, ep);
  double w_rho = h.drho * h.drho / (h.rho_hat * h.rho_hat);
  double w_e = de * de / (h.e_hat * h.e_hat);
  double den = w_rho + w_e + eps;
  double dpdrho_avg = wavg(h.srm, eos_dpdrho(E, Vm[0], em), h.srp, eos_dpdrho(E, Vp[0], ep));
  double Gamma_avg = wavg(h.srm, eos_Gamma(E, Vm[0]), h.srp, eos_Gamma(E, Vp[0]));
  double t1 = (w_e + eps) * dpdrho_avg;
  double t2 = (h.dp - Gamma_avg * h.rho_hat * de) * h.drho / (h.rho_hat * h.rho_hat);
  h.dpdrho_hat = (t1 + t2) / den;
  double g1 = (w_rho + eps) * Gamma_avg;
  double g2 = (h.dp - dpdrho_avg * h.drho) * de / (h.rho_hat * h.e_hat * h.e_hat);
  h.Gamma_hat = (g1 + g2) / den;
  h.c2_hat = h.dpdrho_hat + h.Gamma_hat * pr;
  if (h.c2_hat <= 0) {
    h.c_hat = eps;
    h.c2_hat = h.c_hat * h.c_hat;
  } else {
    h.c_hat = sqrt(h.c2_hat);
  }
}

__device__ __forceinline__ void hllc_star(const DevEos& E, int d, const double* V, double S, double Ss,
                                          double* F) {
  double un = vsel(V, d);
  double rs = V[0] * (S - un) / (S - Ss);
  double U[5];
  eos_p2c(E, V, U);
  double Us[5];
  Us[0] = 1.0;
  Us[1] = (d == 0) ? Ss : V[1];
  Us[2] = (d == 1) ? Ss : V[2];
  Us[3] = (d == 2) ? Ss : V[3];
  Us[4] = U[4] / V[0] + (Ss - un) * (Ss + V[4] / (V[0] * (S - un)));
#pragma unroll
  for (int q = 0; q < 5; q++) Us[q] *= rs;
  flux_phys(E, d, V, F);
#pragma unroll
  for (int q = 0; q < 5; q++) F[q] += S * (Us[q] - U[q]);
}

// err |= 1 where the reference aborts (c^2 < 0 or an undefined S*)
__device__ __forceinline__ void flux_hllc(const DevEos& E, int d, const double* Vm, const double* Vp, double* F,
                                          int& err) {
  RoeHat h;
  roe_hat(E, Vm, Vp, vsel(Vp, d) - vsel(Vm, d), h);
  double em = eos_e(E, Vm[0], Vm[4]);
  double ep = eos_e(E, Vp[0], Vp[4]);
  double cm2 = eos_c2(E, Vm[0], em);
  double cp2 = eos_c2(E, Vp[0], ep);
  if (cm2 < 0 || cp2 < 0) err |= 1;
  double cm = sqrt(cm2), cp = sqrt(cp2);
  double un_hat = wavg(h.srm, vsel(Vm, d), h.srp, vsel(Vp, d));
  double a = vsel(Vm, d) - cm, b = un_hat - h.c_hat;
  double Sm = (b < a) ? b : a;                 // std::min
  double a2 = vsel(Vp, d) + cp, b2 = un_hat + h.c_hat;
  double Sp = (a2 < b2) ? b2 : a2;             // std::max
  if (0.0 <= Sm) {
    flux_phys(E, d, Vm, F);
  } else if (0.0 >= Sp) {
    flux_phys(E, d, Vp, F);
  } else {
    double vm = vsel(Vm, d), vp = vsel(Vp, d);
    double Ss = (Vp[4] - Vm[4] + Vm[0] * vm * (Sm - vm) - Vp[0] * vp * (Sp - vp)) /
                (Vm[0] * (Sm - vm) - Vp[0] * (Sp - vp));
    if (Ss >= 0.0) hllc_star(E, d, Vm, Sm, Ss, F);
    else {
      if (!(Ss <= 0.0)) err |= 1;
      hllc_star(E, d, Vp, Sp, Ss, F);
    }
  }
}

__device__ __forceinline__ void flux_roe(const DevEos& E, int d, double del, const double* Vm, const double* Vp,
                                         double* F) {
  double du = Vp[1] - Vm[1], dv = Vp[2] - Vm[2], dw = Vp[3] - Vm[3];
  double dn = (d == 0) ? du : (d == 1 ? dv : dw);
  double srm = sqrt(Vm[0]), srp = sqrt(Vp[0]);
  double uh = wavg(srm, Vm[1], srp, Vp[1]);
  double vh = wavg(srm, Vm[2], srp, Vp[2]);
  double wh = wavg(srm, Vm[3], srp, Vp[3]);
  double Hh = wavg(srm, eos_H(E, Vm), srp, eos_H(E, Vp));
  RoeHat h;
  roe_hat(E, Vm, Vp, dn, h);
  const double c = h.c_hat, c2 = h.c2_hat, rh = h.rho_hat;
  double un = (d == 0) ? uh : (d == 1 ? vh : wh);
  double vel2 = uh * uh + vh * vh + wh * wh;
  double Hcon = h.e_hat + 0.5 * vel2 - h.dpdrho_hat / h.Gamma_hat;
  double lam[5] = {un - c, un, un, un, un + c};
  double a[5];
  a[0] = (h.dp - rh * c * dn) / (2 * c2);
  a[4] = (h.dp + rh * c * dn) / (2 * c2);
  double r[5][5];
#pragma unroll
  for (int p = 0; p < 5; p++)
#pragma unroll
    for (int q = 0; q < 5; q++) r[p][q] = 0.0;
  r[0][0] = 1; r[0][4] = Hh - un * c;
  r[4][0] = 1; r[4][4] = Hh + un * c;
  r[0][1] = (d == 0) ? un - c : uh; r[0][2] = (d == 1) ? un - c : vh; r[0][3] = (d == 2) ? un - c : wh;
  r[4][1] = (d == 0) ? un + c : uh; r[4][2] = (d == 1) ? un + c : vh; r[4][3] = (d == 2) ? un + c : wh;
  // the entropy/shear vectors take the slot of the normal direction
  // (FluxFcnBase.h:139-154 and the G/H counterparts)
  if (d == 0) {
    r[1][0] = 1; r[1][1] = uh; r[1][2] = vh; r[1][3] = wh; r[1][4] = Hcon;
    r[2][2] = 1; r[2][4] = vh;
    r[3][3] = 1; r[3][4] = wh;
    a[1] = -h.dp / c2 + h.drho; a[2] = rh * dv; a[3] = rh * dw;
  } else if (d == 1) {
    r[1][1] = 1; r[1][4] = uh;
    r[2][0] = 1; r[2][1] = uh; r[2][2] = vh; r[2][3] = wh; r[2][4] = Hcon;
    r[3][3] = 1; r[3][4] = wh;
    a[1] = rh * du; a[2] = -h.dp / c2 + h.drho; a[3] = rh * dw;
  } else {
    r[1][1] = 1; r[1][4] = uh;
    r[2][2] = 1; r[2][4] = vh;
    r[3][0] = 1; r[3][1] = uh; r[3][2] = vh; r[3][3] = wh; r[3][4] = Hcon;
    a[1] = rh * du; a[2] = rh * dv; a[3] = -h.dp / c2 + h.drho;
  }
  double tol = lam[0];
#pragma unroll
  for (int i = 1; i < 5; i++)
    if (fabs(lam[i]) > tol) tol = fabs(lam[i]);
  tol *= del;
#pragma unroll
  for (int i = 0; i < 5; i++)
    if (fabs(lam[i]) < tol) lam[i] = (lam[i] * lam[i] + tol * tol) / (2 * tol);
  double fm[5], fp[5];
  flux_phys(E, d, Vm, fm);
  flux_phys(E, d, Vp, fp);
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double f = 0.5 * (fm[i] + fp[i]);
#pragma unroll
    for (int p = 0; p < 5; p++) f -= 0.5 * fabs(lam[p]) * a[p] * r[p][i];
    F[i] = f;
  }
}

__device__ __forceinline__ void flux_llf(const DevEos& Em, const DevEos& Ep, int d, const double* Vm,
                                         const double* Vp, double* F, int& err) {
  double em = eos_e(Em, Vm[0], Vm[4]);
  double ep = eos_e(Ep, Vp[0], Vp[4]);
  double cm2 = eos_c2(Em, Vm[0], em), cp2 = eos_c2(Ep, Vp[0], ep);
  if (cm2 < 0 || cp2 < 0) err |= 1;
  double x = fabs(vsel(Vm, d)) + sqrt(cm2), y = fabs(vsel(Vp, d)) + sqrt(cp2);
  double a = (x < y) ? y : x;  // std::max
  double fm[5], fp[5], Um[5], Up[5];
  flux_phys(Em, d, Vm, fm);
  flux_phys(Ep, d, Vp, fp);
  eos_p2c(Em, Vm, Um);
  eos_p2c(Ep, Vp, Up);
#pragma unroll
  for (int i = 0; i < 5; i++) F[i] = 0.5 * ((fm[i] + fp[i]) - a * (Up[i] - Um[i]));
}

__device__ __forceinline__ void flux_num(const DevEos& E, int kind, double del, int d, const double* Vm,
                                         const double* Vp, double* F, int& err) {
  if (kind == M2CX_FLUX_HLLC) flux_hllc(E, d, Vm, Vp, F, err);
  else if (kind == M2CX_FLUX_ROE) flux_roe(E, d, del, Vm, Vp, F);
  else flux_llf(E, E, d, Vm, Vp, F, err);
}

// ---------------------------------------------------------------- limiters
__device__ __forceinline__ double lim_gmm(double A, double B, double alpha, double dq0, double dq1) {
  if (dq0 < 0.0 && dq1 < 0.0) {
    double x = alpha * dq0, y = B / (A + 1.0) * (dq0 + dq1), z = alpha * dq1;
    double yz = (y < z) ? z : y;  // std::max(y, z)
    return (x < yz) ? yz : x;     // std::max(x, yz)
  } else if (dq0 > 0.0 && dq1 > 0.0) {
    double x = alpha * dq0, y = B / (A + 1.0) * (dq0 + dq1), z = alpha * dq1;
    double yz = (z < y) ? z : y;  // std::min(y, z)
    return (yz < x) ? yz : x;     // std::min(x, yz)
  }
  return 0.0;
}
__device__ __forceinline__ double ipow_dev(double x, int k) {
  double r = x;
  for (int i = 1; i < k; i++) r *= x;
  return r;
}
__device__ __forceinline__ double lim_va(double A, double B, int k, double dq0, double dq1) {
  if (dq0 * dq1 <= 0) return 0.0;
  if (fabs(dq0 * dq1) < 1e-16) return 0.0;
  double a, b;
  if (k == 2) { a = dq0 * dq0; b = dq1 * dq1; }
  else { a = ipow_dev(dq0, k); b = ipow_dev(dq1, k); }
  return B * (dq1 * a + dq0 * b) / (a + A * b);
}


// ---------------------------------------------------------------- characteristic variables
// Eigenvectors of the flux Jacobian along axis d about a cell state V0
// (FluxFcnBase.h:309-517): L = R^{-1} (left), R (right) and Q = dU/dV.  The
// entries and the sums of products are formed in the reference's order, so
// every product rounds as it does there.
struct CharBasis {
  double e0, c0, dpdrho0, Gamma0, kin0, un0;
  int d;
};
// false when c0^2 < 0, where the reference aborts (FluxFcnBase.h:315-320)
__device__ __forceinline__ bool char_setup(const DevEos& E, const double* V0, int d, CharBasis& b) {
  b.e0 = eos_e(E, V0[0], V0[4]);
  const double c2 = eos_c2(E, V0[0], b.e0);
  b.c0 = sqrt(c2);
  b.dpdrho0 = eos_dpdrho(E, V0[0], b.e0);
  b.Gamma0 = eos_Gamma(E, V0[0]);
  b.kin0 = 0.5 * (V0[1] * V0[1] + V0[2] * V0[2] + V0[3] * V0[3]);
  b.un0 = vsel(V0, d);
  b.d = d;
  return !(c2 < 0);
}
__device__ __forceinline__ void matvec5(const double (&M)[5][5], const double* x, double* y) {
#pragma unroll
  for (int r = 0; r < 5; r++) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) acc += M[r][q] * x[q];
    y[r] = acc;
  }
}
// y = L x (FluxFcnBase.h:336-358)
__device__ __forceinline__ void char_left(const CharBasis& b, const double* V0, const double* x, double* y) {
  const double al = V0[0] * b.Gamma0 / (V0[4] * b.Gamma0 + V0[0] * b.dpdrho0);
  const double be = (b.e0 - b.kin0 + V0[4] / V0[0]) * al;
  double L[5][5];
  L[0][0] = 0.5 * (1.0 - be + b.un0 / b.c0);
  L[4][0] = 0.5 * (1.0 - be - b.un0 / b.c0);
  L[0][4] = al / 2.0;
  L[4][4] = al / 2.0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const double nr = (r == b.d) ? 1.0 : 0.0;
    L[0][1 + r] = -0.5 * (al * V0[1 + r] + nr / b.c0);
    L[4][1 + r] = -0.5 * (al * V0[1 + r] - nr / b.c0);
    L[1 + r][0] = -V0[1 + r] + (be + b.un0) * nr;
    L[1 + r][4] = -al * nr;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double nq = (q == b.d) ? 1.0 : 0.0;
      L[1 + r][1 + q] = (r == q ? 1.0 : 0.0) + al * nr * V0[1 + q] - nr * nq;
    }
  }
  matvec5(L, x, y);
}
// y = R x (FluxFcnBase.h:392-417)
__device__ __forceinline__ void char_right(const CharBasis& b, const double* V0, const double* x, double* y) {
  const double H0 = b.e0 + b.kin0 + V0[4] / V0[0];
  const double par = b.e0 + b.kin0 - b.dpdrho0 / b.Gamma0 - b.un0;
  double R[5][5];
  R[0][0] = 1.0;
  R[0][4] = 1.0;
  R[4][0] = H0 - b.un0 * b.c0;
  R[4][4] = H0 + b.un0 * b.c0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const double nr = (r == b.d) ? 1.0 : 0.0;
    R[0][1 + r] = nr;
    R[1 + r][0] = V0[1 + r] - b.c0 * nr;
    R[1 + r][4] = V0[1 + r] + b.c0 * nr;
    R[4][1 + r] = V0[1 + r] + par * nr;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double nq = (q == b.d) ? 1.0 : 0.0;
      R[1 + r][1 + q] = (r == q ? 1.0 : 0.0) + V0[1 + r] * nq - nr * nq;
    }
  }
  matvec5(R, x, y);
}
// primitive difference -> characteristic: L Q dV (FluxFcnBase.h:309-362)
__device__ __forceinline__ void char_from_prim(const CharBasis& b, const double* V0, const double* dV, double* dW) {
  const double q51 = b.e0 + b.kin0 - b.dpdrho0 / b.Gamma0;
  double t[5];
  t[0] = dV[0];
  t[1] = V0[1] * dV[0] + V0[0] * dV[1];
  t[2] = V0[2] * dV[0] + V0[0] * dV[2];
  t[3] = V0[3] * dV[0] + V0[0] * dV[3];
  t[4] = q51 * dV[0] + V0[0] * V0[1] * dV[1] + V0[0] * V0[2] * dV[2] + V0[0] * V0[3] * dV[3] +
         1.0 / b.Gamma0 * dV[4];
  char_left(b, V0, t, dW);
}
// characteristic -> primitive difference: Q^{-1} R dW (FluxFcnBase.h:367-421)
__device__ __forceinline__ void char_to_prim(const CharBasis& b, const double* V0, const double* dW, double* dV) {
  double t[5];
  char_right(b, V0, dW, t);
  dV[0] = t[0];
  dV[1] = -V0[1] / V0[0] * t[0] + t[1] / V0[0];
  dV[2] = -V0[2] / V0[0] * t[0] + t[2] / V0[0];
  dV[3] = -V0[3] / V0[0] * t[0] + t[3] / V0[0];
  dV[4] = ((-b.e0 + b.kin0) * b.Gamma0 + b.dpdrho0) * t[0] -
          b.Gamma0 * (V0[1] * t[1] + V0[2] * t[2] + V0[3] * t[3] - t[4]);
}

}  // namespace m2c
