// Oracle (test infrastructure only): CPU restatement of the reference's
// numerical flux functions.  Restates FluxFcnBase.h:83-123 (physical
// fluxes), :139-154 (Jacobian eigen-structure used by GenRoe), :287-303 (max
// eigenvalues), FluxFcnHLLC.h:58-245, FluxFcnGenRoe.h:64-221 and
// FluxFcnLLF.h:47-185 and FluxFcnGodunov.h:37-60, keeping the reference's
// operation order.
#pragma once
#include <vector>
#include "eos.h"
#include "riemann.h"

namespace orc {

struct Flux {
  const std::vector<Eos>* vf = nullptr;
  int kind = M2CX_FLUX_HLLC;
  double del = 0.2;   // Harten entropy fix (FluxFcnGenRoe.h:30-31)
  // FluxFcnGodunov's own exact solver (FluxFcnGodunov.h:30); set when kind is
  // M2CX_FLUX_GODUNOV.  `godunov_fatal` records a solve the reference would
  // have ended the run on.
  ExactRiemann* godunov = nullptr;
  mutable bool godunov_fatal = false;
  static constexpr double eps = 1e-10;  // Hu et al. (FluxFcnHLLC.h:24-25)

  const Eos& E(int id) const { return (*vf)[id]; }

  // FluxFcnBase.h:83-123
  void Physical(int dir, const double* V, int id, double* F) const {
    double H = E(id).ComputeTotalEnthalpyPerUnitMass(V);
    double rhoun = V[0] * V[dir + 1];
    F[0] = rhoun;
    F[1] = rhoun * V[1];
    F[2] = rhoun * V[2];
    F[3] = rhoun * V[3];
    F[dir + 1] = rhoun * V[dir + 1] + V[4];
    F[4] = rhoun * H;
  }

  // FluxFcnBase.h:287-303
  void MaxEigenvalues(const double* V, int id, double& lf, double& lg, double& lh, bool& bad) const {
    double e = E(id).GetInternalEnergyPerUnitMass(V[0], V[4]);
    double c = E(id).ComputeSoundSpeedSquare(V[0], e);
    if (c < 0) bad = true;
    c = std::sqrt(c);
    lf = std::fabs(V[1]) + c;
    lg = std::fabs(V[2]) + c;
    lh = std::fabs(V[3]) + c;
  }

  static double Average(double w1, double f1, double w2, double f2) { return (w1 * f1 + w2 * f2) / (w1 + w2); }

  // Shared "hat" state of FluxFcnHLLC.h:67-128 and FluxFcnGenRoe.h:71-140.
  struct Hat {
    double rho_hat, sqrt_rhom, sqrt_rhop, drho, dp, e_hat, c_hat, c_hat_square, dpdrho_hat, Gamma_hat;
  };
  void RoeHat(int dir, const double* Vm, const double* Vp, int id, Hat& h, double diffvelo) const {
    h.rho_hat = std::sqrt(Vm[0] * Vp[0]);
    h.sqrt_rhom = std::sqrt(Vm[0]);
    h.sqrt_rhop = std::sqrt(Vp[0]);
    h.drho = Vp[0] - Vm[0];
    h.dp = Vp[4] - Vm[4];
    (void)dir;
    double diff = diffvelo / (h.sqrt_rhom + h.sqrt_rhop);
    double p_over_rho_hat = Average(h.sqrt_rhom, Vm[4] / Vm[0], h.sqrt_rhop, Vp[4] / Vp[0]) + 0.5 * diff * diff;
    double em = E(id).GetInternalEnergyPerUnitMass(Vm[0], Vm[4]);
    double ep = E(id).GetInternalEnergyPerUnitMass(Vp[0], Vp[4]);
    double de = ep - em;
    h.e_hat = Average(h.sqrt_rhom, em, h.sqrt_rhop, ep);
    double w_rho = h.drho * h.drho / (h.rho_hat * h.rho_hat);
    double w_e = de * de / (h.e_hat * h.e_hat);
    double denominator = w_rho + w_e + eps;
    double dpdrho_roeavg = Average(h.sqrt_rhom, E(id).GetDpdrho(Vm[0], em), h.sqrt_rhop, E(id).GetDpdrho(Vp[0], ep));
    double Gamma_roeavg = Average(h.sqrt_rhom, E(id).GetBigGamma(Vm[0], em), h.sqrt_rhop, E(id).GetBigGamma(Vp[0], ep));
    double t1 = (w_e + eps) * dpdrho_roeavg;
    double t2 = (h.dp - Gamma_roeavg * h.rho_hat * de) * h.drho / (h.rho_hat * h.rho_hat);
    h.dpdrho_hat = (t1 + t2) / denominator;
    double g1 = (w_rho + eps) * Gamma_roeavg;
    double g2 = (h.dp - dpdrho_roeavg * h.drho) * de / (h.rho_hat * h.e_hat * h.e_hat);
    h.Gamma_hat = (g1 + g2) / denominator;
    h.c_hat_square = h.dpdrho_hat + h.Gamma_hat * p_over_rho_hat;
    if (h.c_hat_square <= 0) {
      h.c_hat = eps;
      h.c_hat_square = h.c_hat * h.c_hat;
    } else
      h.c_hat = std::sqrt(h.c_hat_square);
  }

  // FluxFcnHLLC.h:159-192
  void HllcFstar(int dir, const double* V, double S, double Sstar, int id, double* Fs) const {
    double velo = V[dir + 1];
    double rhostar = V[0] * (S - velo) / (S - Sstar);
    double U[5];
    E(id).PrimitiveToConservative(V, U);
    double Us[5];
    Us[0] = 1.0;
    Us[1] = (dir == 0) ? Sstar : V[1];
    Us[2] = (dir == 1) ? Sstar : V[2];
    Us[3] = (dir == 2) ? Sstar : V[3];
    Us[4] = U[4] / V[0] + (Sstar - velo) * (Sstar + V[4] / (V[0] * (S - velo)));
    for (int i = 0; i < 5; i++) Us[i] *= rhostar;
    Physical(dir, V, id, Fs);
    for (int i = 0; i < 5; i++) Fs[i] += S * (Us[i] - U[i]);
  }

  // FluxFcnHLLC.h:58-154, 197-245
  void Hllc(int dir, const double* Vm, const double* Vp, int id, double* F) const {
    Hat h;
    RoeHat(dir, Vm, Vp, id, h, Vp[dir + 1] - Vm[dir + 1]);
    double em = E(id).GetInternalEnergyPerUnitMass(Vm[0], Vm[4]);
    double ep = E(id).GetInternalEnergyPerUnitMass(Vp[0], Vp[4]);
    double cm = std::sqrt(E(id).ComputeSoundSpeedSquare(Vm[0], em));
    double cp = std::sqrt(E(id).ComputeSoundSpeedSquare(Vp[0], ep));
    double un_hat = Average(h.sqrt_rhom, Vm[dir + 1], h.sqrt_rhop, Vp[dir + 1]);
    double Sm = std::min(Vm[dir + 1] - cm, un_hat - h.c_hat);
    double Sp = std::max(Vp[dir + 1] + cp, un_hat + h.c_hat);
    if (0.0 <= Sm)
      Physical(dir, Vm, id, F);
    else if (0.0 >= Sp)
      Physical(dir, Vp, id, F);
    else {
      double vm = Vm[dir + 1], vp = Vp[dir + 1];
      double Sstar = (Vp[4] - Vm[4] + Vm[0] * vm * (Sm - vm) - Vp[0] * vp * (Sp - vp)) /
                     (Vm[0] * (Sm - vm) - Vp[0] * (Sp - vp));
      if (Sstar >= 0.0)
        HllcFstar(dir, Vm, Sm, Sstar, id, F);
      else
        HllcFstar(dir, Vp, Sp, Sstar, id, F);  // Sstar <= 0 (NaN also lands here)
    }
  }

  // FluxFcnGenRoe.h:64-221 (eigenvectors follow FluxFcnBase.h:139-154 and
  // their G/H counterparts)
  void Roe(int dir, const double* Vm, const double* Vp, int id, double* F) const {
    double du = Vp[1] - Vm[1], dv = Vp[2] - Vm[2], dw = Vp[3] - Vm[3];
    double dn = (dir == 0) ? du : (dir == 1 ? dv : dw);
    Hat h;
    double sqrt_rhom = std::sqrt(Vm[0]);
    double sqrt_rhop = std::sqrt(Vp[0]);
    double u_hat = Average(sqrt_rhom, Vm[1], sqrt_rhop, Vp[1]);
    double v_hat = Average(sqrt_rhom, Vm[2], sqrt_rhop, Vp[2]);
    double w_hat = Average(sqrt_rhom, Vm[3], sqrt_rhop, Vp[3]);
    double Hm = E(id).ComputeTotalEnthalpyPerUnitMass(Vm);
    double Hp = E(id).ComputeTotalEnthalpyPerUnitMass(Vp);
    double H_hat = Average(sqrt_rhom, Hm, sqrt_rhop, Hp);
    RoeHat(dir, Vm, Vp, id, h, dn);
    double c = h.c_hat, c2 = h.c_hat_square, rh = h.rho_hat;
    double vel[3] = {u_hat, v_hat, w_hat};
    double un = vel[dir];
    double vel2 = u_hat * u_hat + v_hat * v_hat + w_hat * w_hat;
    double lam[5] = {un - c, un, un, un, un + c};
    double r[5][5];
    // r1, r5: acoustic; r2..r4: entropy + shear, placed as in FluxFcnBase.h
    double Hcon = h.e_hat + 0.5 * vel2 - h.dpdrho_hat / h.Gamma_hat;
    for (int p = 0; p < 5; p++)
      for (int q = 0; q < 5; q++) r[p][q] = 0.0;
    r[0][0] = 1; r[0][1] = u_hat; r[0][2] = v_hat; r[0][3] = w_hat; r[0][4] = H_hat - un * c;
    r[4][0] = 1; r[4][1] = u_hat; r[4][2] = v_hat; r[4][3] = w_hat; r[4][4] = H_hat + un * c;
    r[0][dir + 1] = un - c;
    r[4][dir + 1] = un + c;
    double a[5];
    a[0] = (h.dp - rh * c * dn) / (2 * c2);
    a[4] = (h.dp + rh * c * dn) / (2 * c2);
    if (dir == 0) {
      // r2 = (1,u,v,w,Hcon), r3 = (0,0,1,0,v), r4 = (0,0,0,1,w)
      r[1][0] = 1; r[1][1] = u_hat; r[1][2] = v_hat; r[1][3] = w_hat; r[1][4] = Hcon;
      r[2][2] = 1; r[2][4] = v_hat;
      r[3][3] = 1; r[3][4] = w_hat;
      a[1] = -h.dp / c2 + h.drho; a[2] = rh * dv; a[3] = rh * dw;
    } else if (dir == 1) {
      // r2 = (0,1,0,0,u), r3 = (1,u,v,w,Hcon), r4 = (0,0,0,1,w)
      r[1][1] = 1; r[1][4] = u_hat;
      r[2][0] = 1; r[2][1] = u_hat; r[2][2] = v_hat; r[2][3] = w_hat; r[2][4] = Hcon;
      r[3][3] = 1; r[3][4] = w_hat;
      a[1] = rh * du; a[2] = -h.dp / c2 + h.drho; a[3] = rh * dw;
    } else {
      // r2 = (0,1,0,0,u), r3 = (0,0,1,0,v), r4 = (1,u,v,w,Hcon)
      r[1][1] = 1; r[1][4] = u_hat;
      r[2][2] = 1; r[2][4] = v_hat;
      r[3][0] = 1; r[3][1] = u_hat; r[3][2] = v_hat; r[3][3] = w_hat; r[3][4] = Hcon;
      a[1] = rh * du; a[2] = rh * dv; a[3] = -h.dp / c2 + h.drho;
    }
    // entropy fix, FluxFcnGenRoe.h:186-192, 217-221
    double tol = lam[0];
    for (int i = 1; i < 5; i++)
      if (std::fabs(lam[i]) > tol) tol = std::fabs(lam[i]);
    tol *= del;
    for (int i = 0; i < 5; i++)
      if (std::fabs(lam[i]) < tol) lam[i] = (lam[i] * lam[i] + tol * tol) / (2 * tol);
    double fm[5], fp[5];
    Physical(dir, Vm, id, fm);
    Physical(dir, Vp, id, fp);
    for (int i = 0; i < 5; i++) {
      F[i] = 0.5 * (fm[i] + fp[i]);
      for (int p = 0; p < 5; p++) F[i] -= 0.5 * std::fabs(lam[p]) * a[p] * r[p][i];
    }
  }

  // FluxFcnLLF.h:47-111 (idm == idp) and :119-185 (material interface)
  void Llf(int dir, const double* Vm, int idm, const double* Vp, int idp, double* F) const {
    double em = E(idm).GetInternalEnergyPerUnitMass(Vm[0], Vm[4]);
    double ep = E(idp).GetInternalEnergyPerUnitMass(Vp[0], Vp[4]);
    double cm = std::sqrt(E(idm).ComputeSoundSpeedSquare(Vm[0], em));
    double cp = std::sqrt(E(idp).ComputeSoundSpeedSquare(Vp[0], ep));
    double a = std::max(std::fabs(Vm[dir + 1]) + cm, std::fabs(Vp[dir + 1]) + cp);
    double fm[5], fp[5], Um[5], Up[5];
    Physical(dir, Vm, idm, fm);
    Physical(dir, Vp, idp, fp);
    E(idm).PrimitiveToConservative(Vm, Um);
    E(idp).PrimitiveToConservative(Vp, Up);
    for (int i = 0; i < 5; i++) F[i] = 0.5 * ((fm[i] + fp[i]) - a * (Up[i] - Um[i]));
  }

  // FluxFcnGodunov.h:37-60: the physical flux of the exact Riemann solution
  // (one material on both sides, along the mesh axis) at x/t = 0
  void Godunov(int dir, const double* Vm, const double* Vp, int id, double* F) const {
    double normal[3] = {0.0, 0.0, 0.0};
    normal[dir] = 1.0;
    double Vmid[5], Vsm[5], Vsp[5];
    int midid = id;
    godunov->st = RiemannStats();
    godunov->Solve(normal, Vm, id, Vp, id, Vmid, midid, Vsm, Vsp);
    if (godunov->st.fatal) godunov_fatal = true;
    Physical(dir, Vmid, id, F);
  }

  // FluxFcnBase::ComputeNumericalFluxAtCellInterface dispatch
  void Numerical(int dir, const double* Vm, const double* Vp, int id, double* F) const {
    if (kind == M2CX_FLUX_HLLC) Hllc(dir, Vm, Vp, id, F);
    else if (kind == M2CX_FLUX_ROE) Roe(dir, Vm, Vp, id, F);
    else if (kind == M2CX_FLUX_GODUNOV) Godunov(dir, Vm, Vp, id, F);
    else Llf(dir, Vm, id, Vp, id, F);
  }
};

}  // namespace orc
