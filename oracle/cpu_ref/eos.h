// Oracle (test infrastructure only): CPU restatement of the reference's
// equations of state.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this code, and only as the checker.
//
// Restates VarFcnBase.h:36-298 (shared operations), VarFcnSG.h:64-108,
// VarFcnNASG.h:55-86, VarFcnMG.h:70-99,142-178 and VarFcnDummy.h:18-58.
// Operation order inside every expression follows the reference so that
// results agree to the last bit with the reference built without FMA
// contraction.
#pragma once
#include <cmath>
#include <cfloat>
#include <algorithm>
#include "m2c_amd.h"

namespace orc {

struct Eos {
  int type = M2CX_EOS_SG;
  // SG / NASG
  double gam = 1.4, pc = 0.0, gam1 = 0.4, b = 0.0, q = 0.0, gam_pc = 0.0;
  // MG
  double rho0 = 0, c0 = 0, Gamma0 = 0, s = 0, e0 = 0, rho0_c0_c0 = 0, Gamma0_over_2 = 0, Gamma0_rho0 = 0;
  // clipping (MaterialModelData, IoData.cpp:1048-1053)
  double rhomin = 0.0, pmin = -DBL_MAX, rhomax = DBL_MAX, pmax = DBL_MAX, failsafe_density = 1e-10;
  // dummy state (VarFcnDummy.h:21-26)
  double d_rho = 0, d_p = 0, d_e = 0;

  static Eos from(const m2cx_material& m) {
    Eos e;
    e.type = m.eos;
    e.rhomin = m.rhomin; e.pmin = m.pmin; e.rhomax = m.rhomax; e.pmax = m.pmax;
    e.failsafe_density = m.failsafe_density;
    if (m.eos == M2CX_EOS_SG) {            // VarFcnSG.h:129-136
      e.gam = m.gamma; e.pc = m.pc; e.gam1 = e.gam - 1.0;
    } else if (m.eos == M2CX_EOS_NASG) {   // VarFcnNASG.h:104-113
      e.gam = m.gamma; e.pc = m.pc; e.q = m.q; e.b = m.b;
      e.gam1 = e.gam - 1.0; e.gam_pc = e.gam * e.pc;
    } else if (m.eos == M2CX_EOS_MG) {     // VarFcnMG.h:108-126
      e.rho0 = m.rho0; e.c0 = m.c0; e.Gamma0 = m.Gamma0; e.s = m.s; e.e0 = m.e0;
      e.rho0_c0_c0 = e.rho0 * e.c0 * e.c0;
      e.Gamma0_over_2 = 0.5 * e.Gamma0;
      e.Gamma0_rho0 = e.Gamma0 * e.rho0;
    } else {                               // VarFcnDummy.h:21-23
      e.d_rho = m.dummy_rho; e.d_p = m.dummy_p; e.d_e = m.dummy_e;
    }
    return e;
  }

  bool dummy() const { return type == M2CX_EOS_DUMMY; }

  double GetPressure(double rho, double e) const {
    switch (type) {
      case M2CX_EOS_SG: return gam1 * rho * e - gam * pc;
      case M2CX_EOS_NASG: return gam1 * (e - q) / (1.0 / rho - b) - gam_pc;
      case M2CX_EOS_MG: {
        double eta = 1.0 - rho0 / rho;
        return rho0_c0_c0 * eta * (1.0 - Gamma0_over_2 * eta) / ((1.0 - s * eta) * (1.0 - s * eta)) +
               Gamma0_rho0 * (e - e0);
      }
      default: return d_p;
    }
  }
  double GetInternalEnergyPerUnitMass(double rho, double p) const {
    switch (type) {
      case M2CX_EOS_SG: return (p + gam * pc) / (gam1 * rho);
      case M2CX_EOS_NASG: return (1.0 / gam1) * (p + gam_pc) * (1.0 / rho - b) + q;
      case M2CX_EOS_MG: {
        double eta = 1.0 - rho0 / rho;
        return (p - rho0_c0_c0 * eta * (1.0 - Gamma0_over_2 * eta) / ((1.0 - s * eta) * (1.0 - s * eta))) /
                   Gamma0_rho0 + e0;
      }
      default: return d_e;
    }
  }
  double GetDpdrho(double rho, double e) const {
    switch (type) {
      case M2CX_EOS_SG: return gam1 * e;
      case M2CX_EOS_NASG: {
        double V = 1.0 / rho;
        return gam1 * V * V * (e - q) / ((V - b) * (V - b));
      }
      case M2CX_EOS_MG: {
        double eta = 1.0 - rho0 / rho;
        double S = 1.0 - s * eta;
        return rho0_c0_c0 * (1.0 + (s - Gamma0) * eta) / (S * S * S) * rho0 / (rho * rho);
      }
      default: return 0.0;
    }
  }
  double GetBigGamma(double rho, double /*e*/) const {
    switch (type) {
      case M2CX_EOS_SG: return gam1;
      case M2CX_EOS_NASG: return gam1 / (1.0 - b * rho);
      case M2CX_EOS_MG: return Gamma0_rho0 / rho;
      default: return 0.0;
    }
  }
  // VarFcnBase.h:174-179 / VarFcnDummy.h:50
  double ComputeSoundSpeedSquare(double rho, double e) const {
    if (dummy()) return DBL_MIN;
    return GetDpdrho(rho, e) + GetPressure(rho, e) / rho * GetBigGamma(rho, e);
  }
  // VarFcnBase.h:220-223
  double ComputeTotalEnthalpyPerUnitMass(const double* V) const {
    if (dummy()) return d_e;
    double e = GetInternalEnergyPerUnitMass(V[0], V[4]);
    return e + 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) + V[4] / V[0];
  }
  // CheckState(rho,p): SG VarFcnSG.h:88-100, NASG VarFcnNASG.h:63-75,
  // MG uses the base class VarFcnBase.h:100-117, dummy never fails.
  bool CheckState(double rho, double p) const {
    if (dummy()) return false;
    if (!std::isfinite(rho) || !std::isfinite(p)) return true;
    if (type == M2CX_EOS_SG || type == M2CX_EOS_NASG) return rho <= 0.0 || p + pc <= 0.0;
    if (rho <= 0.0) return true;
    double e = GetInternalEnergyPerUnitMass(rho, p);
    double c2 = GetDpdrho(rho, e) + p / rho * GetBigGamma(rho, e);
    return c2 <= 0;
  }
  // VarFcnBase.h:103-111
  bool CheckState(const double* V) const {
    if (dummy()) return false;
    for (int i = 0; i < 5; i++)
      if (!std::isfinite(V[i])) return true;
    return CheckState(V[0], V[4]);
  }
  // VarFcnBase.h:120-151
  void ConservativeToPrimitive(const double* U, double* V) const {
    if (dummy()) { V[0] = d_rho; V[1] = V[2] = V[3] = 0.0; V[4] = d_p; return; }
    V[0] = U[0];
    double invRho = 1.0 / U[0];
    V[1] = U[1] * invRho;
    V[2] = U[2] * invRho;
    V[3] = U[3] * invRho;
    double e = (U[4] - 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3])) * invRho;
    V[4] = GetPressure(V[0], e);
  }
  void PrimitiveToConservative(const double* V, double* U) const {
    if (dummy()) { U[0] = d_rho; U[1] = U[2] = U[3] = 0.0; U[4] = d_rho * d_e; return; }
    U[0] = V[0];
    U[1] = V[0] * V[1];
    U[2] = V[0] * V[2];
    U[3] = V[0] * V[3];
    double e = GetInternalEnergyPerUnitMass(V[0], V[4]);
    U[4] = V[0] * (e + 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]));
  }
  // VarFcnBase.h:263-298 (U may be null)
  bool ClipDensityAndPressure(double* V, double* U = nullptr) const {
    if (dummy()) {
      V[0] = d_rho; V[1] = V[2] = V[3] = 0.0; V[4] = d_p;
      if (U) { U[0] = d_rho; U[1] = U[2] = U[3] = 0.0; U[4] = d_rho * d_e; }
      return false;
    }
    bool clip = false;
    if (V[0] < rhomin) { V[0] = rhomin; clip = true; }
    if (V[4] < pmin) { V[4] = pmin; clip = true; }
    if (V[0] > rhomax) { V[0] = rhomax; clip = true; }
    if (V[4] > pmax) { V[4] = pmax; clip = true; }
    if (clip && U) PrimitiveToConservative(V, U);
    return clip;
  }
};

}  // namespace orc
