// Oracle (test infrastructure only): CPU restatement of the explicit
// finite-volume step on one rank (no domain decomposition: the reference's
// PETSc halo exchanges are identities on a single subdomain).
#pragma once
#include <map>
#include <string>
#include <vector>
#include <array>
#include "eos.h"
#include "flux.h"
#include "riemann.h"

namespace orc {

using V5 = std::array<double, 5>;

// slope limiters and the van Albada exponent (Reconstructor.h:96-111,
// Reconstructor.cpp:103-121)
double GeneralizedMinMod(double A, double B, double alpha, double dq0, double dq1);
double VanAlbada(double A, double B, int k, double dq0, double dq1);
int SlopeLimiterK(int limiter, double A, double B);

struct Mesh {
  int NX, NY, NZ, PX, PY, PZ;
  std::vector<double> x, y, z, dx, dy, dz;  // indexed i+1 (ghosts included)
  long idx(int i, int j, int k) const { return ((long)(k + 1) * PY + (j + 1)) * PX + (i + 1); }
  long ncells() const { return (long)PX * PY * PZ; }
  void ijk(long c, int& i, int& j, int& k) const {
    i = (int)(c % PX) - 1;
    j = (int)((c / PX) % PY) - 1;
    k = (int)(c / ((long)PX * PY)) - 1;
  }
  bool inside(int i, int j, int k) const { return i >= 0 && i < NX && j >= 0 && j < NY && k >= 0 && k < NZ; }
  double X(int i) const { return x[i + 1]; }
  double Y(int j) const { return y[j + 1]; }
  double Z(int k) const { return z[k + 1]; }
  double DX(int i) const { return dx[i + 1]; }
  double DY(int j) const { return dy[j + 1]; }
  double DZ(int k) const { return dz[k + 1]; }
};

// A "face" ghost of the outer ghost layer (GhostPoint with type FACE,
// SpaceOperator.cpp:472-593).
struct Ghost {
  int i, j, k, ii, jj, kk, side, bc, lsbc;  // side: 0 L,1 R,2 B,3 T,4 K,5 F
};

struct RiemannStore {  // RiemannSolutions.h:12-30, keyed by cell linear index
  std::map<long, std::pair<V5, int>> left, right, bottom, top, back, front;
  void clear() { left.clear(); right.clear(); bottom.clear(); top.clear(); back.clear(); front.clear(); }
};

struct Counters {
  long riemann_calls = 0, riemann_trivial = 0, riemann_errors = 0, riemann_iters = 0, riemann_fevals = 0;
  long riemann_fallbacks = 0, clipped = 0, id_changed = 0, steps = 0;
  long reinit_calls = 0, reinit_iters = 0, unresolved = 0;
};

// A first-layer node of the level-set reinitialization with its neighbours
// across the interface (LevelSetReinitializer.h:55-81).
// a LevelSet[i] block taken from / put into the LevelSet[0] fields of a scheme
m2cx_levelset LSFromScheme(const m2cx_scheme& s);
void SchemeSetLS(m2cx_scheme& s, const m2cx_levelset& L);

struct FirstLayerNode {
  int i, j, k;
  bool s[6];
  int ns;
  double r[6], r0, f;
  double nphi0[3];
};

class Solver {
 public:
  Mesh m;
  m2cx_grid grid;
  m2cx_scheme sch;
  std::vector<Eos> vf;
  int nmat, inactive_id;
  Flux flux;
  ExactRiemann riem;
  ExactRiemann godunov_riem;  // the Godunov flux's own solver (FluxFcnGodunov.h:30)
  // record of interface Riemann problems (17 doubles each: dir, Vm, idm, Vp,
  // idp, err, iterations, f evaluations, fallback), for tests
  bool riem_log_on = false;
  std::vector<double> riem_log;
  // fields (AoS like SpaceVariable3D, ghost width 1)
  std::vector<double> V, V1, V2, Un, U1, R, Vl, Vr, Vb, Vt, Vk, Vf, vol;
  std::vector<double> Urec;  // conservative copy of V for conservative reconstruction
  std::vector<double> A3, B3;  // Reconstructor CoeffA/CoeffB (3 per cell)
  std::vector<int> K3;         // CoeffK
  std::vector<int> LSK3;       // CoeffK of the level set's own reconstruction
  std::vector<int> ID, IDn;
  std::vector<double> Phi, Phi1, Phi_tmp, Rls;
  std::vector<double> cxL, cxR, cyL, cyR, czL, czR;  // FD3 coefficients, 4 per cell
  std::vector<Ghost> ghosts;
  RiemannStore rs;
  Counters cnt;
  double t = 0.0, dt_last = 0.0, dts = 0.0;
  int time_step = 0;
  bool fatal = false;
  std::string msg;

  Solver(const m2cx_grid& g, const m2cx_material* mats, int nmat, const m2cx_scheme& s);

  void Fail(const std::string& s) { if (!fatal) { fatal = true; msg = s; } }

  // SpaceOperator
  void Reconstruct(const std::vector<double>& Vin);
  void ReconstructCell(const std::vector<double>& Vin, long c, int vtype, double* const* fl);
  void CheckReconstructedStates(const std::vector<double>& Vin);
  void ComputeAdvectionFluxes(const std::vector<double>& Vin, std::vector<double>& F, bool capture);
  void ComputeResidual(const std::vector<double>& Vin, std::vector<double>& Rout, bool capture);
  void ConservativeToPrimitive(const std::vector<double>& U, std::vector<double>& Vout);
  void PrimitiveToConservative(const std::vector<double>& Vin, std::vector<double>& U);
  int ClipDensityAndPressure(std::vector<double>& Vio);
  void ApplyBoundaryConditions(std::vector<double>& Vio);
  void ComputeTimeStepSize(const std::vector<double>& Vin, double& dt, double& cfl);
  // LevelSetOperator
  // with `useful`, only the ghosts it marks (the narrow-band reinitializer)
  void LSApplyBoundaryConditions(std::vector<double>& phi, const std::vector<unsigned char>* useful = nullptr);
  void LSComputeResidual(const std::vector<double>& Vin, const std::vector<double>& phi, std::vector<double>& res);
  void LSComputeResidualFVM(const std::vector<double>& Vin, const std::vector<double>& phi, std::vector<double>& res);
  void LSComputeResidualFVMNarrowBand(const std::vector<double>& Vin, const std::vector<double>& phi,
                                      std::vector<double>& res);
  // direction of the 1D Riemann problem at material interfaces
  void GradPhiAtFace(const std::vector<double>& phi, int d, int i, int j, int k, double* dir) const;
  void BimaterialNormal(int d, int i, int j, int k, int myid, int neighborid, double* dir) const;
  // LevelSetReinitializer, full domain (LevelSetReinitializer.cpp:76-176)
  bool ReinitEnabled() const { return sch.reinit_frequency > 0 || sch.reinit_frequency_dt > 0; }
  // LevelSetOperator::Reinitialize (LevelSetOperator.cpp:1338-1363); returns
  // whether it ran
  bool Reinitialize(double time, double dt, int time_step, int special_maxIts, bool must_do);
  void ReinitializeFullDomain(int special_maxIts);
  double reinit_cfl = 0.8;  // the reinitializer's CFL, reduced by its failsafe
  // Narrow-band level set (LevelSetOperator.cpp:84-99; LevelSetReinitializer.
  // cpp:181-300, 440-598, 1045-1133, 1459-2031): bandwidth (0: full domain),
  // the band levels, the useful / active masks and node lists, the list of
  // useful nodes plus one layer, and the persistent phi bounds.
  int nb_bw = 0;
  bool nb_built = false;
  std::vector<int> nbLevel;
  std::vector<unsigned char> nbUseful, nbActive;
  std::vector<long> useful_nodes, active_nodes, plus1_nodes;
  // the finite-volume level set's reconstruction outputs (Phil..Phif, ul..wf,
  // dudx..dwdz of LevelSetOperator), kept between calls: in the band only the
  // active nodes are rewritten, and the fluxes read the others as left
  // (zero before a first write)
  std::vector<double> fvPf, fvUf, fvSl;
  double domain_diagonal = 0.0, phi_max = 0.0, phi_min = 0.0, phi_out_pos = 0.0, phi_out_neg = 0.0;
  void ConstructNarrowBand();
  // Several level sets: level set 0 lives in the members above; level set
  // q >= 1 in lsx[q - 1], exchanged with them by SwapLS(q) while its own
  // operator runs (swapping again puts it back).
  struct LSState {
    m2cx_levelset set{};
    std::vector<double> Phi, Phi1, Phi_tmp, Rls, rTag, rSign, rPhibk, rPhi0, rPhi1, rR, rNormal, rPhiG2;
    std::vector<int> LSK3, nbLevel;
    std::vector<unsigned char> nbUseful, nbActive;
    std::vector<double> fvPf, fvUf, fvSl;
    std::vector<long> useful_nodes, active_nodes, plus1_nodes;
    int nb_bw = 0;
    bool nb_built = false;
    double phi_max = 0, phi_min = 0, phi_out_pos = 0, phi_out_neg = 0, reinit_cfl = 0.8;
  };
  int nls = 0;
  std::vector<LSState> lsx;
  int ls_correction_bucket = 0;  // TimeIntegrator.h:31
  void SetLevelSets(const m2cx_levelset* ls, int n);
  void DeriveLevelSet();
  void SwapLS(int q);
  template <class F>
  void ForEachLS(F fn) {
    for (int q = 0; q < nls; q++) {
      SwapLS(q);
      fn(q);
      SwapLS(q);
    }
  }
  const std::vector<double>& PhiOf(int q) const;
  const std::vector<double>& PhiTmpOf(int q) const;
  int MatOf(int q) const;
  void UpdateMaterialIDByLevelSets();
  int ResolveConflictsInLevelSets();
  // MultiPhaseOperator
  void UpdateMaterialIDAtGhostNodes();
  int UpdateStateVariablesByRiemannSolutions();
  int UpdateStateVariablesByExtrapolation();
  // TimeIntegrator
  void AdvanceOneTimeStep(double time, double dt);
  void UpdateSolutionAfterTimeStepping();
  // Main.cpp:653-712
  int RunSteps(int nsteps);

 private:
  void SetupMesh();
  void SetupGhosts();
  void SetupFD3();
  int Riemann(const double* dir, int d, const double* Vm, int idm, const double* Vp, int idp, double* Vs, int& mid,
              double* Vsm, double* Vsp);

  bool LocalUpdate(int i, int j, int k, int id, bool upwind);
  void FixUnresolvedNodes(std::vector<std::array<int, 3>>& unresolved);
  // reinitialization helpers and work fields (ghost width 1; rNormal 3/cell)
  std::vector<double> rTag, rSign, rPhibk, rPhi0, rPhi1, rR, rNormal;
  double PhiOrZero(const std::vector<double>& phi, int i, int j, int k) const;
  bool TagFirstLayerNodes(std::vector<FirstLayerNode>& fl);
  void EvaluateSignFunction(double eps);
  void ReinitializeFirstLayerNodes(std::vector<FirstLayerNode>& fl);
  void ComputeNormalDirectionBeyondFirstLayer(std::vector<FirstLayerNode>& fl);
  double ComputeReinitResidual(const std::vector<double>& phi, double cfl);
  void ApplyCorrectionToFirstLayerNodes(std::vector<FirstLayerNode>& fl, double cfl);
  double MaxRelativeError();
  // narrow band
  std::vector<double> rPhiG2;  // PhiG2 of the narrow-band reinitializer (ghost width 1 here)
  int OutCount(int i, int j, int k) const {
    return (i < 0 || i >= m.NX) + (j < 0 || j >= m.NY) + (k < 0 || k >= m.NZ);
  }
  bool InGhosted(int i, int j, int k) const {
    return i >= -1 && i <= m.NX && j >= -1 && j <= m.NY && k >= -1 && k <= m.NZ;
  }
  // UsefulG2 has two ghost layers; the second is never set (zero)
  bool UsefulAt(int i, int j, int k) const { return InGhosted(i, j, k) && nbUseful[m.idx(i, j, k)] != 0; }
  void NbAlloc();
  void PropagateNarrowBand();
  void CutOffPhiOutsideBand();
  void CreateUsefulNodesPlusOneLayer();
  void UpdateNarrowBand(const std::vector<long>& first_layer_inc_ghost);
  void UpdatePhiMaxAndPhiMinInBand();
  bool TagFirstLayerNodesInBand(std::vector<FirstLayerNode>& fl, std::vector<long>& flg);
  void EvaluateSignFunctionInBand(double eps);
  void AXPlusBYInBandPlusOne(double a, std::vector<double>& X, double b, const std::vector<double>& Y,
                             bool work_on_ghost);
  double ComputeReinitResidualInBand(const std::vector<double>& phi, double cfl);
  double MaxRelativeErrorInBand();
  void ReinitializeInBand(int special_maxIts);
  // LevelSetReinitializer::ApplyBoundaryConditions (:1249-1360): with the
  // narrow band only the useful ghosts
  void ReinitApplyBoundaryConditions(std::vector<double>& phi);
  void BandNormal(const std::vector<double>& phi, int i, int j, int k, double* nv) const;
  void LSComputeResidualFDMNarrowBand(const std::vector<double>& Vin, const std::vector<double>& phi,
                                      std::vector<double>& res);
  // LevelSetOperator::AXPlusBY (LevelSetOperator.cpp:152-196): the interior,
  // or with the narrow band the useful nodes of the interior
  template <class F>
  void LSInterior(F fn) {
    if (nb_bw) {
      for (long c : useful_nodes) {
        int i, j, k;
        m.ijk(c, i, j, k);
        if (!OutCount(i, j, k)) fn(c);
      }
      return;
    }
    for (int k = 0; k < m.NZ; k++)
      for (int j = 0; j < m.NY; j++)
        for (int i = 0; i < m.NX; i++) fn(m.idx(i, j, k));
  }

 public:
  // Prescribed-velocity level-set tests (LevelSetOperator.cpp:1398-1490,
  // compiled in the reference with LEVELSET_TEST = 1, 2, 3): the flow
  // residual is zero, the velocity is prescribed before every level-set
  // residual, and the material IDs and states are not updated.
  int ls_test = 0;
  void PrescribeVelocityForTesting(std::vector<double>& Vio, const std::vector<double>& phi, double time);
};

}  // namespace orc
