// Oracle (test infrastructure only): several level sets on one subdomain.
// Level set 0 lives in the Solver's level-set members (Phi, Phi1, Phi_tmp,
// Rls, the LevelSet[0] fields of sch, the reinitializer's fields and the
// narrow band); level set q >= 1 is kept in lsx[q - 1] and exchanged with
// them (vector swaps) while its own LevelSetOperator / LevelSetReinitializer
// work runs, so every single-level-set routine serves each level set.  The
// routines that look at all of them restate
//   MultiPhaseOperator::UpdateMaterialIDByLevelSet (MultiPhaseOperator.cpp:153-357)
//   MultiPhaseOperator::ResolveConflictsInLevelSets (:1705-1990)
//   SpaceOperator::CalculateGradPhiAtCellInterface (SpaceOperator.cpp:2211-2268)
// and TimeIntegratorBase::UpdateSolutionAfterTimeStepping (TimeIntegrator.cpp:
// 720-800) for several level sets.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <set>
#include "solver.h"

namespace orc {

m2cx_levelset LSFromScheme(const m2cx_scheme& s) {
  m2cx_levelset L{};
  L.materialid = s.ls_materialid;
  for (int q = 0; q < 6; q++) L.bc[q] = s.ls_bc[q];
  L.solver = s.ls_solver;
  L.flux = s.ls_flux;
  L.delta = s.ls_delta;
  L.rec_type = s.ls_rec_type;
  L.limiter = s.ls_limiter;
  L.gmm_coeff = s.ls_gmm_coeff;
  L.reinit_frequency = s.reinit_frequency;
  L.reinit_maxits = s.reinit_maxits;
  L.reinit_frequency_dt = s.reinit_frequency_dt;
  L.reinit_cfl = s.reinit_cfl;
  L.reinit_tol = s.reinit_tol;
  L.reinit_first_layer = s.reinit_first_layer;
  L.bandwidth = s.ls_bandwidth;
  return L;
}

void SchemeSetLS(m2cx_scheme& s, const m2cx_levelset& L) {
  s.ls_materialid = L.materialid;
  for (int q = 0; q < 6; q++) s.ls_bc[q] = L.bc[q];
  s.ls_solver = L.solver;
  s.ls_flux = L.flux;
  s.ls_delta = L.delta;
  s.ls_rec_type = L.rec_type;
  s.ls_limiter = L.limiter;
  s.ls_gmm_coeff = L.gmm_coeff;
  s.reinit_frequency = L.reinit_frequency;
  s.reinit_maxits = L.reinit_maxits;
  s.reinit_frequency_dt = L.reinit_frequency_dt;
  s.reinit_cfl = L.reinit_cfl;
  s.reinit_tol = L.reinit_tol;
  s.reinit_first_layer = L.reinit_first_layer;
  s.ls_bandwidth = L.bandwidth;
}

// what a level set's settings decide: its ghosts' boundary conditions, the
// CoeffK of its own reconstruction, the reinitializer's CFL and the band
// (LevelSetOperator.cpp:59-110, LevelSetReinitializer.cpp:17-50)
void Solver::DeriveLevelSet() {
  for (auto& gp : ghosts) gp.lsbc = sch.ls_bc[gp.side];
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        const long c = m.idx(i, j, k);
        for (int a = 0; a < 3; a++) LSK3[c * 3 + a] = SlopeLimiterK(sch.ls_limiter, A3[c * 3 + a], B3[c * 3 + a]);
      }
  reinit_cfl = sch.reinit_cfl;
  nb_bw = (sch.n_levelsets > 0 && sch.ls_bandwidth > 0 && sch.ls_bandwidth < INT32_MAX) ? sch.ls_bandwidth : 0;
  if (nb_bw) {
    // LevelSetOperator.cpp:84-99
    if (nb_bw <= 1) Fail("in the narrow-band level set method, bandwidth should be at least 3");
    else if (sch.reinit_frequency <= 0 || sch.reinit_frequency >= nb_bw)
      Fail("the level set reinitialization frequency should be smaller than bandwidth");
  }
}

void Solver::SwapLS(int q) {
  if (q <= 0) return;
  LSState& S = lsx[q - 1];
  const m2cx_levelset mine = LSFromScheme(sch);
  SchemeSetLS(sch, S.set);
  S.set = mine;
  Phi.swap(S.Phi);
  Phi1.swap(S.Phi1);
  Phi_tmp.swap(S.Phi_tmp);
  Rls.swap(S.Rls);
  rTag.swap(S.rTag);
  rSign.swap(S.rSign);
  rPhibk.swap(S.rPhibk);
  rPhi0.swap(S.rPhi0);
  rPhi1.swap(S.rPhi1);
  rR.swap(S.rR);
  rNormal.swap(S.rNormal);
  rPhiG2.swap(S.rPhiG2);
  LSK3.swap(S.LSK3);
  nbLevel.swap(S.nbLevel);
  nbUseful.swap(S.nbUseful);
  nbActive.swap(S.nbActive);
  useful_nodes.swap(S.useful_nodes);
  active_nodes.swap(S.active_nodes);
  plus1_nodes.swap(S.plus1_nodes);
  fvPf.swap(S.fvPf);
  fvUf.swap(S.fvUf);
  fvSl.swap(S.fvSl);
  std::swap(nb_bw, S.nb_bw);
  std::swap(nb_built, S.nb_built);
  std::swap(phi_max, S.phi_max);
  std::swap(phi_min, S.phi_min);
  std::swap(phi_out_pos, S.phi_out_pos);
  std::swap(phi_out_neg, S.phi_out_neg);
  std::swap(reinit_cfl, S.reinit_cfl);
  for (auto& gp : ghosts) gp.lsbc = sch.ls_bc[gp.side];
}

// Main.cpp:285-330 checks, then each level set's operator set up
void Solver::SetLevelSets(const m2cx_levelset* ls, int n) {
  if (n < 0 || n > M2CX_MAX_LEVELSETS) {
    Fail("bad number of level sets");
    return;
  }
  std::set<int> seen;
  for (int q = 0; q < n; q++) {
    const int mid = ls[q].materialid;
    if (mid <= 0 || mid >= nmat) Fail("cannot initialize a level set for tracking this material");
    if (seen.count(mid)) Fail("cannot initialize multiple level sets for the same material");
    seen.insert(mid);
  }
  if (fatal) return;
  sch.n_levelsets = n;
  nls = n;
  lsx.clear();
  if (n == 0) return;
  SchemeSetLS(sch, ls[0]);
  DeriveLevelSet();
  const long nc = m.ncells();
  for (int q = 1; q < n; q++) {
    LSState S;
    S.set = ls[q];
    for (auto* f : {&S.Phi, &S.Phi1, &S.Phi_tmp, &S.Rls}) f->assign(nc, 0.0);
    S.LSK3.assign(nc * 3, 0);
    S.phi_max = -domain_diagonal;
    S.phi_min = domain_diagonal;
    lsx.push_back(std::move(S));
    SwapLS(q);
    DeriveLevelSet();
    SwapLS(q);
  }
}

const std::vector<double>& Solver::PhiOf(int q) const { return q == 0 ? Phi : lsx[q - 1].Phi; }
const std::vector<double>& Solver::PhiTmpOf(int q) const { return q == 0 ? Phi_tmp : lsx[q - 1].Phi_tmp; }
int Solver::MatOf(int q) const { return q == 0 ? sch.ls_materialid : lsx[q - 1].set.materialid; }

// UpdateMaterialIDByLevelSet (MultiPhaseOperator.cpp:153-357), no embedded
// surfaces: a node whose phi changed sign (or reached zero) for some level
// sets takes the material of the one it is now inside (two is an error);
// swept nodes inside none are given, in std::set<Int3> order and in
// iterations, the smallest material ID among their untagged neighbours not
// tracked by a level set; what is left is set to 0.
void Solver::UpdateMaterialIDByLevelSets() {
  const int L = std::max(nls, 1);
  std::vector<char> tag(m.ncells(), 0);
  std::set<std::array<int, 3>> remaining;  // (i, j, k) compares as Int3 does
  std::set<int> tracked;
  for (int q = 0; q < L; q++) tracked.insert(MatOf(q));
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        const long c = m.idx(i, j, k);
        tag[c] = 0;
        if (ID[c] == inactive_id) continue;
        std::set<std::pair<int, int>> swept;
        for (int q = 0; q < L; q++) {
          const double a = PhiOf(q)[c], b = PhiTmpOf(q)[c];
          if (a * b > 0.0) continue;
          swept.insert({q, a < 0 ? -1 : (a == 0 ? 0 : 1)});
        }
        if (swept.empty()) continue;
        int myls = -1;
        for (auto& st : swept) {
          if (st.second != -1) continue;
          if (myls != -1) {
            Fail("a node belongs to two material subdomains");
            return;
          }
          myls = st.first;
        }
        if (myls != -1) {
          ID[c] = MatOf(myls);
        } else {
          tag[c] = 1;
          remaining.insert({i, j, k});
        }
      }
  for (int iter = 0; iter < 100; iter++) {
    if (remaining.empty()) break;
    int progress = 0;
    std::set<std::array<int, 3>> tmp = remaining;
    for (auto& ijk : tmp) {
      const int i = ijk[0], j = ijk[1], k = ijk[2];
      std::set<int> cand;
      for (int nk = k - 1; nk <= k + 1; nk++)
        for (int nj = j - 1; nj <= j + 1; nj++)
          for (int ni = i - 1; ni <= i + 1; ni++) {
            if (!m.inside(ni, nj, nk)) continue;
            const long nc = m.idx(ni, nj, nk);
            if (tag[nc] == 1) continue;
            const int nid = ID[nc];
            if (nid == inactive_id || tracked.count(nid)) continue;
            cand.insert(nid);
          }
      if (cand.empty()) continue;
      const long c = m.idx(i, j, k);
      ID[c] = *cand.begin();
      tag[c] = 0;
      remaining.erase(ijk);
      progress++;
    }
    if (progress == 0) {
      for (auto& ijk : tmp) ID[m.idx(ijk[0], ijk[1], ijk[2])] = 0;
      break;
    }
  }
  UpdateMaterialIDAtGhostNodes();
}

// ResolveConflictsInLevelSets (MultiPhaseOperator.cpp:1705-1990).
// Part I, over the ghosted brick in (k, j, i) order: a node next to the
// boundary of two or more subdomains and inside more than one of them goes
// to the one with the largest |phi| (the first on ties); the others get
// +mean|phi| of the owners, the new owner -mean|phi|.  With the optional
// corrections (LevelSetCorrectionFrequency, at the steps it divides) a node
// next to two or more boundaries inside one subdomain, or on the boundary
// (phi = 0) of some, has all of them set to +-mean|phi| over those level
// sets, uncounted.  Part II (optional only), over the interior in (k, j, i)
// order: a background node (phi >= 0 for every level set) with at most one
// background face neighbour inside the domain goes to the level set with the
// smallest phi, which becomes -phi.  Nodes are updated in place, so later
// nodes see earlier results.  Returns the number of nodes resolved (part I's
// multi-owner nodes and part II's nodes).
int Solver::ResolveConflictsInLevelSets() {
  if (nls == 0) return 0;
  const bool opt = sch.mp_ls_correction_frequency > 0 && time_step % sch.mp_ls_correction_frequency == 0;
  std::vector<std::vector<double>*> phi(nls);
  phi[0] = &Phi;
  for (int q = 1; q < nls; q++) phi[q] = &lsx[q - 1].Phi;
  int resolved = 0;
  std::vector<int> boundaries, owner, inter;
  if (nls >= 2) {
    for (int k = -1; k <= m.NZ; k++)
      for (int j = -1; j <= m.NY; j++)
        for (int i = -1; i <= m.NX; i++) {
          const long c = m.idx(i, j, k);
          boundaries.clear();
          for (int q = 0; q < nls; q++) {
            const std::vector<double>& P = *phi[q];
            const double p = P[c];
            if ((i - 1 >= -1 && p * P[m.idx(i - 1, j, k)] <= 0) || (i + 1 <= m.NX && p * P[m.idx(i + 1, j, k)] <= 0) ||
                (j - 1 >= -1 && p * P[m.idx(i, j - 1, k)] <= 0) || (j + 1 <= m.NY && p * P[m.idx(i, j + 1, k)] <= 0) ||
                (k - 1 >= -1 && p * P[m.idx(i, j, k - 1)] <= 0) || (k + 1 <= m.NZ && p * P[m.idx(i, j, k + 1)] <= 0))
              boundaries.push_back(q);
          }
          if (boundaries.size() <= 1) continue;
          owner.clear();
          inter.clear();
          for (int q : boundaries) {
            const double p = (*phi[q])[c];
            if (p < 0) owner.push_back(q);
            else if (p == 0) inter.push_back(q);
          }
          if (owner.empty() && inter.empty()) continue;
          if (owner.size() <= 1) {
            // one owner, or none with some on their boundary: optional only
            if (!opt) continue;
            double new_phi = 0.0;
            for (int q : boundaries) new_phi += std::fabs((*phi[q])[c]);
            new_phi /= boundaries.size();
            for (int q : boundaries) {
              const bool own = owner.empty() ? q == inter[0] : (*phi[q])[c] < 0;
              (*phi[q])[c] = own ? -new_phi : new_phi;
            }
            continue;
          }
          int new_owner = owner[0];
          double max_phi = std::fabs((*phi[owner[0]])[c]);
          for (size_t w = 1; w < owner.size(); w++)
            if (std::fabs((*phi[owner[w]])[c]) > max_phi) {
              new_owner = owner[w];
              max_phi = std::fabs((*phi[owner[w]])[c]);
            }
          double new_phi = 0.0;
          for (int q : owner) new_phi += std::fabs((*phi[q])[c]);
          new_phi /= owner.size();
          for (int q : owner) (*phi[q])[c] = (q == new_owner) ? -new_phi : new_phi;
          resolved++;
        }
  }
  if (!opt) return resolved;
  // part II: isolated background nodes of the interior
  auto background = [&](long c) {
    for (int q = 0; q < nls; q++)
      if ((*phi[q])[c] < 0) return false;
    return true;
  };
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        const long c = m.idx(i, j, k);
        if (!background(c)) continue;
        int qi = 0;
        const bool nb[6] = {i - 1 >= 0, i + 1 < m.NX, j - 1 >= 0, j + 1 < m.NY, k - 1 >= 0, k + 1 < m.NZ};
        const long nc[6] = {m.idx(i - 1, j, k), m.idx(i + 1, j, k), m.idx(i, j - 1, k),
                            m.idx(i, j + 1, k), m.idx(i, j, k - 1), m.idx(i, j, k + 1)};
        for (int f = 0; f < 6 && qi < 2; f++)
          if (nb[f] && background(nc[f])) qi++;
        if (qi >= 2) continue;
        double min_phi = DBL_MAX;
        int new_owner = -1;
        for (int q = 0; q < nls; q++)
          if ((*phi[q])[c] < min_phi) {
            new_owner = q;
            min_phi = (*phi[q])[c];
          }
        if (new_owner < 0) continue;
        (*phi[new_owner])[c] = -min_phi;
        resolved++;
      }
  return resolved;
}

}  // namespace orc
