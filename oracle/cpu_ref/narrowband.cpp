// Oracle (test infrastructure only): the narrow-band level set on one
// subdomain, a CPU restatement of
//   LevelSetReinitializer.cpp:181-300   ReinitializeInBand
//                            :440-598   TagFirstLayerNodesInBand
//                            :627-647   EvaluateSignFunctionInBand
//                            :1045-1133 ComputeResidualInBand
//                            :1459-1933 UpdatePhiMaxAndPhiMinInBand, ConstructNarrowBand,
//                                       PropagateNarrowBand, CutOffPhiOutsideBand,
//                                       UpdateNarrowBand, CreateUsefulNodesPlusOneLayer
//                            :1937-2031 AXPlusBYInBandPlusOne, CalculateMaximumRelativeErrorInBand
//   LevelSetOperator.cpp:880-955        ComputeResidualFDM_NarrowBand
// The reinitializer's work fields (Tag, Sign, R, Phi0, Phi1, Phibk, PhiG2,
// NormalDir) persist between calls and each step writes only the nodes the
// reference writes, so values left from earlier calls are read where the
// reference reads them.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <set>
#include "solver.h"

namespace orc {

static const int kLargeInt = 31415927;  // large_int (LevelSetReinitializer.cpp:14)

// the fields a narrow-band run keeps from its start (zero, as VecSet leaves
// a new SpaceVariable3D; Tag is set to 0 at LevelSetReinitializer.cpp:37)
void Solver::NbAlloc() {
  const long n = m.ncells();
  if ((long)nbLevel.size() == n) return;
  nbLevel.assign(n, kLargeInt);
  nbUseful.assign(n, 0);
  nbActive.assign(n, 0);
  for (auto* f : {&rTag, &rSign, &rPhibk, &rPhi0, &rPhi1, &rR, &rPhiG2}) f->assign(n, 0.0);
  rNormal.assign(n * 3, 0.0);
}

// ConstructNarrowBand (:1479-1592): levels 0 (phi = 0) and 1 (a sign change
// or zero product towards a populated face neighbour) over the ghosted
// domain, edges and corners excluded, then the higher levels, the cut-off of
// phi outside the band and the list of useful nodes plus one layer.
void Solver::ConstructNarrowBand() {
  NbAlloc();
  useful_nodes.clear();
  active_nodes.clear();
  auto pop = [&](int i, int j, int k) { return OutCount(i, j, k) < 2; };
  for (int k = -1; k <= m.NZ; k++)
    for (int j = -1; j <= m.NY; j++)
      for (int i = -1; i <= m.NX; i++) {
        const long c = m.idx(i, j, k);
        nbLevel[c] = kLargeInt;
        nbUseful[c] = 0;
        nbActive[c] = 0;
        if (!pop(i, j, k)) continue;
        const double p = Phi[c];
        int lev = -1;
        if (p == 0) {
          lev = 0;
        } else if ((i - 1 >= -1 && pop(i - 1, j, k) && p * Phi[m.idx(i - 1, j, k)] <= 0) ||
                   (i + 1 <= m.NX && pop(i + 1, j, k) && p * Phi[m.idx(i + 1, j, k)] <= 0) ||
                   (j - 1 >= -1 && pop(i, j - 1, k) && p * Phi[m.idx(i, j - 1, k)] <= 0) ||
                   (j + 1 <= m.NY && pop(i, j + 1, k) && p * Phi[m.idx(i, j + 1, k)] <= 0) ||
                   (k - 1 >= -1 && pop(i, j, k - 1) && p * Phi[m.idx(i, j, k - 1)] <= 0) ||
                   (k + 1 <= m.NZ && pop(i, j, k + 1) && p * Phi[m.idx(i, j, k + 1)] <= 0)) {
          lev = 1;
        }
        if (lev < 0) continue;
        nbLevel[c] = lev;
        nbUseful[c] = 1;
        nbActive[c] = 1;
        useful_nodes.push_back(c);
        active_nodes.push_back(c);
      }
  PropagateNarrowBand();
  CutOffPhiOutsideBand();
  CreateUsefulNodesPlusOneLayer();
  nb_built = true;
}

// PropagateNarrowBand (:1596-1728): level b (2..bandwidth) for the unlevelled
// populated face neighbours of the nodes levelled before; levels below the
// bandwidth are active.
void Solver::PropagateNarrowBand() {
  for (int band = 2; band <= nb_bw; band++) {
    const size_t size = useful_nodes.size();
    for (size_t q = 0; q < size; q++) {
      int i, j, k;
      m.ijk(useful_nodes[q], i, j, k);
      const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
      for (int d = 0; d < 6; d++) {
        const int a = nb[d][0], b = nb[d][1], e = nb[d][2];
        if (!InGhosted(a, b, e) || OutCount(a, b, e) >= 2) continue;
        const long c = m.idx(a, b, e);
        if (nbLevel[c] != kLargeInt) continue;
        nbLevel[c] = band;
        nbUseful[c] = 1;
        useful_nodes.push_back(c);
        if (band < nb_bw) {
          nbActive[c] = 1;
          active_nodes.push_back(c);
        }
      }
    }
  }
}

// CutOffPhiOutsideBand (:1732-1773)
void Solver::CutOffPhiOutsideBand() {
  phi_max = -domain_diagonal;
  phi_min = domain_diagonal;
  for (long c : useful_nodes) {
    phi_max = std::max(Phi[c], phi_max);
    phi_min = std::min(Phi[c], phi_min);
  }
  phi_out_pos = phi_max * 10.0;
  phi_out_neg = phi_min * 10.0;
  if (phi_out_pos <= 0) phi_out_pos = domain_diagonal;  // no interface
  if (phi_out_neg >= 0) phi_out_neg = -domain_diagonal;
  const long n = m.ncells();
  for (long c = 0; c < n; c++)
    if (!nbUseful[c]) Phi[c] = Phi[c] >= 0 ? phi_out_pos : phi_out_neg;
}

// CreateUsefulNodesPlusOneLayer (:1890-1933): the useful nodes and their
// populated face neighbours that are not useful
void Solver::CreateUsefulNodesPlusOneLayer() {
  std::set<long> extra;
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    for (int d = 0; d < 6; d++) {
      const int a = nb[d][0], b = nb[d][1], e = nb[d][2];
      if (!InGhosted(a, b, e) || OutCount(a, b, e) >= 2) continue;
      const long q = m.idx(a, b, e);
      if (!nbUseful[q]) extra.insert(q);
    }
  }
  plus1_nodes = useful_nodes;
  plus1_nodes.insert(plus1_nodes.end(), extra.begin(), extra.end());
}

// UpdatePhiMaxAndPhiMinInBand (:1459-1476): the running bounds over the
// useful nodes inside the domain
void Solver::UpdatePhiMaxAndPhiMinInBand() {
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (OutCount(i, j, k)) continue;
    phi_max = std::max(phi_max, Phi[c]);
    phi_min = std::min(phi_min, Phi[c]);
  }
}

// TagFirstLayerNodesInBand (:440-598): tags of the useful nodes, the first
// layer inside the domain with its set S, and the tagged nodes including
// ghosts (the seeds of the new band)
bool Solver::TagFirstLayerNodesInBand(std::vector<FirstLayerNode>& fl, std::vector<long>& flg) {
  fl.clear();
  flg.clear();
  auto pop = [&](int i, int j, int k) { return OutCount(i, j, k) < 2; };
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    rTag[c] = 0;
    if (!pop(i, j, k)) continue;
    const double p = Phi[c];
    const bool hit = (i - 1 >= -1 && pop(i - 1, j, k) && p * Phi[m.idx(i - 1, j, k)] <= 0) ||
                     (i + 1 <= m.NX && pop(i + 1, j, k) && p * Phi[m.idx(i + 1, j, k)] <= 0) ||
                     (j - 1 >= -1 && pop(i, j - 1, k) && p * Phi[m.idx(i, j - 1, k)] <= 0) ||
                     (j + 1 <= m.NY && pop(i, j + 1, k) && p * Phi[m.idx(i, j + 1, k)] <= 0) ||
                     (k - 1 >= -1 && pop(i, j, k - 1) && p * Phi[m.idx(i, j, k - 1)] <= 0) ||
                     (k + 1 <= m.NZ && pop(i, j, k + 1) && p * Phi[m.idx(i, j, k + 1)] <= 0);
    if (!hit) continue;
    rTag[c] = 1;
    flg.push_back(c);
    if (!OutCount(i, j, k)) {
      FirstLayerNode nd{};
      nd.i = i; nd.j = j; nd.k = k;
      fl.push_back(nd);
    }
  }
  for (auto& nd : fl) {
    const int i = nd.i, j = nd.j, k = nd.k;
    const double p = Phi[m.idx(i, j, k)];
    const long nb[6] = {m.idx(i - 1, j, k), m.idx(i + 1, j, k), m.idx(i, j - 1, k),
                        m.idx(i, j + 1, k), m.idx(i, j, k - 1), m.idx(i, j, k + 1)};
    double sum = 0.0;
    for (int q = 0; q < 6; q++) {
      if (p * Phi[nb[q]] < 0) {
        nd.s[q] = true;
        nd.r[q] = p / Phi[nb[q]];
        nd.ns++;
        sum += Phi[nb[q]];
      }
    }
    nd.r0 = (sum != 0) ? p / sum : 0.0;
  }
  return !fl.empty();
}

// UpdateNarrowBand (:1777-1886)
void Solver::UpdateNarrowBand(const std::vector<long>& flg) {
  const std::vector<long> backup = useful_nodes;
  for (long c : useful_nodes) {
    nbLevel[c] = kLargeInt;
    nbUseful[c] = 0;
    nbActive[c] = 0;
  }
  useful_nodes.clear();
  active_nodes.clear();
  for (long c : flg) {
    nbUseful[c] = 1;
    useful_nodes.push_back(c);
    nbActive[c] = 1;
    active_nodes.push_back(c);
    nbLevel[c] = (Phi[c] == 0) ? 0 : 1;
  }
  PropagateNarrowBand();
  // nodes that left the band: phi and the residual cut off
  for (long c : backup)
    if (!nbUseful[c]) {
      Phi[c] = (Phi[c] >= 0) ? phi_out_pos : phi_out_neg;
      rR[c] = 0.0;
    }
  // nodes that just entered it: the running bounds instead of the cut-off
  for (long c : useful_nodes) {
    if (Phi[c] > 0.9 * phi_out_pos)
      Phi[c] = phi_max * 1.1;
    else if (Phi[c] < 0.9 * phi_out_neg)
      Phi[c] = phi_min * 1.1;
  }
  CreateUsefulNodesPlusOneLayer();
}

// EvaluateSignFunctionInBand (:627-647)
void Solver::EvaluateSignFunctionInBand(double eps) {
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    const double factor = eps * std::min(m.DX(i), std::min(m.DY(j), m.DZ(k)));
    rSign[c] = Phi[c] / std::sqrt(Phi[c] * Phi[c] + factor * factor);
  }
}

// AXPlusBYInBandPlusOne (:1937-1962)
void Solver::AXPlusBYInBandPlusOne(double a, std::vector<double>& X, double b, const std::vector<double>& Y,
                                   bool work_on_ghost) {
  for (long c : plus1_nodes) {
    if (!work_on_ghost) {
      int i, j, k;
      m.ijk(c, i, j, k);
      if (OutCount(i, j, k)) continue;
    }
    X[c] = a * X[c] + b * Y[c];
  }
}

// ComputeResidualInBand (:1045-1133): the full-domain residual with the
// differences towards nodes outside the band set to zero
double Solver::ComputeReinitResidualInBand(const std::vector<double>& phi, double cfl) {
  const int flt = sch.reinit_first_layer;
  const bool fix = flt == M2CX_REINIT_FIXED || flt == M2CX_REINIT_CONSTRAINED1 || flt == M2CX_REINIT_CONSTRAINED2;
  double maxres = 0.0;
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (OutCount(i, j, k)) continue;
    if (fix && rTag[c] != 0) {
      rR[c] = 0.0;
      continue;
    }
    const double dx = std::min(m.DX(i), std::min(m.DY(j), m.DZ(k)));
    const double dt = cfl * dx;
    const double p = phi[c];
    const double a = UsefulAt(i - 1, j, k) ? (p - phi[m.idx(i - 1, j, k)]) / (m.X(i) - m.X(i - 1)) : 0.0;
    const double b = UsefulAt(i + 1, j, k) ? (phi[m.idx(i + 1, j, k)] - p) / (m.X(i + 1) - m.X(i)) : 0.0;
    const double cc = UsefulAt(i, j - 1, k) ? (p - phi[m.idx(i, j - 1, k)]) / (m.Y(j) - m.Y(j - 1)) : 0.0;
    const double d = UsefulAt(i, j + 1, k) ? (phi[m.idx(i, j + 1, k)] - p) / (m.Y(j + 1) - m.Y(j)) : 0.0;
    const double e = UsefulAt(i, j, k - 1) ? (p - phi[m.idx(i, j, k - 1)]) / (m.Z(k) - m.Z(k - 1)) : 0.0;
    const double f = UsefulAt(i, j, k + 1) ? (phi[m.idx(i, j, k + 1)] - p) / (m.Z(k + 1) - m.Z(k)) : 0.0;
    const double ap = std::max(a, 0.0), am = std::min(a, 0.0);
    const double bp = std::max(b, 0.0), bm = std::min(b, 0.0);
    const double cp = std::max(cc, 0.0), cm = std::min(cc, 0.0);
    const double dp = std::max(d, 0.0), dm = std::min(d, 0.0);
    const double ep = std::max(e, 0.0), em = std::min(e, 0.0);
    const double fp = std::max(f, 0.0), fm = std::min(f, 0.0);
    double lr;
    if (p >= 0)
      lr = std::sqrt(std::max(ap * ap, bm * bm) + std::max(cp * cp, dm * dm) + std::max(ep * ep, fm * fm)) - 1.0;
    else
      lr = std::sqrt(std::max(am * am, bp * bp) + std::max(cm * cm, dp * dp) + std::max(em * em, fp * fp)) - 1.0;
    if (std::fabs(lr) <= 1.0)
      rR[c] = -dt * rSign[c] * lr;
    else
      rR[c] = -dt * rSign[c] * lr / std::fabs(lr);
    if (rTag[c] == 0) maxres = std::max(maxres, std::fabs(lr));
  }
  return maxres;
}

// CalculateMaximumRelativeErrorInBand (:1998-2031)
double Solver::MaxRelativeErrorInBand() {
  double err = 0.0;
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (OutCount(i, j, k)) continue;
    double e = std::fabs(Phi[c] - rPhi0[c]);
    if (Phi[c] != 0) e /= std::fabs(Phi[c]);
    err = std::max(err, e);
  }
  return err;
}

void Solver::ReinitApplyBoundaryConditions(std::vector<double>& phi) {
  LSApplyBoundaryConditions(phi, nb_bw ? &nbUseful : nullptr);
}

// ReinitializeInBand (:181-300)
void Solver::ReinitializeInBand(int special_maxIts) {
  if (!nb_built) {
    Fail("the narrow band of the level set has not been constructed");
    return;
  }
  UpdatePhiMaxAndPhiMinInBand();
  std::vector<FirstLayerNode> fl;
  std::vector<long> flg;
  if (!TagFirstLayerNodesInBand(fl, flg)) return;
  cnt.reinit_calls++;
  UpdateNarrowBand(flg);
  EvaluateSignFunctionInBand(1.0);
  const int flt = sch.reinit_first_layer;
  if (flt == M2CX_REINIT_CONSTRAINED1 || flt == M2CX_REINIT_CONSTRAINED2) {
    AXPlusBYInBandPlusOne(0.0, rPhi1, 1.0, Phi, true);
    ReinitializeFirstLayerNodes(fl);
    ReinitApplyBoundaryConditions(Phi);
  }
  AXPlusBYInBandPlusOne(0.0, rPhibk, 1.0, Phi, false);
  const double tol = sch.reinit_tol;
  for (;;) {  // RETRY_NarrowBand
    double residual = 0.0, dphi_max = 0.0;
    const int maxIts = (special_maxIts > 0) ? special_maxIts : sch.reinit_maxits;
    const double cfl = reinit_cfl;
    int iter;
    for (iter = 0; iter < maxIts; iter++) {
      AXPlusBYInBandPlusOne(0.0, rPhi0, 1.0, Phi, false);
      residual = ComputeReinitResidualInBand(Phi, cfl);
      if (residual < tol) break;
      AXPlusBYInBandPlusOne(0.0, rPhi1, 1.0, Phi, false);
      AXPlusBYInBandPlusOne(1.0, rPhi1, 1.0, rR, false);
      ReinitApplyBoundaryConditions(rPhi1);
      ComputeReinitResidualInBand(rPhi1, cfl);
      AXPlusBYInBandPlusOne(0.25, rPhi1, 0.75, Phi, false);
      AXPlusBYInBandPlusOne(1.0, rPhi1, 0.25, rR, false);
      ReinitApplyBoundaryConditions(rPhi1);
      ComputeReinitResidualInBand(rPhi1, cfl);
      AXPlusBYInBandPlusOne(1.0 / 3.0, Phi, 2.0 / 3.0, rPhi1, false);
      AXPlusBYInBandPlusOne(1.0, Phi, 2.0 / 3.0, rR, false);
      ReinitApplyBoundaryConditions(Phi);
      ApplyCorrectionToFirstLayerNodes(fl, cfl);
      cnt.reinit_iters++;
      dphi_max = MaxRelativeErrorInBand();
      if (dphi_max < tol) break;
    }
    if (iter < maxIts) return;
    // failsafe (:274-296)
    if (dphi_max < 0.02) return;
    if (reinit_cfl < 0.02) {
      Fail("level-set reinitialization failed to converge");
      return;
    }
    reinit_cfl /= 1.5;
    AXPlusBYInBandPlusOne(0.0, Phi, 1.0, rPhibk, false);
  }
}

// ComputeResidualFDM_NarrowBand (LevelSetOperator.cpp:880-955): the residual
// cleared on the useful nodes and computed on the active ones inside the
// domain, with FD3 derivatives where the node two away is useful and
// one-sided differences otherwise
void Solver::LSComputeResidualFDMNarrowBand(const std::vector<double>& Vin, const std::vector<double>& phi,
                                            std::vector<double>& res) {
  auto P = [&](int i, int j, int k) { return phi[m.idx(i, j, k)]; };
  auto fd = [&](const std::vector<double>& C, long c, double v0, double v1, double v2, double v3) {
    const double* cc = &C[c * 4];
    return cc[0] * v0 + cc[1] * v1 + cc[2] * v2 + cc[3] * v3;
  };
  for (long c : useful_nodes) res[c] = 0.0;
  for (long c : active_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (OutCount(i, j, k)) continue;
    const double* v = &Vin[c * 5];
    double dm, dp;
    dm = UsefulAt(i - 2, j, k) ? fd(cxL, c, P(i - 2, j, k), P(i - 1, j, k), P(i, j, k), P(i + 1, j, k))
                               : (P(i, j, k) - P(i - 1, j, k)) / (m.X(i) - m.X(i - 1));
    dp = UsefulAt(i + 2, j, k) ? fd(cxR, c, P(i - 1, j, k), P(i, j, k), P(i + 1, j, k), P(i + 2, j, k))
                               : (P(i + 1, j, k) - P(i, j, k)) / (m.X(i + 1) - m.X(i));
    double r = v[1] >= 0 ? dm * v[1] : dp * v[1];
    dm = UsefulAt(i, j - 2, k) ? fd(cyL, c, P(i, j - 2, k), P(i, j - 1, k), P(i, j, k), P(i, j + 1, k))
                               : (P(i, j, k) - P(i, j - 1, k)) / (m.Y(j) - m.Y(j - 1));
    dp = UsefulAt(i, j + 2, k) ? fd(cyR, c, P(i, j - 1, k), P(i, j, k), P(i, j + 1, k), P(i, j + 2, k))
                               : (P(i, j + 1, k) - P(i, j, k)) / (m.Y(j + 1) - m.Y(j));
    r += v[2] >= 0 ? dm * v[2] : dp * v[2];
    dm = UsefulAt(i, j, k - 2) ? fd(czL, c, P(i, j, k - 2), P(i, j, k - 1), P(i, j, k), P(i, j, k + 1))
                               : (P(i, j, k) - P(i, j, k - 1)) / (m.Z(k) - m.Z(k - 1));
    dp = UsefulAt(i, j, k + 2) ? fd(czR, c, P(i, j, k - 1), P(i, j, k), P(i, j, k + 1), P(i, j, k + 2))
                               : (P(i, j, k + 1) - P(i, j, k)) / (m.Z(k + 1) - m.Z(k));
    r += v[3] >= 0 ? dm * v[3] : dp * v[3];
    res[c] = r * -1.0;
  }
}

}  // namespace orc
