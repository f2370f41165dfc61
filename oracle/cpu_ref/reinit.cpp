// Oracle (test infrastructure only): level-set reinitialization over the full
// domain, a CPU restatement of LevelSetReinitializer (LevelSetReinitializer.cpp)
// on one subdomain, plus the prescribed-velocity level-set tests.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include "solver.h"

namespace orc {

// The reference's Phi never writes its edge and corner ghosts (outside the
// domain in two or more directions): they keep the zero of VecSet
// (SpaceVariable.cpp:203-207).  The tagging loop reads them as such.
double Solver::PhiOrZero(const std::vector<double>& phi, int i, int j, int k) const {
  int out = (i < 0 || i >= m.NX) + (j < 0 || j >= m.NY) + (k < 0 || k >= m.NZ);
  return out >= 2 ? 0.0 : phi[m.idx(i, j, k)];
}

// LevelSetOperator::Reinitialize (LevelSetOperator.cpp:1338-1363) with
// isTimeToWrite (Utils.cpp:337-357) deciding when.
bool Solver::Reinitialize(double time, double dt, int time_step, int special_maxIts, bool must_do) {
  if (must_do && !ReinitEnabled()) {
    Fail("reinitialization of the level set is requested, but not specified");
    return false;
  }
  if (!ReinitEnabled()) return false;
  bool due;
  if (must_do) due = time > -100.0 + 0.1 * dt;
  else if (sch.reinit_frequency_dt > 0)
    due = std::floor(time / sch.reinit_frequency_dt) != std::floor((time - dt) / sch.reinit_frequency_dt);
  else
    due = sch.reinit_frequency > 0 && time_step % sch.reinit_frequency == 0;
  if (!due) return false;
  if (nb_bw)
    ReinitializeInBand(special_maxIts);
  else
    ReinitializeFullDomain(special_maxIts);
  return true;
}

// TagFirstLayerNodes (LevelSetReinitializer.cpp:304-436): tags over the
// ghosted domain except edges and corners; the first layer (interior nodes
// with a sign change or a zero product towards a face neighbour) with its
// set S of neighbours strictly across the interface.
bool Solver::TagFirstLayerNodes(std::vector<FirstLayerNode>& fl) {
  fl.clear();
  const std::vector<double>& phi = Phi;
  for (int k = -1; k <= m.NZ; k++)
    for (int j = -1; j <= m.NY; j++)
      for (int i = -1; i <= m.NX; i++) {
        long c = m.idx(i, j, k);
        rTag[c] = 0.0;
        int out = (i < 0 || i >= m.NX) + (j < 0 || j >= m.NY) + (k < 0 || k >= m.NZ);
        if (out >= 2) continue;
        const double p = phi[c];
        bool hit = (i - 1 >= -1 && p * PhiOrZero(phi, i - 1, j, k) <= 0) ||
                   (i + 1 <= m.NX && p * PhiOrZero(phi, i + 1, j, k) <= 0) ||
                   (j - 1 >= -1 && p * PhiOrZero(phi, i, j - 1, k) <= 0) ||
                   (j + 1 <= m.NY && p * PhiOrZero(phi, i, j + 1, k) <= 0) ||
                   (k - 1 >= -1 && p * PhiOrZero(phi, i, j, k - 1) <= 0) ||
                   (k + 1 <= m.NZ && p * PhiOrZero(phi, i, j, k + 1) <= 0);
        if (!hit) continue;
        rTag[c] = 1.0;
        if (out == 0) {
          FirstLayerNode n{};
          n.i = i; n.j = j; n.k = k;
          fl.push_back(n);
        }
      }
  for (auto& n : fl) {
    const int i = n.i, j = n.j, k = n.k;
    const double p = phi[m.idx(i, j, k)];
    const long nb[6] = {m.idx(i - 1, j, k), m.idx(i + 1, j, k), m.idx(i, j - 1, k),
                        m.idx(i, j + 1, k), m.idx(i, j, k - 1), m.idx(i, j, k + 1)};
    double sum = 0.0;
    for (int q = 0; q < 6; q++) {
      if (p * phi[nb[q]] < 0) {
        n.s[q] = true;
        n.r[q] = p / phi[nb[q]];
        n.ns++;
        sum += phi[nb[q]];
      }
    }
    n.r0 = (sum != 0) ? p / sum : 0.0;
  }
  return !fl.empty();
}

// EvaluateSignFunctionFullDomain (:601-622), eps = 1
void Solver::EvaluateSignFunction(double eps) {
  for (int k = -1; k <= m.NZ; k++)
    for (int j = -1; j <= m.NY; j++)
      for (int i = -1; i <= m.NX; i++) {
        long c = m.idx(i, j, k);
        double factor = eps * std::min(m.DX(i), std::min(m.DY(j), m.DZ(k)));
        rSign[c] = Phi[c] / std::sqrt(Phi[c] * Phi[c] + factor * factor);
      }
}

// CentralDifferenceLocal (LevelSetReinitializer.h:118-123)
static inline double central_diff(double phi0, double phi1, double phi2, double x0, double x1, double x2) {
  double c0 = -(x2 - x1) / ((x1 - x0) * (x2 - x0));
  double c1 = 1.0 / (x1 - x0) - 1.0 / (x2 - x1);
  double c2 = (x1 - x0) / ((x2 - x0) * (x2 - x1));
  return c0 * phi0 + c1 * phi1 + c2 * phi2;
}

// DifferentiateInFirstLayer (:906-955)
static double differentiate_first_layer(double x0, double x1, double x2, double tag0, double tag2, double phi0,
                                        double phi1, double phi2, double phi00, double phi3, double eps) {
  bool u0 = tag0 != 0, u2 = tag2 != 0;
  if (!u0 && !u2) return 0.0;
  double d0 = phi1 - phi0, d1 = phi2 - phi1;
  bool condB = d0 * d1 < 0 || phi0 * phi00 < 0 || phi2 * phi3 < 0;
  if (condB) {
    if (u0 && (phi0 * phi2 < 0) && std::fabs(d0 + eps) < std::fabs(d1)) u0 = false;
    if (u2 && (phi0 * phi2 < 0) && std::fabs(d1 + eps) < std::fabs(d0)) u2 = false;
  }
  if (!u0 && !u2) return 0.0;
  if (u0) {
    if (u2) return central_diff(phi0, phi1, phi2, x0, x1, x2);
    return (phi1 - phi0) / (x1 - x0);
  }
  return (phi2 - phi1) / (x2 - x1);
}

// ReinitializeFirstLayerNodes (:650-873): phi = phi/|grad phi| on the first
// layer with one-sided or central differences chosen by the tags (Hartmann et
// al. 2008, Eq. 21a), then the CR-1 / CR-2 averaging on the nodes whose
// curvature points away from the interface.  The copy of phi with two ghost
// layers (PhiG2) only ever holds the first ghost layer; its second layer
// stays zero, which the reference reads in y at the top boundary (its
// `j+2<=NY` test has j+2 on both branches).  PopulatePhiG2 (:877-902) copies
// every node over the full domain, only the useful nodes with the narrow band
// (PhiG2 keeps its older values elsewhere).
void Solver::ReinitializeFirstLayerNodes(std::vector<FirstLayerNode>& fl) {
  if (nb_bw)
    for (long c : useful_nodes) rPhiG2[c] = rPhi1[c];
  const std::vector<double>& phig = nb_bw ? rPhiG2 : rPhi1;  // copy of Phi including ghosts
  auto G = [&](int i, int j, int k) {
    if (i < -1 || i > m.NX || j < -1 || j > m.NY || k < -1 || k > m.NZ) return 0.0;
    return phig[m.idx(i, j, k)];
  };
  const int NX = m.NX, NY = m.NY, NZ = m.NZ;
  for (auto& n : fl) {
    const int i = n.i, j = n.j, k = n.k;
    const long c = m.idx(i, j, k);
    const double epsx = 1.0e-3 * m.DX(i), epsy = 1.0e-3 * m.DY(j), epsz = 1.0e-3 * m.DZ(k);
    double g[3];
    g[0] = differentiate_first_layer(m.X(i - 1), m.X(i), m.X(i + 1), rTag[m.idx(i - 1, j, k)], rTag[m.idx(i + 1, j, k)],
                                     G(i - 1, j, k), G(i, j, k), G(i + 1, j, k),
                                     (i - 2 >= -1) ? G(i - 2, j, k) : G(i - 1, j, k),
                                     (i + 2 <= NX) ? G(i + 2, j, k) : G(i + 1, j, k), epsx);
    g[1] = differentiate_first_layer(m.Y(j - 1), m.Y(j), m.Y(j + 1), rTag[m.idx(i, j - 1, k)], rTag[m.idx(i, j + 1, k)],
                                     G(i, j - 1, k), G(i, j, k), G(i, j + 1, k),
                                     (j - 2 >= -1) ? G(i, j - 2, k) : G(i, j - 1, k), G(i, j + 2, k), epsy);
    (void)NY;
    g[2] = differentiate_first_layer(m.Z(k - 1), m.Z(k), m.Z(k + 1), rTag[m.idx(i, j, k - 1)], rTag[m.idx(i, j, k + 1)],
                                     G(i, j, k - 1), G(i, j, k), G(i, j, k + 1),
                                     (k - 2 >= -1) ? G(i, j, k - 2) : G(i, j, k - 1),
                                     (k + 2 <= NZ) ? G(i, j, k + 2) : G(i, j, k + 1), epsz);
    double norm = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    for (int d = 0; d < 3; d++) n.nphi0[d] = g[d] / norm;
    if (norm == 0)
      Phi[c] = phig[c];
    else
      Phi[c] = phig[c] / norm;
  }
  const int flt = sch.reinit_first_layer;
  if (flt != M2CX_REINIT_CONSTRAINED1 && flt != M2CX_REINIT_CONSTRAINED2) return;

  ComputeNormalDirectionBeyondFirstLayer(fl);
  auto Nrm = [&](int i, int j, int k, int d) { return rNormal[m.idx(i, j, k) * 3 + d]; };
  // step 1: the correction set, Eq. (14) of Hartmann et al. (2010)
  for (auto& n : fl) {
    const int i = n.i, j = n.j, k = n.k;
    double kx, ky, kz;
    if (i - 1 >= 0 && i + 1 < NX)
      kx = central_diff(Nrm(i - 1, j, k, 0), Nrm(i, j, k, 0), Nrm(i + 1, j, k, 0), m.X(i - 1), m.X(i), m.X(i + 1));
    else if (i - 1 >= 0)
      kx = (Nrm(i, j, k, 0) - Nrm(i - 1, j, k, 0)) / (m.X(i) - m.X(i - 1));
    else if (i + 1 < NX)
      kx = (Nrm(i + 1, j, k, 0) - Nrm(i, j, k, 0)) / (m.X(i + 1) - m.X(i));
    else
      kx = 0.0;
    if (j - 1 >= 0 && j + 1 < NY)
      ky = central_diff(Nrm(i, j - 1, k, 1), Nrm(i, j, k, 1), Nrm(i, j + 1, k, 1), m.Y(j - 1), m.Y(j), m.Y(j + 1));
    else if (j - 1 >= 0)
      ky = (Nrm(i, j, k, 1) - Nrm(i, j - 1, k, 1)) / (m.Y(j) - m.Y(j - 1));
    else if (j + 1 < NY)
      ky = (Nrm(i, j + 1, k, 1) - Nrm(i, j, k, 1)) / (m.Y(j + 1) - m.Y(j));
    else
      ky = 0.0;
    if (k - 1 >= 0 && k + 1 < NZ)
      kz = central_diff(Nrm(i, j, k - 1, 2), Nrm(i, j, k, 2), Nrm(i, j, k + 1, 2), m.Z(k - 1), m.Z(k), m.Z(k + 1));
    else if (k - 1 >= 0)
      kz = (Nrm(i, j, k, 2) - Nrm(i, j, k - 1, 2)) / (m.Z(k) - m.Z(k - 1));
    else if (k + 1 < NZ)
      kz = (Nrm(i, j, k + 1, 2) - Nrm(i, j, k, 2)) / (m.Z(k + 1) - m.Z(k));
    else
      kz = 0.0;
    double curv = kx + ky + kz;
    double pg = phig[m.idx(i, j, k)];
    n.f = (curv * pg < 0 || (curv == 0.0 && pg < 0)) ? 0.0 : DBL_MAX;
  }
  // step 2: the correction
  for (auto& n : fl) {
    if (n.f != 0.0) continue;
    const int i = n.i, j = n.j, k = n.k;
    const long nb[6] = {m.idx(i - 1, j, k), m.idx(i + 1, j, k), m.idx(i, j - 1, k),
                        m.idx(i, j + 1, k), m.idx(i, j, k - 1), m.idx(i, j, k + 1)};
    const double pg = phig[m.idx(i, j, k)];
    if (flt == M2CX_REINIT_CONSTRAINED1) {
      double sum = 0.0;
      for (int q = 0; q < 6; q++)
        if (n.s[q]) sum += Phi[nb[q]] / phig[nb[q]];
      n.f = (n.ns != 0) ? pg * sum / (double)n.ns : 0.0;
    } else {
      double s1 = 0.0, s2 = 0.0;
      for (int q = 0; q < 6; q++)
        if (n.s[q]) { s1 += Phi[nb[q]]; s2 += phig[nb[q]]; }
      n.f = (s2 != 0) ? pg * s1 / s2 : 0.0;
    }
  }
  // step 3: apply
  for (auto& n : fl) {
    if (n.f == DBL_MAX) continue;
    Phi[m.idx(n.i, n.j, n.k)] = n.f;
    n.f = 0.0;
  }
}

// The unit normal at one interior node of the narrow band (the narrow-band
// branches of ComputeNormalDirectionBeyondFirstLayer, :1155-1208, and of
// LevelSetOperator::ComputeNormalDirectionCentralDifferencing_NarrowBand,
// LevelSetOperator.cpp:1545-1592): central differences where both
// neighbours are useful -- in y the reference tests the upper neighbour twice,
// so the lower one may be outside the band -- one-sided where one of them is,
// zero where neither is; divided by its norm when that is nonzero.
void Solver::BandNormal(const std::vector<double>& phi, int i, int j, int k, double* nv) const {
  const long c = m.idx(i, j, k);
  const double p = phi[c];
  auto P = [&](int a, int b, int e) { return phi[m.idx(a, b, e)]; };
  if (UsefulAt(i + 1, j, k) && UsefulAt(i - 1, j, k))
    nv[0] = central_diff(P(i - 1, j, k), p, P(i + 1, j, k), m.X(i - 1), m.X(i), m.X(i + 1));
  else if (UsefulAt(i + 1, j, k))
    nv[0] = (P(i + 1, j, k) - p) / (m.X(i + 1) - m.X(i));
  else if (UsefulAt(i - 1, j, k))
    nv[0] = (p - P(i - 1, j, k)) / (m.X(i) - m.X(i - 1));
  else
    nv[0] = 0.0;
  if (UsefulAt(i, j + 1, k) && UsefulAt(i, j + 1, k))
    nv[1] = central_diff(P(i, j - 1, k), p, P(i, j + 1, k), m.Y(j - 1), m.Y(j), m.Y(j + 1));
  else if (UsefulAt(i, j + 1, k))
    nv[1] = (P(i, j + 1, k) - p) / (m.Y(j + 1) - m.Y(j));
  else if (UsefulAt(i, j - 1, k))
    nv[1] = (p - P(i, j - 1, k)) / (m.Y(j) - m.Y(j - 1));
  else
    nv[1] = 0.0;
  if (UsefulAt(i, j, k + 1) && UsefulAt(i, j, k - 1))
    nv[2] = central_diff(P(i, j, k - 1), p, P(i, j, k + 1), m.Z(k - 1), m.Z(k), m.Z(k + 1));
  else if (UsefulAt(i, j, k + 1))
    nv[2] = (P(i, j, k + 1) - p) / (m.Z(k + 1) - m.Z(k));
  else if (UsefulAt(i, j, k - 1))
    nv[2] = (p - P(i, j, k - 1)) / (m.Z(k) - m.Z(k - 1));
  else
    nv[2] = 0.0;
  const double nr = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  if (nr != 0.0) {
    const double inv = 1.0 / nr;  // Vec3D::operator/= multiplies by the reciprocal
    nv[0] *= inv; nv[1] *= inv; nv[2] *= inv;
  }
}

// ComputeNormalDirectionBeyondFirstLayer (:1136-1245)
void Solver::ComputeNormalDirectionBeyondFirstLayer(std::vector<FirstLayerNode>& fl) {
  for (auto& n : fl)
    for (int d = 0; d < 3; d++) rNormal[m.idx(n.i, n.j, n.k) * 3 + d] = n.nphi0[d];
  if (nb_bw) {
    for (long c : useful_nodes) {
      int i, j, k;
      m.ijk(c, i, j, k);
      if (OutCount(i, j, k) || rTag[c] != 0) continue;
      BandNormal(Phi, i, j, k, &rNormal[c * 3]);
    }
    return;
  }
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        if (rTag[c] != 0) continue;
        double* nv = &rNormal[c * 3];
        nv[0] = central_diff(Phi[m.idx(i - 1, j, k)], Phi[c], Phi[m.idx(i + 1, j, k)], m.X(i - 1), m.X(i), m.X(i + 1));
        nv[1] = central_diff(Phi[m.idx(i, j - 1, k)], Phi[c], Phi[m.idx(i, j + 1, k)], m.Y(j - 1), m.Y(j), m.Y(j + 1));
        nv[2] = central_diff(Phi[m.idx(i, j, k - 1)], Phi[c], Phi[m.idx(i, j, k + 1)], m.Z(k - 1), m.Z(k), m.Z(k + 1));
        double nr = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        if (nr != 0.0) {
          double inv = 1.0 / nr;  // Vec3D::operator/= multiplies by the reciprocal
          nv[0] *= inv; nv[1] *= inv; nv[2] *= inv;
        }
      }
}

// ComputeResidualFullDomain (:960-1041): Godunov upwind |grad phi| - 1 times
// the smoothed sign, pseudo-time step cfl*min(dx,dy,dz); first-layer nodes
// fixed for the Fixed / CR-1 / CR-2 treatments.  Returns the largest
// |local residual| over nodes off the first layer.
double Solver::ComputeReinitResidual(const std::vector<double>& phi, double cfl) {
  const int flt = sch.reinit_first_layer;
  const bool fix = flt == M2CX_REINIT_FIXED || flt == M2CX_REINIT_CONSTRAINED1 || flt == M2CX_REINIT_CONSTRAINED2;
  double maxres = 0.0;
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        if (fix && rTag[c] != 0) {
          rR[c] = 0.0;
          continue;
        }
        double dx = std::min(m.DX(i), std::min(m.DY(j), m.DZ(k)));
        double dt = cfl * dx;
        double p = phi[c];
        double a = (p - phi[m.idx(i - 1, j, k)]) / (m.X(i) - m.X(i - 1));
        double b = (phi[m.idx(i + 1, j, k)] - p) / (m.X(i + 1) - m.X(i));
        double cc = (p - phi[m.idx(i, j - 1, k)]) / (m.Y(j) - m.Y(j - 1));
        double d = (phi[m.idx(i, j + 1, k)] - p) / (m.Y(j + 1) - m.Y(j));
        double e = (p - phi[m.idx(i, j, k - 1)]) / (m.Z(k) - m.Z(k - 1));
        double f = (phi[m.idx(i, j, k + 1)] - p) / (m.Z(k + 1) - m.Z(k));
        double ap = std::max(a, 0.0), am = std::min(a, 0.0);
        double bp = std::max(b, 0.0), bm = std::min(b, 0.0);
        double cp = std::max(cc, 0.0), cm = std::min(cc, 0.0);
        double dp = std::max(d, 0.0), dm = std::min(d, 0.0);
        double ep = std::max(e, 0.0), em = std::min(e, 0.0);
        double fp = std::max(f, 0.0), fm = std::min(f, 0.0);
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

// ApplyCorrectionToFirstLayerNodes (:1364-1455): HCR-1 / HCR-2 forcing of the
// first layer towards the interface position of the original phi (Hartmann
// et al. 2010), then the boundary conditions.
void Solver::ApplyCorrectionToFirstLayerNodes(std::vector<FirstLayerNode>& fl, double cfl) {
  const int flt = sch.reinit_first_layer;
  if (flt != M2CX_REINIT_ITERATIVE_CONSTRAINED1 && flt != M2CX_REINIT_ITERATIVE_CONSTRAINED2) return;
  for (auto& n : fl) {
    const int i = n.i, j = n.j, k = n.k;
    const long c = m.idx(i, j, k);
    const long nb[6] = {m.idx(i - 1, j, k), m.idx(i + 1, j, k), m.idx(i, j - 1, k),
                        m.idx(i, j + 1, k), m.idx(i, j, k - 1), m.idx(i, j, k + 1)};
    n.f = 0.0;
    bool out = false;
    for (int q = 0; q < 6; q++)
      if (n.s[q] && Phi[c] * Phi[nb[q]] >= 0) { out = true; break; }
    if (out) continue;
    double sum = 0.0;
    if (flt == M2CX_REINIT_ITERATIVE_CONSTRAINED1) {
      for (int q = 0; q < 6; q++)
        if (n.s[q]) sum += n.r[q] * Phi[nb[q]];
      n.f = (n.ns != 0) ? sum / (double)n.ns - Phi[c] : 0.0 - Phi[c];
    } else {
      for (int q = 0; q < 6; q++)
        if (n.s[q]) sum += Phi[nb[q]];
      n.f = n.r0 * sum - Phi[c];
    }
    n.f /= std::max(m.DX(i), std::max(m.DY(j), m.DZ(k)));
  }
  const double beta = 0.5;
  for (auto& n : fl) {
    const int i = n.i, j = n.j, k = n.k;
    double dt = cfl * std::min(m.DX(i), std::min(m.DY(j), m.DZ(k)));
    Phi[m.idx(i, j, k)] += dt * beta * n.f;
  }
  ReinitApplyBoundaryConditions(Phi);
}

// CalculateMaximumRelativeErrorFullDomain (:1966-1994)
double Solver::MaxRelativeError() {
  double err = 0.0;
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        double e = std::fabs(Phi[c] - rPhi0[c]);
        if (Phi[c] != 0) e /= std::fabs(Phi[c]);
        err = std::max(err, e);
      }
  return err;
}

// ReinitializeFullDomain (LevelSetReinitializer.cpp:76-176)
void Solver::ReinitializeFullDomain(int special_maxIts) {
  const long n = m.ncells();
  if (rTag.size() != (size_t)n) {
    for (auto* f : {&rTag, &rSign, &rPhibk, &rPhi0, &rPhi1, &rR}) f->assign(n, 0.0);
    rNormal.assign(n * 3, 0.0);
  }
  std::vector<FirstLayerNode> fl;
  if (!TagFirstLayerNodes(fl)) return;
  cnt.reinit_calls++;
  EvaluateSignFunction(1.0);
  const int flt = sch.reinit_first_layer;
  if (flt == M2CX_REINIT_CONSTRAINED1 || flt == M2CX_REINIT_CONSTRAINED2) {
    for (long c = 0; c < n; c++) rPhi1[c] = 0.0 * rPhi1[c] + 1.0 * Phi[c];
    ReinitializeFirstLayerNodes(fl);
    LSApplyBoundaryConditions(Phi);
  }
  rPhibk = Phi;
  const double tol = sch.reinit_tol;
  auto interior = [&](auto fn) {
    for (int k = 0; k < m.NZ; k++)
      for (int j = 0; j < m.NY; j++)
        for (int i = 0; i < m.NX; i++) fn(m.idx(i, j, k));
  };
  for (;;) {  // RETRY_FullDomain
    double residual = 0.0, dphi_max = 0.0;
    const int maxIts = (special_maxIts > 0) ? special_maxIts : sch.reinit_maxits;
    const double cfl = reinit_cfl;
    int iter;
    for (iter = 0; iter < maxIts; iter++) {
      interior([&](long c) { rPhi0[c] = 0.0 * rPhi0[c] + 1.0 * Phi[c]; });
      residual = ComputeReinitResidual(Phi, cfl);
      if (residual < tol) break;
      interior([&](long c) { rPhi1[c] = 0.0 * rPhi1[c] + 1.0 * Phi[c]; });
      interior([&](long c) { rPhi1[c] = 1.0 * rPhi1[c] + 1.0 * rR[c]; });
      LSApplyBoundaryConditions(rPhi1);
      ComputeReinitResidual(rPhi1, cfl);
      interior([&](long c) { rPhi1[c] = 0.25 * rPhi1[c] + 0.75 * Phi[c]; });
      interior([&](long c) { rPhi1[c] = 1.0 * rPhi1[c] + 0.25 * rR[c]; });
      LSApplyBoundaryConditions(rPhi1);
      ComputeReinitResidual(rPhi1, cfl);
      interior([&](long c) { Phi[c] = 1.0 / 3.0 * Phi[c] + 2.0 / 3.0 * rPhi1[c]; });
      interior([&](long c) { Phi[c] = 1.0 * Phi[c] + 2.0 / 3.0 * rR[c]; });
      LSApplyBoundaryConditions(Phi);
      ApplyCorrectionToFirstLayerNodes(fl, cfl);
      cnt.reinit_iters++;
      dphi_max = MaxRelativeError();
      if (dphi_max < tol) break;
    }
    if (iter < maxIts) return;
    // failsafe (:155-173)
    if (dphi_max < 0.02) return;
    if (reinit_cfl < 0.02) {
      Fail("level-set reinitialization failed to converge");
      return;
    }
    reinit_cfl /= 1.5;
    Phi = rPhibk;
  }
}

// LevelSetOperator::PrescribeVelocityFieldForTesting (LevelSetOperator.cpp:
// 1398-1490) with the normals of ComputeNormalDirectionCentralDifferencing_
// FullDomain (:1509-1540) for test 3.
void Solver::PrescribeVelocityForTesting(std::vector<double>& Vio, const std::vector<double>& phi, double time) {
  const double pi = 2.0 * std::acos(0);
  if (ls_test == 1 || ls_test == 2) {
    for (int k = -1; k <= m.NZ; k++)
      for (int j = -1; j <= m.NY; j++)
        for (int i = -1; i <= m.NX; i++) {
          double* v = &Vio[m.idx(i, j, k) * 5];
          const double x = m.X(i), y = m.Y(j);
          if (ls_test == 1) {
            double v1 = std::fabs(y - 50.0) < 1.0e-10 ? 1.0e-10 : 50.0 - y;
            double v2 = std::fabs(x - 50.0) < 1.0e-10 ? 1.0e-10 : x - 50.0;
            v[1] = pi / 314.0 * v1;
            v[2] = pi / 314.0 * v2;
            v[3] = 0.0;
          } else {
            const double T = 8.0;
            v[1] = 2.0 * std::pow(std::sin(pi * x), 2) * std::sin(pi * y) * std::cos(pi * y) * std::cos(pi * time / T);
            v[2] = -2.0 * std::sin(pi * x) * std::cos(pi * x) * std::pow(std::sin(pi * y), 2) * std::cos(pi * time / T);
            v[3] = 0.0;
          }
        }
  } else if (ls_test == 3 && nb_bw) {
    // the useful nodes only; the others keep their velocity (:1468-1483)
    const double s = time < 1.0 ? 1.0 : -1.0;
    for (long c : useful_nodes) {
      int i, j, k;
      m.ijk(c, i, j, k);
      double* v = &Vio[c * 5];
      if (OutCount(i, j, k)) {
        v[1] = v[2] = v[3] = 0.0;
        continue;
      }
      double nv[3];
      BandNormal(phi, i, j, k, nv);
      v[1] = s * nv[0];
      v[2] = s * nv[1];
      v[3] = 0.0;
    }
  } else if (ls_test == 3) {
    const double s = time < 1.0 ? 1.0 : -1.0;
    for (int k = -1; k <= m.NZ; k++)
      for (int j = -1; j <= m.NY; j++)
        for (int i = -1; i <= m.NX; i++) {
          double* v = &Vio[m.idx(i, j, k) * 5];
          if (!m.inside(i, j, k)) {
            v[1] = v[2] = v[3] = 0.0;
            continue;
          }
          long c = m.idx(i, j, k);
          double nv[3];
          nv[0] = central_diff(phi[m.idx(i - 1, j, k)], phi[c], phi[m.idx(i + 1, j, k)], m.X(i - 1), m.X(i), m.X(i + 1));
          nv[1] = central_diff(phi[m.idx(i, j - 1, k)], phi[c], phi[m.idx(i, j + 1, k)], m.Y(j - 1), m.Y(j), m.Y(j + 1));
          nv[2] = central_diff(phi[m.idx(i, j, k - 1)], phi[c], phi[m.idx(i, j, k + 1)], m.Z(k - 1), m.Z(k), m.Z(k + 1));
          double nr = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
          if (nr != 0.0) {
            double inv = 1.0 / nr;
            nv[0] *= inv; nv[1] *= inv; nv[2] *= inv;
          }
          v[1] = s * nv[0];
          v[2] = s * nv[1];
          v[3] = 0.0;
        }
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// Level-set finite-volume scheme (LevelSetOperator::ComputeResidualFVM,
// LevelSetOperator.cpp:966-1300) with its own reconstruction settings
// (LevelSet[0].Reconstruction) and flux (LevelSet[0].Flux).
// ---------------------------------------------------------------------------
namespace orc {

static inline double ls_sigma(const m2cx_scheme& s, double A, double B, int K, double dq0, double dq1) {
  if (s.ls_rec_type != M2CX_REC_LINEAR) return 0.0;
  if (s.ls_limiter == M2CX_LIM_GENERALIZED_MINMOD) return GeneralizedMinMod(A, B, s.ls_gmm_coeff, dq0, dq1);
  if (s.ls_limiter == M2CX_LIM_VANALBADA) return VanAlbada(A, B, K, dq0, dq1);
  return 0.5 * (dq0 + dq1);
}

// ComputeLocalAdvectionFlux (:1229-1262)
static double ls_face_flux(const m2cx_scheme& s, double phim, double phip, double um, double up) {
  double flux = 0.0;
  if (s.ls_flux == M2CX_LSFLUX_UPWIND) {
    double vmid = 0.5 * (um + up);
    if (vmid > 0) flux = phim * um;
    else if (vmid < 0) flux = phip * up;
    else flux = 0.5 * (phim * um + phip * up);
  } else if (s.ls_flux == M2CX_LSFLUX_LLF) {
    flux = 0.5 * (phim * um + phip * up) - 0.5 * std::max(std::fabs(um), std::fabs(up)) * (phip - phim);
  } else {  // Roe with Harten's entropy fix
    double lam = std::fabs(0.5 * (um + up));
    double tol = std::max(std::fabs(um), std::fabs(up)) * s.ls_delta;
    if (lam < tol) lam = (lam * lam + tol * tol) / (2 * tol);
    flux = 0.5 * (phim * um + phip * up) - 0.5 * lam * (phip - phim);
  }
  return flux;
}

void Solver::LSComputeResidualFVM(const std::vector<double>& Vin, const std::vector<double>& phi,
                                  std::vector<double>& res) {
  const long n = m.ncells();
  // face values: [6][n] for phi (l r b t k f), [6][n] for the velocity
  // component of each axis (minus, plus per axis), slopes [3][n]
  std::vector<double> pf(6 * n, 0.0), uf(6 * n, 0.0), sl(3 * n, 0.0);
  // Reconstructor::ReconstructIn1D (Reconstructor.cpp:922-1054) for u, v, w
  // and Reconstruct (:335-790) for phi, interior cells
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        const long c = m.idx(i, j, k);
        const long st[3] = {1, m.PX, (long)m.PX * m.PY};
        const double dd[3] = {m.DX(i), m.DY(j), m.DZ(k)};
        for (int a = 0; a < 3; a++) {
          const double A = A3[c * 3 + a], B = B3[c * 3 + a];
          const int K = LSK3[c * 3 + a];
          double u0 = Vin[(c - st[a]) * 5 + 1 + a], u1 = Vin[c * 5 + 1 + a], u2 = Vin[(c + st[a]) * 5 + 1 + a];
          double sig = ls_sigma(sch, A, B, K, u1 - u0, u2 - u1);
          uf[(2 * a) * n + c] = u1 - 0.5 * sig;
          uf[(2 * a + 1) * n + c] = u1 + 0.5 * sig;
          sl[a * n + c] = sig / dd[a];
          double p0 = phi[c - st[a]], p1 = phi[c], p2 = phi[c + st[a]];
          double sp = ls_sigma(sch, A, B, K, p1 - p0, p2 - p1);
          pf[(2 * a) * n + c] = p1 - 0.5 * sp;
          pf[(2 * a + 1) * n + c] = p1 + 0.5 * sp;
        }
      }
  // ghost faces by the mesh boundary conditions (ReconstructIn1D :1048-1140,
  // Reconstruct :807-880; one dof: symmetry flips the velocity only)
  for (auto& gp : ghosts) {
    const long c = m.idx(gp.i, gp.j, gp.k), im = m.idx(gp.ii, gp.jj, gp.kk);
    const int a = gp.side / 2;
    const bool lo = (gp.side % 2) == 0;
    // the ghost's face towards the domain: plus face of a low ghost, minus
    // face of a high one; the image's face on the boundary: its minus face
    // for a low ghost, its plus face for a high one
    const int gface = lo ? 2 * a + 1 : 2 * a, iface = lo ? 2 * a : 2 * a + 1;
    switch (gp.bc) {
      case M2CX_BC_INLET: case M2CX_BC_INLET2: case M2CX_BC_OVERSET:
        uf[gface * n + c] = Vin[c * 5 + 1 + a];
        pf[gface * n + c] = phi[c];
        for (int e = 0; e < 3; e++) sl[e * n + c] = 0.0;
        break;
      case M2CX_BC_OUTLET:
        uf[gface * n + c] = uf[iface * n + im];
        pf[gface * n + c] = pf[iface * n + im];
        for (int e = 0; e < 3; e++) sl[e * n + c] = sl[e * n + im];
        break;
      default:  // symmetry, slip and stick walls
        uf[gface * n + c] = -uf[iface * n + im];
        pf[gface * n + c] = pf[iface * n + im];
        for (int e = 0; e < 3; e++) sl[e * n + c] = -sl[e * n + im];
        break;
    }
  }
  // ComputeAdvectionFlux (:1079-1150): fluxes through the minus faces of
  // interior cells and of the upper ghost layers, accumulated in loop order
  std::fill(res.begin(), res.end(), 0.0);
  for (int k = 0; k <= m.NZ; k++)
    for (int j = 0; j <= m.NY; j++)
      for (int i = 0; i <= m.NX; i++) {
        const long c = m.idx(i, j, k);
        if (k != m.NZ && j != m.NY) {
          const long cm = m.idx(i - 1, j, k);
          double f = ls_face_flux(sch, pf[1 * n + cm], pf[0 * n + c], uf[1 * n + cm], uf[0 * n + c]) / m.DX(i);
          res[cm] -= f;
          res[c] += f;
        }
        if (k != m.NZ && i != m.NX) {
          const long cm = m.idx(i, j - 1, k);
          double f = ls_face_flux(sch, pf[3 * n + cm], pf[2 * n + c], uf[3 * n + cm], uf[2 * n + c]) / m.DY(j);
          res[cm] -= f;
          res[c] += f;
        }
        if (j != m.NY && i != m.NX) {
          const long cm = m.idx(i, j, k - 1);
          double f = ls_face_flux(sch, pf[5 * n + cm], pf[4 * n + c], uf[5 * n + cm], uf[4 * n + c]) / m.DZ(k);
          res[cm] -= f;
          res[c] += f;
        }
      }
  // AddSourceTerm (:1266-1285)
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        const long c = m.idx(i, j, k);
        res[c] += phi[c] * (sl[0 * n + c] + sl[1 * n + c] + sl[2 * n + c]);
      }
}

// LevelSetOperator::ComputeResidualFVM in a narrow band (LevelSetOperator.cpp:
// 966-1062, 1160-1232, 1288-1306): ReconstructInBand reconstructs the
// velocity components and phi at the ACTIVE nodes only (Reconstructor::
// ReconstructIn1D and Reconstruct with Active as the selection, :922-1054,
// 335-790), face ghosts by the boundary conditions where active; the
// outputs live in buffers kept between calls, so a flux through a face of a
// useful node that is not active reads what the buffers held (zero before
// the first write, else that node's values from the last call in which it
// was active).  The residual is zeroed on the useful nodes and the fluxes
// through the minus faces of each useful node are added in the order of the
// useful-node list (ComputeAdvectionFluxInBand, :1160-1232), which fixes the
// summation order of each node's terms; the source term on the active nodes
// inside the domain.
void Solver::LSComputeResidualFVMNarrowBand(const std::vector<double>& Vin, const std::vector<double>& phi,
                                            std::vector<double>& res) {
  const long n = m.ncells();
  if ((long)fvPf.size() != 6 * n) {
    fvPf.assign(6 * n, 0.0);
    fvUf.assign(6 * n, 0.0);
    fvSl.assign(3 * n, 0.0);
  }
  std::vector<double>&pf = fvPf, &uf = fvUf, &sl = fvSl;
  // the scalar field of ReconstructInBand holds the velocity component on the
  // useful nodes; the active nodes' stencils lie on useful nodes, so it is
  // read here from V itself
  for (long c : active_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (i < 0 || i >= m.NX || j < 0 || j >= m.NY || k < 0 || k >= m.NZ) continue;
    const long st[3] = {1, m.PX, (long)m.PX * m.PY};
    const double dd[3] = {m.DX(i), m.DY(j), m.DZ(k)};
    for (int a = 0; a < 3; a++) {
      const double A = A3[c * 3 + a], B = B3[c * 3 + a];
      const int K = LSK3[c * 3 + a];
      double u0 = Vin[(c - st[a]) * 5 + 1 + a], u1 = Vin[c * 5 + 1 + a], u2 = Vin[(c + st[a]) * 5 + 1 + a];
      double sig = ls_sigma(sch, A, B, K, u1 - u0, u2 - u1);
      uf[(2 * a) * n + c] = u1 - 0.5 * sig;
      uf[(2 * a + 1) * n + c] = u1 + 0.5 * sig;
      sl[a * n + c] = sig / dd[a];
      double p0 = phi[c - st[a]], p1 = phi[c], p2 = phi[c + st[a]];
      double sp = ls_sigma(sch, A, B, K, p1 - p0, p2 - p1);
      pf[(2 * a) * n + c] = p1 - 0.5 * sp;
      pf[(2 * a + 1) * n + c] = p1 + 0.5 * sp;
    }
  }
  for (auto& gp : ghosts) {
    const long c = m.idx(gp.i, gp.j, gp.k), im = m.idx(gp.ii, gp.jj, gp.kk);
    if (!nbActive[c]) continue;
    const int a = gp.side / 2;
    const bool lo = (gp.side % 2) == 0;
    const int gface = lo ? 2 * a + 1 : 2 * a, iface = lo ? 2 * a : 2 * a + 1;
    switch (gp.bc) {
      case M2CX_BC_INLET: case M2CX_BC_INLET2: case M2CX_BC_OVERSET:
        uf[gface * n + c] = Vin[c * 5 + 1 + a];
        pf[gface * n + c] = phi[c];
        for (int e = 0; e < 3; e++) sl[e * n + c] = 0.0;
        break;
      case M2CX_BC_OUTLET:
        uf[gface * n + c] = uf[iface * n + im];
        pf[gface * n + c] = pf[iface * n + im];
        for (int e = 0; e < 3; e++) sl[e * n + c] = sl[e * n + im];
        break;
      default:  // symmetry, slip and stick walls
        uf[gface * n + c] = -uf[iface * n + im];
        pf[gface * n + c] = pf[iface * n + im];
        for (int e = 0; e < 3; e++) sl[e * n + c] = -sl[e * n + im];
        break;
    }
  }
  for (long c : useful_nodes) res[c] = 0.0;
  for (long c : useful_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (i == -1 || j == -1 || k == -1) continue;
    if (k != m.NZ && j != m.NY) {
      const long cm = m.idx(i - 1, j, k);
      double f = ls_face_flux(sch, pf[1 * n + cm], pf[0 * n + c], uf[1 * n + cm], uf[0 * n + c]) / m.DX(i);
      if (nbActive[cm]) res[cm] -= f;
      if (nbActive[c]) res[c] += f;
    }
    if (k != m.NZ && i != m.NX) {
      const long cm = m.idx(i, j - 1, k);
      double f = ls_face_flux(sch, pf[3 * n + cm], pf[2 * n + c], uf[3 * n + cm], uf[2 * n + c]) / m.DY(j);
      if (nbActive[cm]) res[cm] -= f;
      if (nbActive[c]) res[c] += f;
    }
    if (j != m.NY && i != m.NX) {
      const long cm = m.idx(i, j, k - 1);
      double f = ls_face_flux(sch, pf[5 * n + cm], pf[4 * n + c], uf[5 * n + cm], uf[4 * n + c]) / m.DZ(k);
      if (nbActive[cm]) res[cm] -= f;
      if (nbActive[c]) res[c] += f;
    }
  }
  for (long c : active_nodes) {
    int i, j, k;
    m.ijk(c, i, j, k);
    if (i < 0 || i >= m.NX || j < 0 || j >= m.NY || k < 0 || k >= m.NZ) continue;
    res[c] += phi[c] * (sl[0 * n + c] + sl[1 * n + c] + sl[2 * n + c]);
  }
}

// MultiPhaseOperator::UpdateStateVariablesByExtrapolation (MultiPhaseOperator.
// cpp:1155-1272): a cell that changed material takes the weighted average of
// its neighbours (3x3x3) of the same material that did not change, each
// weighted by how much its flow points at the cell (Upwind) or by its inverse
// distance (PhaseChangeDirection = All, :1220-1221).
int Solver::UpdateStateVariablesByExtrapolation() {
  std::vector<std::array<int, 3>> unresolved;
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        const long c = m.idx(i, j, k);
        if (ID[c] == IDn[c]) continue;
        if (ID[c] == inactive_id) continue;
        cnt.id_changed++;
        const double x0[3] = {m.X(i), m.Y(j), m.Z(k)};
        double sum_weight = 0.0, vsum[5] = {0, 0, 0, 0, 0};
        for (int nk = k - 1; nk <= k + 1; nk++)
          for (int nj = j - 1; nj <= j + 1; nj++)
            for (int ni = i - 1; ni <= i + 1; ni++) {
              const long q = m.idx(ni, nj, nk);
              if (ID[q] != ID[c]) continue;
              if (ID[q] != IDn[q]) continue;
              if (!m.inside(ni, nj, nk)) continue;
              const double* vn = &V[q * 5];
              double weight;
              if (sch.mp_phase_dir == M2CX_PHASEDIR_ALL) {
                double d[3] = {x0[0] - m.X(ni), x0[1] - m.Y(nj), x0[2] - m.Z(nk)};
                weight = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
              } else {
                double v1[3] = {vn[1], vn[2], vn[3]};
                double v1n = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
                if (v1n != 0) {
                  double inv = 1.0 / v1n;
                  v1[0] *= inv; v1[1] *= inv; v1[2] *= inv;
                }
                double d[3] = {x0[0] - m.X(ni), x0[1] - m.Y(nj), x0[2] - m.Z(nk)};
                double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double inv = 1.0 / dn;
                d[0] *= inv; d[1] *= inv; d[2] *= inv;
                weight = std::max(0.0, d[0] * v1[0] + d[1] * v1[1] + d[2] * v1[2]);
              }
              if (weight > 0) {
                sum_weight += weight;
                for (int p = 0; p < 5; p++) vsum[p] += weight * vn[p];
              }
            }
        if (sum_weight == 0)
          unresolved.push_back({k, j, i});
        else
          for (int p = 0; p < 5; p++) V[c * 5 + p] = vsum[p] / sum_weight;
      }
  if (!unresolved.empty()) {
    cnt.unresolved += (long)unresolved.size();
    FixUnresolvedNodes(unresolved);
  }
  return 0;
}

}  // namespace orc
