// Oracle (test infrastructure only): see solver.h.
#include "solver.h"
#include <cmath>
#include <cstdio>
#include <set>

namespace orc {

static inline double* at5(std::vector<double>& f, long c) { return &f[c * 5]; }
static inline const double* at5(const std::vector<double>& f, long c) { return &f[c * 5]; }

// Integer power for the van Albada limiter (Reconstructor.h:105-111 calls
// pow(dq,k)); k = 2 on uniform meshes, where dq*dq is the correctly rounded
// value of pow(dq,2).
static inline double ipow(double x, int k) {
  double r = x;
  for (int i = 1; i < k; i++) r *= x;
  return r;
}

// Reconstructor.h:96-103
double GeneralizedMinMod(double A, double B, double alpha, double dq0, double dq1) {
  if (dq0 < 0.0 && dq1 < 0.0) return std::max(alpha * dq0, std::max(B / (A + 1.0) * (dq0 + dq1), alpha * dq1));
  else if (dq0 > 0.0 && dq1 > 0.0) return std::min(alpha * dq0, std::min(B / (A + 1.0) * (dq0 + dq1), alpha * dq1));
  return 0.0;
}
// Reconstructor.h:104-111
double VanAlbada(double A, double B, int k, double dq0, double dq1) {
  if (dq0 * dq1 <= 0) return 0.0;
  if (std::fabs(dq0 * dq1) < 1e-16) return 0.0;
  double a = ipow(dq0, k), b = ipow(dq1, k);
  return B * (dq1 * a + dq0 * b) / (a + A * b);
}

// Reconstructor.cpp:103-121
int SlopeLimiterK(int limiter, double A, double B) {
  if (limiter != M2CX_LIM_VANALBADA) return 0;
  int k;
  for (k = 2; k < 1000; k++) {
    double rhs = 2.0 / (1.0 + std::pow((k - 1.0) / k, k - 1) / k) * std::min(A, 1.0);
    if (B <= rhs) break;
  }
  return k;
}

// d/dx at 0 of the Lagrange interpolant through nodes d[0..3]: the solution
// of the 4x4 Taylor system of GradientCalculatorFD3.cpp:457-469 (which the
// reference solves with Eigen's column-pivoting QR), evaluated in extended
// precision and rounded once.
static void FD3Coefficients(double d0, double d1, double d2, double d3, double* c) {
  long double d[4] = {d0, d1, d2, d3};
  for (int m = 0; m < 4; m++) {
    long double s = 0;
    for (int l = 0; l < 4; l++) {
      if (l == m) continue;
      long double prod = 1.0L / (d[m] - d[l]);
      for (int n = 0; n < 4; n++) {
        if (n == m || n == l) continue;
        prod *= (0.0L - d[n]) / (d[m] - d[n]);
      }
      s += prod;
    }
    c[m] = (double)s;
  }
}

Solver::Solver(const m2cx_grid& g, const m2cx_material* mats, int nmat_, const m2cx_scheme& s)
    : grid(g), sch(s), nmat(nmat_), riem(vf, RiemannParams::from(s)), godunov_riem(vf, RiemannParams::from(s)) {
  for (int i = 0; i < nmat; i++) vf.push_back(Eos::from(mats[i]));
  // INACTIVE_MATERIAL_ID = number of materials; a dummy EOS sits there
  // (Main.cpp:127-133)
  m2cx_material d{};
  d.eos = M2CX_EOS_DUMMY;
  vf.push_back(Eos::from(d));
  inactive_id = nmat;
  flux.vf = &vf;
  flux.kind = s.flux;
  flux.del = s.entropy_fix;
  if (s.flux == M2CX_FLUX_GODUNOV) flux.godunov = &godunov_riem;
  SetupMesh();
  SetupGhosts();
  SetupFD3();
  long n = m.ncells();
  for (auto* f : {&V, &V1, &V2, &Un, &U1, &R, &Vl, &Vr, &Vb, &Vt, &Vk, &Vf}) f->assign(n * 5, 0.0);
  ID.assign(n, 0);
  IDn.assign(n, 0);
  for (auto* f : {&Phi, &Phi1, &Phi_tmp, &Rls}) f->assign(n, 0.0);
  // domain_diagonal (Main.cpp:207-209)
  domain_diagonal = std::sqrt(std::pow(g.xmax - g.x0, 2) + std::pow(g.ymax - g.y0, 2) + std::pow(g.zmax - g.z0, 2));
  phi_max = -domain_diagonal;
  phi_min = domain_diagonal;
  nls = s.n_levelsets;
  // TimeIntegrator.cpp:46-49
  if (nls > 0 && s.mp_corrections_until_reinit < 1)
    Fail("MultiPhase::CorrectionsUntilReinitialization must be greater than 0");
  DeriveLevelSet();  // cfl = iod_ls.reinit.cfl (LevelSetReinitializer.cpp:39), the band
}

// MeshGenerator.cpp:161-173, SpaceOperator.cpp:173-350
void Solver::SetupMesh() {
  m.NX = grid.nx; m.NY = grid.ny; m.NZ = grid.nz;
  m.PX = m.NX + 2; m.PY = m.NY + 2; m.PZ = m.NZ + 2;
  auto axis = [](int N, double a0, double a1, std::vector<double>& c, std::vector<double>& d) {
    double h = (a1 - a0) / N;
    c.assign(N + 2, 0.0);
    d.assign(N + 2, h);
    for (int i = 0; i < N; i++) c[i + 1] = a0 + 0.5 * h + i * h;
    c[0] = c[1] - d[1];
    c[N + 1] = c[N] + d[N];
    d[0] = d[1];
    d[N + 1] = d[N];
  };
  axis(m.NX, grid.x0, grid.xmax, m.x, m.dx);
  axis(m.NY, grid.y0, grid.ymax, m.y, m.dy);
  axis(m.NZ, grid.z0, grid.zmax, m.z, m.dz);
  vol.assign(m.ncells(), 0.0);
  A3.assign(m.ncells() * 3, 0.0);
  B3.assign(m.ncells() * 3, 0.0);
  K3.assign(m.ncells() * 3, 0);
  LSK3.assign(m.ncells() * 3, 0);
  for (int k = -1; k <= m.NZ; k++)
    for (int j = -1; j <= m.NY; j++)
      for (int i = -1; i <= m.NX; i++) vol[m.idx(i, j, k)] = m.DX(i) * m.DY(j) * m.DZ(k);
  // Reconstructor::Setup (Reconstructor.cpp:53-99)
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        double* A = &A3[c * 3];
        double* B = &B3[c * 3];
        A[0] = (m.DX(i - 1) + m.DX(i)) / (m.DX(i) + m.DX(i + 1));
        A[1] = (m.DY(j - 1) + m.DY(j)) / (m.DY(j) + m.DY(j + 1));
        A[2] = (m.DZ(k - 1) + m.DZ(k)) / (m.DZ(k) + m.DZ(k + 1));
        B[0] = 2.0 * m.DX(i) / (m.DX(i) + m.DX(i + 1));
        B[1] = 2.0 * m.DY(j) / (m.DY(j) + m.DY(j + 1));
        B[2] = 2.0 * m.DZ(k) / (m.DZ(k) + m.DZ(k + 1));
        for (int a = 0; a < 3; a++) K3[c * 3 + a] = SlopeLimiterK(sch.limiter, A[a], B[a]);
        for (int a = 0; a < 3; a++) LSK3[c * 3 + a] = SlopeLimiterK(sch.ls_limiter, A[a], B[a]);
      }
}

// SpaceOperator::CreateGhostNodeLists (SpaceOperator.cpp:472-593): face ghosts
void Solver::SetupGhosts() {
  ghosts.clear();
  for (int k = -1; k <= m.NZ; k++)
    for (int j = -1; j <= m.NY; j++)
      for (int i = -1; i <= m.NX; i++) {
        int out = (i < 0 || i >= m.NX) + (j < 0 || j >= m.NY) + (k < 0 || k >= m.NZ);
        if (out != 1) continue;
        Ghost gp{i, j, k, i, j, k, 0, 0, 0};
        if (i < 0) { gp.ii = -i - 1; gp.side = 0; }
        else if (i >= m.NX) { gp.ii = m.NX + (i - m.NX) - 1; gp.side = 1; }
        else if (j < 0) { gp.jj = -j - 1; gp.side = 2; }
        else if (j >= m.NY) { gp.jj = m.NY + (j - m.NY) - 1; gp.side = 3; }
        else if (k < 0) { gp.kk = -k - 1; gp.side = 4; }
        else { gp.kk = m.NZ + (k - m.NZ) - 1; gp.side = 5; }
        gp.bc = grid.bc[gp.side];
        gp.lsbc = sch.ls_bc[gp.side];
        ghosts.push_back(gp);
      }
}

// GradientCalculatorFD3::CalculateCoefficientsLeftBiased / RightBiased
// (GradientCalculatorFD3.cpp:300-453), including the right-biased y test
// `i<NY-1` of :400.
void Solver::SetupFD3() {
  long n = m.ncells();
  for (auto* f : {&cxL, &cxR, &cyL, &cyR, &czL, &czR}) f->assign(n * 4, 0.0);
  auto first = [](double d, double* c) { c[0] = 0.0; c[1] = -1.0 / d; c[2] = 1.0 / d; c[3] = 0.0; };
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k) * 4;
        // left-biased: i-2, i-1, i, i+1
        if (i > 0) FD3Coefficients(m.X(i - 2) - m.X(i), m.X(i - 1) - m.X(i), 0.0, m.X(i + 1) - m.X(i), &cxL[c]);
        else first(m.X(i) - m.X(i - 1), &cxL[c]);
        if (j > 0) FD3Coefficients(m.Y(j - 2) - m.Y(j), m.Y(j - 1) - m.Y(j), 0.0, m.Y(j + 1) - m.Y(j), &cyL[c]);
        else first(m.Y(j) - m.Y(j - 1), &cyL[c]);
        if (k > 0) FD3Coefficients(m.Z(k - 2) - m.Z(k), m.Z(k - 1) - m.Z(k), 0.0, m.Z(k + 1) - m.Z(k), &czL[c]);
        else first(m.Z(k) - m.Z(k - 1), &czL[c]);
        // right-biased: i-1, i, i+1, i+2
        if (i < m.NX - 1) FD3Coefficients(m.X(i - 1) - m.X(i), 0.0, m.X(i + 1) - m.X(i), m.X(i + 2) - m.X(i), &cxR[c]);
        else first(m.X(i + 1) - m.X(i), &cxR[c]);
        if (i < m.NY - 1 && j + 2 <= m.NY)
          FD3Coefficients(m.Y(j - 1) - m.Y(j), 0.0, m.Y(j + 1) - m.Y(j), m.Y(j + 2) - m.Y(j), &cyR[c]);
        else if (i < m.NY - 1) {
          // stencil reaches the unpopulated second ghost layer; the residual
          // never reads this coefficient set (LevelSetOperator.cpp:840-845)
          cyR[c] = cyR[c + 1] = cyR[c + 2] = cyR[c + 3] = NAN;
        } else first(m.Y(j + 1) - m.Y(j), &cyR[c]);
        if (k < m.NZ - 1) FD3Coefficients(m.Z(k - 1) - m.Z(k), 0.0, m.Z(k + 1) - m.Z(k), m.Z(k + 2) - m.Z(k), &czR[c]);
        else first(m.Z(k + 1) - m.Z(k), &czR[c]);
      }
}

// ---------------------------------------------------------------------------
// SpaceOperator::ApplyBoundaryConditions (SpaceOperator.cpp:738-821)
// ---------------------------------------------------------------------------
void Solver::ApplyBoundaryConditions(std::vector<double>& Vio) {
  for (auto& gp : ghosts) {
    double* v = at5(Vio, m.idx(gp.i, gp.j, gp.k));
    const double* im = at5(Vio, m.idx(gp.ii, gp.jj, gp.kk));
    switch (gp.bc) {
      case M2CX_BC_INLET: for (int p = 0; p < 5; p++) v[p] = grid.inlet[p]; break;
      case M2CX_BC_INLET2: for (int p = 0; p < 5; p++) v[p] = grid.inlet2[p]; break;
      case M2CX_BC_SLIPWALL:
      case M2CX_BC_SYMMETRY:
        for (int p = 0; p < 5; p++) v[p] = im[p];
        v[gp.side / 2 + 1] *= -1.0;
        break;
      case M2CX_BC_STICKWALL:
        v[0] = im[0]; v[1] = -1.0 * im[1]; v[2] = -1.0 * im[2]; v[3] = -1.0 * im[3]; v[4] = im[4];
        break;
      case M2CX_BC_OUTLET: for (int p = 0; p < 5; p++) v[p] = im[p]; break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------------------
// Characteristic variables of the Euler equations along axis `dir`, about the
// cell state V0 (FluxFcnBase.h:309-517).  L = R^{-1} holds the left and R
// the right eigenvectors of dF/dU; Q = dU/dV.  Entries and sums are formed in
// the reference's order, so every product rounds the same way.
// ---------------------------------------------------------------------------
struct CharBasis {
  double e0, c0, dpdrho0, Gamma0, kin0, un0, n[3];
};

// false: c0^2 < 0, where the reference aborts (FluxFcnBase.h:315-320)
static bool CharSetup(const Eos& E, const double* V0, int dir, CharBasis& b) {
  b.e0 = E.GetInternalEnergyPerUnitMass(V0[0], V0[4]);
  double c2 = E.ComputeSoundSpeedSquare(V0[0], b.e0);
  if (c2 < 0) return false;
  b.c0 = std::sqrt(c2);
  b.dpdrho0 = E.GetDpdrho(V0[0], b.e0);
  b.Gamma0 = E.GetBigGamma(V0[0], b.e0);
  b.kin0 = 0.5 * (V0[1] * V0[1] + V0[2] * V0[2] + V0[3] * V0[3]);
  b.un0 = V0[dir + 1];
  for (int q = 0; q < 3; q++) b.n[q] = (q == dir) ? 1.0 : 0.0;
  return true;
}

// FluxFcnBase.h:336-352 (and :445-461)
static void LeftEigenvectors(const CharBasis& b, const double* V0, double L[5][5]) {
  const double al = V0[0] * b.Gamma0 / (V0[4] * b.Gamma0 + V0[0] * b.dpdrho0);
  const double be = (b.e0 - b.kin0 + V0[4] / V0[0]) * al;
  L[0][0] = 0.5 * (1.0 - be + b.un0 / b.c0);
  L[4][0] = 0.5 * (1.0 - be - b.un0 / b.c0);
  L[0][4] = al / 2.0;
  L[4][4] = al / 2.0;
  for (int r = 0; r < 3; r++) {
    L[0][1 + r] = -0.5 * (al * V0[1 + r] + b.n[r] / b.c0);
    L[4][1 + r] = -0.5 * (al * V0[1 + r] - b.n[r] / b.c0);
    L[1 + r][0] = -V0[1 + r] + (be + b.un0) * b.n[r];
    L[1 + r][4] = -al * b.n[r];
    for (int q = 0; q < 3; q++) L[1 + r][1 + q] = (double)(r == q) + al * b.n[r] * V0[1 + q] - b.n[r] * b.n[q];
  }
}

// FluxFcnBase.h:392-410 (and :497-510)
static void RightEigenvectors(const CharBasis& b, const double* V0, double R[5][5]) {
  const double H0 = b.e0 + b.kin0 + V0[4] / V0[0];
  const double par = b.e0 + b.kin0 - b.dpdrho0 / b.Gamma0 - b.un0;
  R[0][0] = 1.0;
  R[0][4] = 1.0;
  R[4][0] = H0 - b.un0 * b.c0;
  R[4][4] = H0 + b.un0 * b.c0;
  for (int r = 0; r < 3; r++) {
    R[0][1 + r] = b.n[r];
    R[1 + r][0] = V0[1 + r] - b.c0 * b.n[r];
    R[1 + r][4] = V0[1 + r] + b.c0 * b.n[r];
    R[4][1 + r] = V0[1 + r] + par * b.n[r];
    for (int q = 0; q < 3; q++) R[1 + r][1 + q] = (double)(r == q) + V0[1 + r] * b.n[q] - b.n[r] * b.n[q];
  }
}

static void MatVec5(const double M[5][5], const double* x, double* y) {
  for (int r = 0; r < 5; r++) {
    double acc = 0.0;
    for (int q = 0; q < 5; q++) acc += M[r][q] * x[q];
    y[r] = acc;
  }
}

// primitive difference -> characteristic (FluxFcnBase.h:309-362): L Q dV
static void PrimToChar(const CharBasis& b, const double* V0, const double* dV, double* dW) {
  double t[5], L[5][5];
  const double q51 = b.e0 + b.kin0 - b.dpdrho0 / b.Gamma0;
  t[0] = dV[0];
  t[1] = V0[1] * dV[0] + V0[0] * dV[1];
  t[2] = V0[2] * dV[0] + V0[0] * dV[2];
  t[3] = V0[3] * dV[0] + V0[0] * dV[3];
  t[4] = q51 * dV[0] + V0[0] * V0[1] * dV[1] + V0[0] * V0[2] * dV[2] + V0[0] * V0[3] * dV[3] + 1.0 / b.Gamma0 * dV[4];
  LeftEigenvectors(b, V0, L);
  MatVec5(L, t, dW);
}

// characteristic -> primitive difference (FluxFcnBase.h:367-421): Q^{-1} R dW
static void CharToPrim(const CharBasis& b, const double* V0, const double* dW, double* dV) {
  double t[5], R[5][5];
  RightEigenvectors(b, V0, R);
  MatVec5(R, dW, t);
  dV[0] = t[0];
  dV[1] = -V0[1] / V0[0] * t[0] + t[1] / V0[0];
  dV[2] = -V0[2] / V0[0] * t[0] + t[2] / V0[0];
  dV[3] = -V0[3] / V0[0] * t[0] + t[3] / V0[0];
  dV[4] = ((-b.e0 + b.kin0) * b.Gamma0 + b.dpdrho0) * t[0] -
          b.Gamma0 * (V0[1] * t[1] + V0[2] * t[2] + V0[3] * t[3] - t[4]);
}

// ---------------------------------------------------------------------------
// Reconstructor::Reconstruct (Reconstructor.cpp:335-918).  Slopes are limited
// in primitive, conservative or characteristic variables
// (ReconstructionData::VariableType, IoData.h:670-671).
// ---------------------------------------------------------------------------
void Solver::ReconstructCell(const std::vector<double>& Vin, long c, int vtype, double* const* fl) {
  int i, j, k;
  m.ijk(c, i, j, k);
  const double* v = at5(Vin, c);
  const int myid = ID[c];
  const Eos& E = vf[myid];
  const bool cons = (vtype == M2CX_VAR_CONSERVATIVE || vtype == M2CX_VAR_CONSERVATIVE_CHARACTERISTIC);
  const bool chr = (vtype == M2CX_VAR_PRIMITIVE_CHARACTERISTIC || vtype == M2CX_VAR_CONSERVATIVE_CHARACTERISTIC);
  const std::vector<double>& Q = cons ? Urec : Vin;  // the variables whose slopes are limited
  const double* q0 = at5(Q, c);
  long nb[6] = {m.idx(i - 1, j, k), m.idx(i + 1, j, k), m.idx(i, j - 1, k),
                m.idx(i, j + 1, k), m.idx(i, j, k - 1), m.idx(i, j, k + 1)};
  // Step 1.1: one-sided differences (:472-574); zero next to an inactive cell
  double dq[6][5];
  for (int p = 0; p < 5; p++) {
    dq[0][p] = (ID[nb[0]] == inactive_id) ? 0.0 : q0[p] - Q[nb[0] * 5 + p];
    dq[1][p] = (ID[nb[1]] == inactive_id) ? 0.0 : Q[nb[1] * 5 + p] - q0[p];
    dq[2][p] = (ID[nb[2]] == inactive_id) ? 0.0 : q0[p] - Q[nb[2] * 5 + p];
    dq[3][p] = (ID[nb[3]] == inactive_id) ? 0.0 : Q[nb[3] * 5 + p] - q0[p];
    dq[4][p] = (ID[nb[4]] == inactive_id) ? 0.0 : q0[p] - Q[nb[4] * 5 + p];
    dq[5][p] = (ID[nb[5]] == inactive_id) ? 0.0 : Q[nb[5] * 5 + p] - q0[p];
  }
  CharBasis cb[3];
  if (chr) {
    for (int a = 0; a < 3; a++)
      if (!CharSetup(E, v, a, cb[a])) {
        char b[160];
        snprintf(b, sizeof b, "c0^2 < 0 in a characteristic transform at (%d,%d,%d): state violates hyperbolicity",
                 i, j, k);
        Fail(b);
        for (int a2 = 0; a2 < 3; a2++) CharSetup(vf[inactive_id], v, a2, cb[a2]);
        break;
      }
    for (int f = 0; f < 6; f++) {
      double w[5];
      if (vtype == M2CX_VAR_PRIMITIVE_CHARACTERISTIC) PrimToChar(cb[f / 2], v, dq[f], w);
      else {
        double L[5][5];
        LeftEigenvectors(cb[f / 2], v, L);
        MatVec5(L, dq[f], w);
      }
      for (int p = 0; p < 5; p++) dq[f][p] = w[p];
    }
  }
  // Step 2.2: limited slopes, component by component (:577-611)
  double alpha = sch.gmm_coeff;
  double sig[3][5];
  for (int p = 0; p < 5; p++)
    for (int a = 0; a < 3; a++) {
      double A = A3[c * 3 + a], B = B3[c * 3 + a];
      double d0 = dq[2 * a][p], d1 = dq[2 * a + 1][p];
      if (sch.limiter == M2CX_LIM_GENERALIZED_MINMOD) sig[a][p] = GeneralizedMinMod(A, B, alpha, d0, d1);
      else if (sch.limiter == M2CX_LIM_VANALBADA) sig[a][p] = VanAlbada(A, B, K3[c * 3 + a], d0, d1);
      else sig[a][p] = 0.5 * (d0 + d1);
    }
  // Step 2.3: back to primitive / conservative differences (:613-633)
  if (chr) {
    for (int a = 0; a < 3; a++) {
      double w[5];
      if (vtype == M2CX_VAR_PRIMITIVE_CHARACTERISTIC) CharToPrim(cb[a], v, sig[a], w);
      else {
        double R[5][5];
        RightEigenvectors(cb[a], v, R);
        MatVec5(R, sig[a], w);
      }
      for (int p = 0; p < 5; p++) sig[a][p] = w[p];
    }
  }
  // Step 2.4 (:638-648)
  if (sch.slope_near_interface == M2CX_SLOPE_ZERO) {
    for (int a = 0; a < 3; a++)
      if (ID[c] != ID[nb[2 * a]] || ID[c] != ID[nb[2 * a + 1]])
        for (int p = 0; p < 5; p++) sig[a][p] = 0.0;
  }
  // Step 2.5: face values (:665-714)
  if (!cons) {
    for (int p = 0; p < 5; p++) {
      fl[0][p] = v[p] - 0.5 * sig[0][p];
      fl[1][p] = v[p] + 0.5 * sig[0][p];
      fl[2][p] = v[p] - 0.5 * sig[1][p];
      fl[3][p] = v[p] + 0.5 * sig[1][p];
      fl[4][p] = v[p] - 0.5 * sig[2][p];
      fl[5][p] = v[p] + 0.5 * sig[2][p];
    }
  } else {
    for (int f = 0; f < 6; f++) {
      double u2[5];
      for (int p = 0; p < 5; p++) u2[p] = (f & 1) ? q0[p] + 0.5 * sig[f / 2][p] : q0[p] - 0.5 * sig[f / 2][p];
      E.ConservativeToPrimitive(u2, fl[f]);
    }
  }
}

void Solver::Reconstruct(const std::vector<double>& Vin) {
  const long n = m.ncells();
  if (sch.rec_type == M2CX_REC_CONSTANT) {
    for (long c = 0; c < n * 5; c++) {
      double x = Vin[c];
      Vl[c] = 0.0 * Vl[c] + 1.0 * x; Vr[c] = 0.0 * Vr[c] + 1.0 * x;
      Vb[c] = 0.0 * Vb[c] + 1.0 * x; Vt[c] = 0.0 * Vt[c] + 1.0 * x;
      Vk[c] = 0.0 * Vk[c] + 1.0 * x; Vf[c] = 0.0 * Vf[c] + 1.0 * x;
    }
    return;
  }
  // V -> U over the ghosted domain, each cell with its own material (:400-409)
  if (sch.var_type == M2CX_VAR_CONSERVATIVE || sch.var_type == M2CX_VAR_CONSERVATIVE_CHARACTERISTIC) {
    Urec.assign(n * 5, 0.0);
    for (long c = 0; c < n; c++) vf[ID[c]].PrimitiveToConservative(at5(Vin, c), at5(Urec, c));
  }
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        const double* v = at5(Vin, c);
        double* fl[6] = {at5(Vl, c), at5(Vr, c), at5(Vb, c), at5(Vt, c), at5(Vk, c), at5(Vf, c)};
        int myid = ID[c];
        if (myid == inactive_id) {
          for (int f = 0; f < 6; f++)
            for (int p = 0; p < 5; p++) fl[f][p] = v[p];
          continue;
        }
        ReconstructCell(Vin, c, sch.var_type, fl);
        if (fatal) return;
        // Step 2.6 (:722-778): the else-if chain stops at the first failing
        // face.  Other variable types retry the whole cell in primitive
        // variables; in primitive variables that face becomes constant.
        for (int f = 0; f < 6; f++)
          if (vf[myid].CheckState(fl[f])) {
            if (sch.var_type != M2CX_VAR_PRIMITIVE) {
              ReconstructCell(Vin, c, M2CX_VAR_PRIMITIVE, fl);
              for (int f2 = 0; f2 < 6; f2++)
                if (vf[myid].CheckState(fl[f2])) {
                  for (int p = 0; p < 5; p++) fl[f2][p] = v[p];
                  break;
                }
            } else {
              for (int p = 0; p < 5; p++) fl[f][p] = v[p];
            }
            break;
          }
      }
  // Step 3: faces of the outer ghost layer (:807-889)
  for (auto& gp : ghosts) {
    long g = m.idx(gp.i, gp.j, gp.k), im = m.idx(gp.ii, gp.jj, gp.kk);
    std::vector<double>* dstf[6] = {&Vr, &Vl, &Vt, &Vb, &Vf, &Vk};  // face of the ghost towards the domain
    std::vector<double>* srcf[6] = {&Vl, &Vr, &Vb, &Vt, &Vk, &Vf};  // face of the image on the boundary
    double* dst = at5(*dstf[gp.side], g);
    const double* src = at5(*srcf[gp.side], im);
    const double* vg = at5(Vin, g);
    switch (gp.bc) {
      case M2CX_BC_INLET:
      case M2CX_BC_INLET2:
      case M2CX_BC_OVERSET:
        for (int p = 0; p < 5; p++) dst[p] = vg[p];
        break;
      case M2CX_BC_OUTLET:
        for (int p = 0; p < 5; p++) dst[p] = src[p];
        break;
      case M2CX_BC_SYMMETRY:
      case M2CX_BC_SLIPWALL:
        for (int p = 0; p < 5; p++) dst[p] = (p == gp.side / 2 + 1) ? -src[p] : src[p];
        break;
      case M2CX_BC_STICKWALL:
        for (int p = 0; p < 5; p++) dst[p] = (p >= 1 && p <= 3) ? -src[p] : src[p];
        break;
      default: break;
    }
  }
}

// SpaceOperator::CheckReconstructedStates (SpaceOperator.cpp:2420-2569)
void Solver::CheckReconstructedStates(const std::vector<double>& Vin) {
  bool error = false;
  for (int k = -1; k <= m.NZ; k++)
    for (int j = -1; j <= m.NY; j++)
      for (int i = -1; i <= m.NX; i++) {
        int boundary = (k == -1 || k == m.NZ) + (j == -1 || j == m.NY) + (i == -1 || i == m.NX);
        if (boundary >= 2) continue;
        long c = m.idx(i, j, k);
        int myid = ID[c];
        if (myid == inactive_id) continue;
        const Eos& E = vf[myid];
        const double* v = at5(Vin, c);
        double* fl[6] = {at5(Vl, c), at5(Vr, c), at5(Vb, c), at5(Vt, c), at5(Vk, c), at5(Vf, c)};
        auto clipf = [&](double* s) {
          if (E.ClipDensityAndPressure(s)) {
            for (int p = 0; p < 5; p++) s[p] = v[p];
            cnt.clipped++;
          }
        };
        if (boundary == 0) {
          for (int f = 0; f < 6; f++) clipf(fl[f]);
          error = E.CheckState(fl[0]) || E.CheckState(fl[1]) || E.CheckState(fl[2]) || E.CheckState(fl[3]) ||
                  E.CheckState(fl[4]) || E.CheckState(fl[5]);
        } else {
          int f = -1;
          if (i == -1) f = 1;
          else if (i == m.NX) f = 0;
          else if (j == -1) f = 3;
          else if (j == m.NY) f = 2;
          else if (k == -1) f = 5;
          else if (k == m.NZ) f = 4;
          clipf(fl[f]);
          error = error || E.CheckState(fl[f]);
        }
        if (error) {
          char b[160];
          snprintf(b, sizeof b, "Reconstructed state at (%d,%d,%d) violates hyperbolicity. matid = %d", i, j, k, myid);
          Fail(b);
          return;
        }
      }
}

// CentralDifferenceLocal (SpaceOperator.h:223-228)
static inline double central_diff3(double phi0, double phi1, double phi2, double x0, double x1, double x2) {
  double c0 = -(x2 - x1) / ((x1 - x0) * (x2 - x0));
  double c1 = 1.0 / (x1 - x0) - 1.0 / (x2 - x1);
  double c2 = (x1 - x0) / ((x2 - x0) * (x2 - x1));
  return c0 * phi0 + c1 * phi1 + c2 * phi2;
}

// SpaceOperator::CalculateGradientAtCellInterface (SpaceOperator.cpp:
// 2272-2415): grad phi at face d of cell (i, j, k) (between i-1 and i for
// d = 0), normal part by the two cells, tangential parts from central
// differences in the two columns weighted to the face, normalized.
void Solver::GradPhiAtFace(const std::vector<double>& phi, int d, int i, int j, int k, double* dir) const {
  auto P = [&](int a, int b, int c) { return phi[m.idx(a, b, c)]; };
  auto X = [&](int a) { return m.X(a); };
  auto Y = [&](int a) { return m.Y(a); };
  auto Z = [&](int a) { return m.Z(a); };
  // derivative along axis e in the column through (a, b, c), one-sided at the
  // domain's ghost rows as the reference does
  auto tang = [&](int e, int a, int b, int c, int Nl) -> double {
    int q = e == 0 ? a : (e == 1 ? b : c);
    auto at = [&](int o) { return e == 0 ? P(a + o, b, c) : (e == 1 ? P(a, b + o, c) : P(a, b, c + o)); };
    auto co = [&](int o) { return e == 0 ? X(a + o) : (e == 1 ? Y(b + o) : Z(c + o)); };
    if (q < 0) return (at(1) - at(0)) / (co(1) - co(0));
    if (q >= Nl) return (at(0) - at(-1)) / (co(0) - co(-1));
    return central_diff3(at(-1), at(0), at(1), co(-1), co(0), co(1));
  };
  dir[0] = dir[1] = dir[2] = 0.0;
  int im = i, jm = j, km = k;
  double xc, xm, dxc;
  if (d == 0) { im = i - 1; xc = X(i); xm = X(i - 1); dxc = m.DX(i); }
  else if (d == 1) { jm = j - 1; xc = Y(j); xm = Y(j - 1); dxc = m.DY(j); }
  else { km = k - 1; xc = Z(k); xm = Z(k - 1); dxc = m.DZ(k); }
  dir[d] = (P(i, j, k) - P(im, jm, km)) / (xc - xm);
  const double xh = xc - 0.5 * dxc;
  const double dx = xc - xm;
  const double c1 = (xc - xh) / dx;
  const double c2 = (xh - xm) / dx;
  const int Ns[3] = {m.NX, m.NY, m.NZ};
  for (int e = 0; e < 3; e++) {
    if (e == d) continue;
    const double t1 = tang(e, im, jm, km, Ns[e]);
    const double t2 = tang(e, i, j, k, Ns[e]);
    dir[e] = c1 * t1 + c2 * t2;
  }
  double nrm = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (nrm != 0.0) {
    double inv = 1.0 / nrm;  // Vec3D::operator/= multiplies by the reciprocal
    dir[0] *= inv; dir[1] *= inv; dir[2] *= inv;
  }
}

// SpaceOperator::GetNormalForBimaterialRiemann and CalculateGradPhiAtCellInterface
// (SpaceOperator.cpp:2186-2270): the face's direction of the 1D Riemann
// problem (myid: this cell, neighborid: the cell on the minus side).  One
// level set is involved when one side is material 0 (its gradient, turned
// when the neighbour is material 0), two otherwise (the difference of their
// unit gradients, normalized).
void Solver::BimaterialNormal(int d, int i, int j, int k, int myid, int neighborid, double* dir) const {
  dir[0] = dir[1] = dir[2] = 0.0;
  if (sch.riemann_normal == M2CX_RNORMAL_MESH) {
    dir[d] = 1.0;
    return;
  }
  const int L = std::max(nls, 1);
  const int my_mat = (myid == 0) ? neighborid : myid;
  const int nb_mat = (neighborid == 0) ? myid : neighborid;
  int my_ls = -1, nb_ls = -1;
  for (int q = 0; q < L; q++) {
    if (my_ls < 0 && MatOf(q) == my_mat) my_ls = q;
    if (nb_ls < 0 && MatOf(q) == nb_mat) nb_ls = q;
  }
  if (my_ls < 0 || nb_ls < 0) {  // the reference asserts
    dir[d] = 1.0;
    return;
  }
  if (my_ls == nb_ls) {
    GradPhiAtFace(PhiOf(my_ls), d, i, j, k, dir);
    if (neighborid == 0)
      for (int q = 0; q < 3; q++) dir[q] *= -1.0;
    if (std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) == 0.0) dir[d] = 1.0;
  } else {
    double n1[3], n2[3];
    GradPhiAtFace(PhiOf(my_ls), d, i, j, k, n1);
    GradPhiAtFace(PhiOf(nb_ls), d, i, j, k, n2);
    for (int q = 0; q < 3; q++) dir[q] = n2[q] - n1[q];
    const double nr = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (nr != 0.0) {
      const double inv = 1.0 / nr;  // Vec3D::operator/= multiplies by the reciprocal
      dir[0] *= inv; dir[1] *= inv; dir[2] *= inv;
    } else {
      dir[d] = 1.0;
    }
  }
  if (dir[d] < 0.0) dir[d] *= -1.0;  // aligned with the positive grid axis
  if (sch.riemann_normal == M2CX_RNORMAL_AVERAGE) {
    dir[d] += 1.0;
    double inv = 1.0 / std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] *= inv; dir[1] *= inv; dir[2] *= inv;
  }
}

int Solver::Riemann(const double* dir, int d, const double* Vm, int idm, const double* Vp, int idp, double* Vs,
                    int& mid, double* Vsm, double* Vsp) {
  riem.st = RiemannStats();
  int err = riem.Solve(dir, Vm, idm, Vp, idp, Vs, mid, Vsm, Vsp);
  cnt.riemann_calls++;
  cnt.riemann_trivial += riem.st.trivial;
  cnt.riemann_iters += riem.st.it_main;
  cnt.riemann_fevals += riem.st.fevals;
  cnt.riemann_fallbacks += riem.st.fallback;
  if (riem_log_on) {  // test-side record of every interface problem
    riem_log.push_back(d);
    for (int p = 0; p < 5; p++) riem_log.push_back(Vm[p]);
    riem_log.push_back(idm);
    for (int p = 0; p < 5; p++) riem_log.push_back(Vp[p]);
    riem_log.push_back(idp);
    riem_log.push_back(err);
    riem_log.push_back(riem.st.it_main);
    riem_log.push_back(riem.st.fevals);
    riem_log.push_back(riem.st.fallback);
  }
  if (riem.st.fatal) Fail("Exact Riemann solver reached a fatal state");
  return err;
}

// ---------------------------------------------------------------------------
// SpaceOperator::ComputeAdvectionFluxes (SpaceOperator.cpp:1432-2076),
// without embedded boundaries and surface tension.
// ---------------------------------------------------------------------------
void Solver::ComputeAdvectionFluxes(const std::vector<double>& Vin, std::vector<double>& F, bool capture) {
  if (capture) rs.clear();
  Reconstruct(Vin);
  CheckReconstructedStates(Vin);
  if (fatal) return;
  std::fill(F.begin(), F.end(), 0.0);
  bool llf_if = (sch.mp_flux == M2CX_MPFLUX_LLF);
  bool const_if = (sch.mp_recon == M2CX_MPREC_CONSTANT);
  for (int k = 0; k <= m.NZ; k++)
    for (int j = 0; j <= m.NY; j++)
      for (int i = 0; i <= m.NX; i++) {
        long c = m.idx(i, j, k);
        int myid = ID[c];
        for (int d = 0; d < 3; d++) {
          if (d == 0 && !(k != m.NZ && j != m.NY)) continue;
          if (d == 1 && !(k != m.NZ && i != m.NX)) continue;
          if (d == 2 && !(j != m.NY && i != m.NX)) continue;
          long nbr = (d == 0) ? m.idx(i - 1, j, k) : (d == 1 ? m.idx(i, j - 1, k) : m.idx(i, j, k - 1));
          const std::vector<double>& FM = (d == 0) ? Vr : (d == 1 ? Vt : Vf);  // minus side: face of neighbour
          const std::vector<double>& FP = (d == 0) ? Vl : (d == 1 ? Vb : Vk);  // plus side: face of this cell
          int nid = ID[nbr];
          double f1[5] = {0, 0, 0, 0, 0}, f2[5] = {0, 0, 0, 0, 0};
          if (nid != myid) {
            if (nid != inactive_id && myid != inactive_id) {
              const double* Vm = const_if ? at5(Vin, nbr) : at5(FM, nbr);
              const double* Vp = const_if ? at5(Vin, c) : at5(FP, c);
              if (llf_if) {
                flux.Llf(d, Vm, nid, Vp, myid, f1);
                for (int p = 0; p < 5; p++) f2[p] = f1[p];
              } else {
                double Vmid[5], Vsm[5], Vsp[5];
                int midid = -1;
                double dir[3];
                BimaterialNormal(d, i, j, k, myid, nid, dir);
                int err = Riemann(dir, d, Vm, nid, Vp, myid, Vmid, midid, Vsm, Vsp);
                if (err) cnt.riemann_errors++;
                vf[nid].ClipDensityAndPressure(Vsm);
                vf[myid].ClipDensityAndPressure(Vsp);
                if (capture && !err) {
                  V5 sm, sp;
                  for (int p = 0; p < 5; p++) { sm[p] = Vsm[p]; sp[p] = Vsp[p]; }
                  if (d == 0) { rs.left[c] = {sm, nid}; rs.right[nbr] = {sp, myid}; }
                  else if (d == 1) { rs.bottom[c] = {sm, nid}; rs.top[nbr] = {sp, myid}; }
                  else { rs.back[c] = {sm, nid}; rs.front[nbr] = {sp, myid}; }
                }
                if (sch.mp_flux == M2CX_MPFLUX_EXACT) {
                  flux.Physical(d, Vmid, midid, f1);
                  for (int p = 0; p < 5; p++) f2[p] = f1[p];
                } else {
                  flux.Numerical(d, Vm, Vsm, nid, f1);
                  flux.Numerical(d, Vsp, Vp, myid, f2);
                }
              }
            }
          } else if (myid != inactive_id) {
            flux.Numerical(d, at5(FM, nbr), at5(FP, c), myid, f1);
            for (int p = 0; p < 5; p++) f2[p] = f1[p];
          }
          double area = (d == 0) ? m.DY(j) * m.DZ(k) : (d == 1 ? m.DX(i) * m.DZ(k) : m.DX(i) * m.DY(j));
          double* fm = at5(F, nbr);
          double* fc = at5(F, c);
          for (int p = 0; p < 5; p++) fm[p] += area * f1[p];
          for (int p = 0; p < 5; p++) fc[p] -= area * f2[p];
        }
      }
  if (flux.godunov_fatal) {
    flux.godunov_fatal = false;
    Fail("Exact Riemann solver reached a fatal state (Godunov flux)");
  }
}

// SpaceOperator::ComputeResidual (SpaceOperator.cpp:2573-2641)
void Solver::ComputeResidual(const std::vector<double>& Vin, std::vector<double>& Rout, bool capture) {
  if (ls_test) {  // SpaceOperator.cpp:2581-2583: no flow residual in the level-set tests
    std::fill(Rout.begin(), Rout.end(), 0.0);
    return;
  }
  ComputeAdvectionFluxes(Vin, Rout, capture);
  if (fatal) return;
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        double cst = 1.0 / (-vol[c]);  // Vec5D::operator/= (Vector5D.h:174-185)
        for (int p = 0; p < 5; p++) Rout[c * 5 + p] *= cst;
      }
}

void Solver::ConservativeToPrimitive(const std::vector<double>& U, std::vector<double>& Vout) {
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        vf[ID[c]].ConservativeToPrimitive(at5(U, c), at5(Vout, c));
      }
}

void Solver::PrimitiveToConservative(const std::vector<double>& Vin, std::vector<double>& U) {
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        vf[ID[c]].PrimitiveToConservative(at5(Vin, c), at5(U, c));
      }
}

// SpaceOperator::ClipDensityAndPressure (SpaceOperator.cpp:646-693), checkState on
int Solver::ClipDensityAndPressure(std::vector<double>& Vio) {
  int n = 0;
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        n += (int)vf[ID[c]].ClipDensityAndPressure(at5(Vio, c));
        if (vf[ID[c]].CheckState(at5(Vio, c))) {
          char b[160];
          snprintf(b, sizeof b, "State variables at (%d,%d,%d) violate hyperbolicity. matid = %d", i, j, k, ID[c]);
          Fail(b);
          return n;
        }
      }
  cnt.clipped += n;
  return n;
}

// SpaceOperator::ComputeTimeStepSize (SpaceOperator.cpp:1223-1330)
void Solver::ComputeTimeStepSize(const std::vector<double>& Vin, double& dt, double& cfl) {
  double dxmin = DBL_MAX;
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        int myid = ID[c];
        if (myid == inactive_id) continue;
        const double* v = at5(Vin, c);
        double c2 = vf[myid].ComputeSoundSpeedSquare(v[0], vf[myid].GetInternalEnergyPerUnitMass(v[0], v[4]));
        if (c2 < 0) { Fail("c^2 < 0 in ComputeTimeStepSize"); return; }
        double lf, lg, lh;
        bool bad = false;
        flux.MaxEigenvalues(v, myid, lf, lg, lh, bad);
        if (m.NX > 1) dxmin = std::min(dxmin, m.DX(i) / lf);
        if (m.NY > 1) dxmin = std::min(dxmin, m.DY(j) / lg);
        if (m.NZ > 1) dxmin = std::min(dxmin, m.DZ(k) / lh);
      }
  if (sch.time_step > 0) {  // Time.TimeStep (SpaceOperator.cpp:1315-1321)
    dt = sch.time_step;
    cfl = dt / dxmin;
  } else {
    cfl = sch.cfl;
    dt = cfl * dxmin;
  }
}

// ---------------------------------------------------------------------------
// LevelSetOperator (LevelSetOperator.cpp:408-508, 796-870) with
// GradientCalculatorFD3::CalculateFirstDerivativeAtNodes (:54-171)
// ---------------------------------------------------------------------------
void Solver::LSApplyBoundaryConditions(std::vector<double>& phi, const std::vector<unsigned char>* useful) {
  for (auto& gp : ghosts) {
    int i = gp.i, j = gp.j, k = gp.k;
    if (useful && !(*useful)[m.idx(i, j, k)]) continue;
    double& ph = phi[m.idx(i, j, k)];
    double im = phi[m.idx(gp.ii, gp.jj, gp.kk)];
    if (gp.lsbc == M2CX_LSBC_ZERO_NEUMANN) {
      ph = im;
    } else if (gp.lsbc == M2CX_LSBC_LINEAR_EXTRAPOLATION || gp.lsbc == M2CX_LSBC_NON_NEGATIVE) {
      double r, r1, r2, f1, f2;
      int s = gp.side;
      if (s == 0) {
        if (i + 2 < m.NX) {
          r = m.X(i); r1 = m.X(i + 1); f1 = phi[m.idx(i + 1, j, k)]; r2 = m.X(i + 2); f2 = phi[m.idx(i + 2, j, k)];
          ph = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
        } else ph = im;
      } else if (s == 1) {
        if (i - 2 >= 0) {
          r = m.X(i); r1 = m.X(i - 1); f1 = phi[m.idx(i - 1, j, k)]; r2 = m.X(i - 2); f2 = phi[m.idx(i - 2, j, k)];
          ph = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
        } else ph = im;
      } else if (s == 2) {
        if (j + 2 < m.NY) {
          r = m.Y(j); r1 = m.Y(j + 1); f1 = phi[m.idx(i, j + 1, k)]; r2 = m.Y(j + 2); f2 = phi[m.idx(i, j + 2, k)];
          ph = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
        } else ph = im;
      } else if (s == 3) {
        if (j - 2 >= 0) {
          r = m.Y(j); r1 = m.Y(j - 1); f1 = phi[m.idx(i, j - 1, k)]; r2 = m.Y(j - 2); f2 = phi[m.idx(i, j - 2, k)];
          ph = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
        } else ph = im;
      } else if (s == 4) {
        if (k + 2 < m.NZ) {
          r = m.Z(k); r1 = m.Z(k + 1); f1 = phi[m.idx(i, j, k + 1)]; r2 = m.Z(k + 2); f2 = phi[m.idx(i, j, k + 2)];
          ph = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
        } else ph = im;
      } else {
        if (k - 2 >= 0) {
          r = m.Z(k); r1 = m.Z(k - 1); f1 = phi[m.idx(i, j, k - 1)]; r2 = m.Z(k - 2); f2 = phi[m.idx(i, j, k - 2)];
          ph = f1 + (f2 - f1) / (r2 - r1) * (r - r1);
        } else ph = im;
      }
      if (gp.lsbc == M2CX_LSBC_NON_NEGATIVE) {
        int ind = s / 2;
        double cg = (ind == 0) ? m.X(i) : (ind == 1 ? m.Y(j) : m.Z(k));
        double ci = (ind == 0) ? m.X(gp.ii) : (ind == 1 ? m.Y(gp.jj) : m.Z(gp.kk));
        double d2wall = 0.5 * std::fabs(ci - cg);
        ph = std::max(-d2wall, ph);
      }
    }
  }
}

void Solver::LSComputeResidual(const std::vector<double>& Vin, const std::vector<double>& phi,
                               std::vector<double>& res) {
  if (nb_bw && !nb_built) {
    Fail("the narrow band of the level set has not been constructed");
    return;
  }
  if (sch.ls_solver == M2CX_LSSOLVER_FINITE_VOLUME) {
    if (nb_bw)
      LSComputeResidualFVMNarrowBand(Vin, phi, res);
    else
      LSComputeResidualFVM(Vin, phi, res);
    return;
  }
  if (nb_bw) {
    LSComputeResidualFDMNarrowBand(Vin, phi, res);
    return;
  }
  auto P = [&](int i, int j, int k) { return phi[m.idx(i, j, k)]; };
  auto fd = [&](const std::vector<double>& C, long c, double v0, double v1, double v2, double v3) {
    const double* cc = &C[c * 4];
    return cc[0] * v0 + cc[1] * v1 + cc[2] * v2 + cc[3] * v3;
  };
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        double a, b, cc, d, e, f;
        a = (i - 2 >= -1) ? fd(cxL, c, P(i - 2, j, k), P(i - 1, j, k), P(i, j, k), P(i + 1, j, k))
                          : (P(i, j, k) - P(i - 1, j, k)) / (m.X(i) - m.X(i - 1));
        b = (i + 2 <= m.NX) ? fd(cxR, c, P(i - 1, j, k), P(i, j, k), P(i + 1, j, k), P(i + 2, j, k))
                            : (P(i + 1, j, k) - P(i, j, k)) / (m.X(i + 1) - m.X(i));
        cc = (j - 2 >= -1) ? fd(cyL, c, P(i, j - 2, k), P(i, j - 1, k), P(i, j, k), P(i, j + 1, k))
                           : (P(i, j, k) - P(i, j - 1, k)) / (m.Y(j) - m.Y(j - 1));
        d = (j + 2 <= m.NY) ? fd(cyR, c, P(i, j - 1, k), P(i, j, k), P(i, j + 1, k), P(i, j + 2, k))
                            : (P(i, j + 1, k) - P(i, j, k)) / (m.Y(j + 1) - m.Y(j));
        e = (k - 2 >= -1) ? fd(czL, c, P(i, j, k - 2), P(i, j, k - 1), P(i, j, k), P(i, j, k + 1))
                          : (P(i, j, k) - P(i, j, k - 1)) / (m.Z(k) - m.Z(k - 1));
        f = (k + 2 <= m.NZ) ? fd(czR, c, P(i, j, k - 1), P(i, j, k), P(i, j, k + 1), P(i, j, k + 2))
                            : (P(i, j, k + 1) - P(i, j, k)) / (m.Z(k + 1) - m.Z(k));
        const double* v = at5(Vin, c);
        double r = (v[1] >= 0 ? a * v[1] : b * v[1]) + (v[2] >= 0 ? cc * v[2] : d * v[2]) +
                   (v[3] >= 0 ? e * v[3] : f * v[3]);
        res[c] = r * -1.0;
      }
}

// ---------------------------------------------------------------------------
// MultiPhaseOperator (MultiPhaseOperator.cpp:919-942, 980-1150, 1278-1525),
// no embedded surfaces; the material-ID update is in levelsets.cpp.
// ---------------------------------------------------------------------------
void Solver::UpdateMaterialIDAtGhostNodes() {
  for (auto& gp : ghosts) ID[m.idx(gp.i, gp.j, gp.k)] = ID[m.idx(gp.ii, gp.jj, gp.kk)];
}

bool Solver::LocalUpdate(int i, int j, int k, int id, bool upwind) {
  long c = m.idx(i, j, k);
  double* v = at5(V, c);
  int counter = 0;
  double sum_weight = 0.0;
  struct Dir { std::map<long, std::pair<V5, int>>* mp; long nb; int comp; double sign; };
  Dir dirs[6] = {{&rs.left, m.idx(i - 1, j, k), 1, 1.0},  {&rs.right, m.idx(i + 1, j, k), 1, -1.0},
                 {&rs.bottom, m.idx(i, j - 1, k), 2, 1.0}, {&rs.top, m.idx(i, j + 1, k), 2, -1.0},
                 {&rs.back, m.idx(i, j, k - 1), 3, 1.0},   {&rs.front, m.idx(i, j, k + 1), 3, -1.0}};
  for (auto& d : dirs) {
    auto it = d.mp->find(c);
    if (it == d.mp->end()) continue;
    const double* vn = at5(V, d.nb);
    bool upw = (d.sign > 0) ? (vn[d.comp] > 0) : (vn[d.comp] < 0);
    if (it->second.second == id && (!upwind || upw)) {
      double nrm = std::sqrt(vn[1] * vn[1] + vn[2] * vn[2] + vn[3] * vn[3]);
      double weight = upwind ? ((d.sign > 0 ? vn[d.comp] : -vn[d.comp]) / nrm) : 1.0;
      sum_weight += weight;
      const V5& s = it->second.first;
      if (counter == 0)
        for (int p = 0; p < 5; p++) v[p] = weight * s[p];
      else
        for (int p = 0; p < 5; p++) v[p] += weight * s[p];
      counter++;
    }
  }
  if (sum_weight > 0.0) {
    double cst = 1.0 / sum_weight;
    for (int p = 0; p < 5; p++) v[p] *= cst;
  }
  return counter > 0;
}

int Solver::UpdateStateVariablesByRiemannSolutions() {
  std::vector<std::array<int, 3>> unresolved;  // (k,j,i) order as the reference stores it
  for (int k = 0; k < m.NZ; k++)
    for (int j = 0; j < m.NY; j++)
      for (int i = 0; i < m.NX; i++) {
        long c = m.idx(i, j, k);
        if (ID[c] == IDn[c]) continue;
        if (ID[c] == inactive_id) continue;
        cnt.id_changed++;
        bool ok = LocalUpdate(i, j, k, ID[c], true);
        if (!ok) ok = LocalUpdate(i, j, k, ID[c], false);
        if (!ok) unresolved.push_back({k, j, i});
      }
  // FixUnresolvedNodes with apply_failsafe_density = On resolves every cell
  // (a failsafe density where no neighbour helps) and reports none still
  // unresolved (MultiPhaseOperator.cpp:1486-1525), so the level-set update
  // and forced reinitialization of TimeIntegrator.cpp:771-797 do not run.
  if (!unresolved.empty()) {
    cnt.unresolved += (long)unresolved.size();
    FixUnresolvedNodes(unresolved);
  }
  return 0;
}

// MultiPhaseOperator::FixUnresolvedNodes (MultiPhaseOperator.cpp:1278-1525),
// apply_failsafe_density = On.  `reset` is declared once for all nodes, as in
// the reference; with PhaseChangeDirection = All the upwind accumulation still
// writes the node (and sets `reset`) but its result is not used (:1396).
void Solver::FixUnresolvedNodes(std::vector<std::array<int, 3>>& unresolved) {
  bool reset = false;
  auto is_unres = [&](int ni, int nj, int nk) {
    for (auto& u : unresolved)
      if (nk == u[0] && nj == u[1] && ni == u[2]) return true;
    return false;
  };
  for (auto& u : unresolved) {
    int k = u[0], j = u[1], i = u[2];
    long c = m.idx(i, j, k);
    double* v = at5(V, c);
    double x0[3] = {m.X(i), m.Y(j), m.Z(k)};
    double sum_weight = 0.0, sum_weight2 = 0.0;
    double vtmp[5] = {0, 0, 0, 0, 0};
    for (int nk = k - 1; nk <= k + 1; nk++)
      for (int nj = j - 1; nj <= j + 1; nj++)
        for (int ni = i - 1; ni <= i + 1; ni++) {
          if (!m.inside(ni, nj, nk)) continue;
          long nc = m.idx(ni, nj, nk);
          if (ID[nc] != ID[c]) continue;
          if (nk == k && nj == j && ni == i) continue;
          if (is_unres(ni, nj, nk)) continue;
          const double* vn = at5(V, nc);
          double v1[3] = {vn[1], vn[2], vn[3]};
          double v1norm = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
          if (v1norm != 0) {
            double cst = 1.0 / v1norm;
            v1[0] *= cst; v1[1] *= cst; v1[2] *= cst;
          }
          double x1x0[3] = {x0[0] - m.X(ni), x0[1] - m.Y(nj), x0[2] - m.Z(nk)};
          double x1x0norm = std::sqrt(x1x0[0] * x1x0[0] + x1x0[1] * x1x0[1] + x1x0[2] * x1x0[2]);
          double cst = 1.0 / x1x0norm;
          x1x0[0] *= cst; x1x0[1] *= cst; x1x0[2] *= cst;
          double weight = std::max(0.0, x1x0[0] * v1[0] + x1x0[1] * v1[1] + x1x0[2] * v1[2]);
          if (weight > 0) {
            sum_weight += weight;
            if (reset)
              for (int p = 0; p < 5; p++) v[p] += weight * vn[p];
            else {
              for (int p = 0; p < 5; p++) v[p] = weight * vn[p];
              reset = true;
            }
          }
          for (int p = 0; p < 5; p++) vtmp[p] += 1.0 / x1x0norm * vn[p];
          sum_weight2 += 1.0 / x1x0norm;
        }
    if (sch.mp_phase_dir == M2CX_PHASEDIR_UPWIND && sum_weight > 0) {
      double cst = 1.0 / sum_weight;
      for (int p = 0; p < 5; p++) v[p] *= cst;
      continue;
    }
    if (sum_weight2 > 0) {
      for (int p = 0; p < 5; p++) v[p] = vtmp[p] / sum_weight2;
      continue;
    }
    double density = 0.0;
    for (int layer = 1; layer <= 10; layer++) {
      for (int nk = k - layer; nk <= k + layer; nk++)
        for (int nj = j - layer; nj <= j + layer; nj++)
          for (int ni = i - layer; ni <= i + layer; ni++) {
            if (!m.inside(ni, nj, nk)) continue;
            long nc = m.idx(ni, nj, nk);
            if (ID[nc] != ID[c]) continue;
            if (nk == k && nj == j && ni == i) continue;
            if (is_unres(ni, nj, nk)) continue;
            double dxv[3] = {m.X(ni) - x0[0], m.Y(nj) - x0[1], m.Z(nk) - x0[2]};
            double dist = std::sqrt(dxv[0] * dxv[0] + dxv[1] * dxv[1] + dxv[2] * dxv[2]);
            sum_weight += 1.0 / dist;
            density += (1.0 / dist) * V[nc * 5];
          }
      if (sum_weight > 0) { v[0] = density / sum_weight; break; }
    }
    if (sum_weight == 0) v[0] = vf[ID[c]].failsafe_density;
  }
}

// ---------------------------------------------------------------------------
// TimeIntegrator (TimeIntegrator.cpp:167-715) and the end-of-step update
// (:720-843)
// ---------------------------------------------------------------------------
void Solver::AdvanceOneTimeStep(double time, double dt) {
  (void)time;
  const bool ls = sch.n_levelsets > 0;
  // Phi_tmp = Phi(n) including ghosts; with the narrow band only the useful
  // nodes after the first step (TimeIntegrator.cpp:333-339)
  if (ls) ForEachLS([&](int) {
    if (!nb_bw || time_step == 1) Phi_tmp = Phi;
    else
      for (long c : useful_nodes) Phi_tmp[c] = 0.0 * Phi_tmp[c] + 1.0 * Phi[c];
  });
  if (sch.integrator == M2CX_TI_FORWARD_EULER) {
    ComputeResidual(V, R, true);
    if (fatal) return;
    PrimitiveToConservative(V, Un);
    for (long q = 0; q < (long)Un.size(); q++) Un[q] = 1.0 * Un[q] + dt * R[q];
    ConservativeToPrimitive(Un, V);
    ClipDensityAndPressure(V);
    if (fatal) return;
    ApplyBoundaryConditions(V);
    if (ls) ForEachLS([&](int) {
      if (ls_test) PrescribeVelocityForTesting(V, Phi, time - dt);
      LSComputeResidual(V, Phi, Rls);
      LSInterior([&](long c) { Phi[c] = 1.0 * Phi[c] + dt * Rls[c]; });
      LSApplyBoundaryConditions(Phi);
    });
  } else {
    const bool rk3 = (sch.integrator == M2CX_TI_RK3);
    auto interior = [&](auto fn) {
      for (int k = 0; k < m.NZ; k++)
        for (int j = 0; j < m.NY; j++)
          for (int i = 0; i < m.NX; i++) fn(m.idx(i, j, k));
    };
    // stage 1
    ComputeResidual(V, R, true);
    if (fatal) return;
    PrimitiveToConservative(V, Un);
    interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 0.0 * U1[c * 5 + p] + 1.0 * Un[c * 5 + p]; });
    interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 1.0 * U1[c * 5 + p] + dt * R[c * 5 + p]; });
    ConservativeToPrimitive(U1, V1);
    int clipped = ClipDensityAndPressure(V1);
    if (fatal) return;
    if (clipped) PrimitiveToConservative(V1, U1);
    ApplyBoundaryConditions(V1);
    if (ls) ForEachLS([&](int) {
      if (ls_test) PrescribeVelocityForTesting(V, Phi, time - dt);
      LSComputeResidual(V, Phi, Rls);
      LSInterior([&](long c) { Phi1[c] = 0.0 * Phi1[c] + 1.0 * Phi[c]; });
      LSInterior([&](long c) { Phi1[c] = 1.0 * Phi1[c] + dt * Rls[c]; });
      LSApplyBoundaryConditions(Phi1);
    });
    if (!rk3) {
      // stage 2 of RK2
      ComputeResidual(V1, R, false);
      if (fatal) return;
      const double hdt = 0.5 * dt;
      interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 0.5 * U1[c * 5 + p] + 0.5 * Un[c * 5 + p]; });
      interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 1.0 * U1[c * 5 + p] + hdt * R[c * 5 + p]; });
      ConservativeToPrimitive(U1, V);
      ClipDensityAndPressure(V);
      if (fatal) return;
      ApplyBoundaryConditions(V);
      if (ls) ForEachLS([&](int) {
        if (ls_test) PrescribeVelocityForTesting(V1, Phi1, time);
        LSComputeResidual(V1, Phi1, Rls);
        LSInterior([&](long c) { Phi[c] = 0.5 * Phi[c] + 0.5 * Phi1[c]; });
        LSInterior([&](long c) { Phi[c] = 1.0 * Phi[c] + hdt * Rls[c]; });
        LSApplyBoundaryConditions(Phi);
      });
    } else {
      // stage 2 of RK3
      ComputeResidual(V1, R, false);
      if (fatal) return;
      const double qdt = 0.25 * dt;
      interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 0.25 * U1[c * 5 + p] + 0.75 * Un[c * 5 + p]; });
      interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 1.0 * U1[c * 5 + p] + qdt * R[c * 5 + p]; });
      ConservativeToPrimitive(U1, V2);
      clipped = ClipDensityAndPressure(V2);
      if (fatal) return;
      if (clipped) PrimitiveToConservative(V2, U1);
      ApplyBoundaryConditions(V2);
      if (ls) ForEachLS([&](int) {
        if (ls_test) PrescribeVelocityForTesting(V1, Phi1, time);
        LSComputeResidual(V1, Phi1, Rls);
        LSInterior([&](long c) { Phi1[c] = 0.25 * Phi1[c] + 0.75 * Phi[c]; });
        LSInterior([&](long c) { Phi1[c] = 1.0 * Phi1[c] + qdt * Rls[c]; });
        LSApplyBoundaryConditions(Phi1);
      });
      // stage 3
      ComputeResidual(V2, R, false);
      if (fatal) return;
      const double tdt = 2.0 / 3.0 * dt;
      interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 2.0 / 3.0 * U1[c * 5 + p] + 1.0 / 3.0 * Un[c * 5 + p]; });
      interior([&](long c) { for (int p = 0; p < 5; p++) U1[c * 5 + p] = 1.0 * U1[c * 5 + p] + tdt * R[c * 5 + p]; });
      ConservativeToPrimitive(U1, V);
      ClipDensityAndPressure(V);
      if (fatal) return;
      ApplyBoundaryConditions(V);
      if (ls) ForEachLS([&](int) {
        if (ls_test) PrescribeVelocityForTesting(V2, Phi1, time - 0.5 * dt);
        LSComputeResidual(V2, Phi1, Rls);
        LSInterior([&](long c) { Phi[c] = 1.0 / 3.0 * Phi[c] + 2.0 / 3.0 * Phi1[c]; });
        LSInterior([&](long c) { Phi[c] = 1.0 * Phi[c] + tdt * Rls[c]; });
        LSApplyBoundaryConditions(Phi);
      });
    }
  }
  if (ls) UpdateSolutionAfterTimeStepping();
}

// TimeIntegratorBase::UpdateSolutionAfterTimeStepping (TimeIntegrator.cpp:
// 720-795) for one level set and no embedded surfaces: ID update, the
// level-set reinitialization at its frequency, then the state update.
void Solver::UpdateSolutionAfterTimeStepping() {
  // overlapping subdomains of several level sets (:737-741)
  const int resolved = ResolveConflictsInLevelSets();
  if (resolved) ForEachLS([&](int) { LSApplyBoundaryConditions(Phi); });
  IDn = ID;
  if (!ls_test) UpdateMaterialIDByLevelSets();
  if (fatal) return;
  // reinitialization: all level sets once enough overlaps were resolved,
  // otherwise each at its own frequency (:757-768)
  ls_correction_bucket += resolved;
  const bool must = ls_correction_bucket >= sch.mp_corrections_until_reinit;
  if (must) ls_correction_bucket = 0;
  ForEachLS([&](int) { Reinitialize(t, dts, time_step, 0, must); });
  if (fatal || ls_test) return;
  if (sch.mp_phase_change == M2CX_PHASECHANGE_EXTRAPOLATION)
    UpdateStateVariablesByExtrapolation();
  else
    UpdateStateVariablesByRiemannSolutions();
  if (fatal) return;
  ClipDensityAndPressure(V);
  if (fatal) return;
  ApplyBoundaryConditions(V);
}

// Main.cpp:653-712 (no concurrent programs, no subcycling)
int Solver::RunSteps(int nsteps) {
  double tmax = (sch.max_time > 0) ? sch.max_time : DBL_MAX;
  for (int s = 0; s < nsteps; s++) {
    if (!(t < tmax - 1.0e-6 * dts)) break;
    time_step++;
    double dt, cfl;
    ComputeTimeStepSize(V, dt, cfl);
    if (fatal) return -1;
    if (t + dt > tmax) {
      cfl *= (tmax - t) / dt;
      dt = tmax - t;
    }
    dts = dt;
    t += dt;
    dt_last = dt;
    AdvanceOneTimeStep(t, dt);
    if (fatal) return -1;
    cnt.steps++;
  }
  return 0;
}

}  // namespace orc
