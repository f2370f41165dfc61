// Multi-material part of the explicit step on gfx950:
//  - classification of material-interface faces into a compact list, once
//    per time step (IDs do not change between RK stages);
//  - the exact two-material Riemann solve over that list, each stage, giving
//    the two interface fluxes and, in stage 1, the star states kept for the
//    end-of-step update (SpaceOperator.cpp:1613-1690);
//  - level-set FD3 advection stages (LevelSetOperator.cpp:796-870,
//    GradientCalculatorFD3.cpp:54-171) and boundary fill (:408-508);
//  - the end-of-step material update: IDs from the level-set sign change with
//    the neighbour vote for swept cells (MultiPhaseOperator.cpp:153-357),
//    ghost IDs (:919-942), and the Riemann-solution state update of cells that
//    changed material (:980-1150) with its failsafe (:1278-1525).
#pragma once
#include <string>
#include <vector>
#include "m2c_kernels.h"
#include "m2c_riemann.h"

namespace m2c {

// Lanes of an interface Riemann launch: 1024 waves, one per SIMD of the 256
// CUs (the exact solver's register use leaves room for one resident wave per
// SIMD).  Each lane owns 2 x 512 integration-path points (24 KiB).
constexpr int RIEM_LANES = 1024 * 64;
// integration-path points per lane and path of the Godunov face solves
constexpr int GODUNOV_PATHCAP = 512;

// LevelSet[0] finite-volume settings (LevelSetSchemeData, IoData.h:802-830)
struct LSFvm {
  int on, flux, rec, limiter;
  double gmm, delta;
  const int* lsk;  // van Albada exponents of the level set's limiter, laid out as DevGrid::kt
};

struct MPBuffers {
  int* fidx[3] = {};
  IfaceList L{};
  double *phi = nullptr, *phi1 = nullptr, *phin = nullptr, *rls = nullptr, *phitmp = nullptr;
  double* fd3 = nullptr;  // FD3 coefficient tables (see mp_create)
  long off_cxL = 0, off_cxR = 0, off_cyL = 0, off_cyR3 = 0, off_cyR1 = 0, off_czL = 0, off_czR = 0;
  int lsid = 1, nls = 0;
  LSFvm fv{};  // finite-volume level-set settings (fv.lsk owned here)
  int lsbc[6] = {1, 1, 1, 1, 1, 1};
  RiemannCfg rc{};
  double* slab = nullptr;
  int lanes = 0, pathcap = 0;
  int* remaining = nullptr;
  int* changed = nullptr;
  int* unresolved = nullptr;
  unsigned char* tag = nullptr;
  unsigned char* done = nullptr;
  unsigned char* IDn = nullptr;
  int list_cap = 0;
  // learned p* guesses per interface slot (see m2c_mlp.h); null = acoustic
  float* pguess = nullptr;
  int guess = 0;
};

// FD3 coefficients: d/dx at 0 of the Lagrange interpolant through d[0..3]
// (the 4x4 Taylor system GradientCalculatorFD3.cpp:457-469 solves with Eigen),
// in extended precision, rounded once.
static inline void fd3_coeffs(double d0, double d1, double d2, double d3, double* c) {
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

inline int mp_create(MPBuffers& B, DevGrid& g, const m2cx_scheme& sch, std::vector<double>* ax, hipStream_t st) {
  const long n = g.ncells;
  const long nin = (long)g.NX * g.NY * g.NZ;
  for (int d = 0; d < 3; d++) {
    if (hipMalloc(&B.fidx[d], n * 4) != hipSuccess) return M2CX_E_OOM;
    if (hipMemsetAsync(B.fidx[d], 0, n * 4, st) != hipSuccess) return M2CX_E_HIP;
  }
  // interface faces: a level-set surface crosses at most a few faces per
  // cell of a cross-section; 16 x (sum of the three cross-sections) covers
  // any bubble population here, and overflow is reported as an error
  const long sections = (long)g.NX * g.NY + (long)g.NY * g.NZ + (long)g.NX * g.NZ;
  long cap = std::max<long>(1 << 16, std::min<long>(std::min<long>(nin, 16 * sections), 1L << 22));
  B.L.cap = (int)cap;
  if (hipMalloc(&B.L.cell, cap * 4) || hipMalloc(&B.L.dir, cap * 4) || hipMalloc(&B.L.f1, cap * 40) ||
      hipMalloc(&B.L.f2, cap * 40) || hipMalloc(&B.L.vsm, cap * 40) || hipMalloc(&B.L.vsp, cap * 40) ||
      hipMalloc(&B.L.err, cap * 4) || hipMalloc(&B.L.stats, cap * 12) ||
      hipMalloc(&B.L.hist, 2 * M2CX_RIEMANN_HIST_BINS * 8))
    return M2CX_E_OOM;
  if (hipMemsetAsync(B.L.hist, 0, 2 * M2CX_RIEMANN_HIST_BINS * 8, st) != hipSuccess) return M2CX_E_HIP;
  for (double** p : {&B.phi, &B.phi1, &B.phin, &B.rls, &B.phitmp}) {
    if (hipMalloc(p, n * 8) != hipSuccess) return M2CX_E_OOM;
    if (hipMemsetAsync(*p, 0, n * 8, st) != hipSuccess) return M2CX_E_HIP;
  }
  // FD3 tables (GradientCalculatorFD3.cpp:300-453), per axis because the mesh
  // is a tensor product; cyR1 holds the first-order right-biased y stencil the
  // reference uses where its `i<NY-1` test (:400) fails.
  const std::vector<double>&x = ax[0], &y = ax[1], &z = ax[2];
  const int NX = g.NX, NY = g.NY, NZ = g.NZ;
  std::vector<double> t;
  auto put = [&](long& off, int N, auto fn) {
    off = (long)t.size();
    t.resize(t.size() + (size_t)N * 4);
    for (int i = 0; i < N; i++) fn(i, &t[off + 4L * i]);
  };
  auto first = [](double d, double* c) { c[0] = 0.0; c[1] = -1.0 / d; c[2] = 1.0 / d; c[3] = 0.0; };
  auto C = [](const std::vector<double>& v, int i) { return v[i + 1]; };
  put(B.off_cxL, NX, [&](int i, double* c) {
    if (i > 0) fd3_coeffs(C(x, i - 2) - C(x, i), C(x, i - 1) - C(x, i), 0.0, C(x, i + 1) - C(x, i), c);
    else first(C(x, i) - C(x, i - 1), c);
  });
  put(B.off_cxR, NX, [&](int i, double* c) {
    if (i < NX - 1) fd3_coeffs(C(x, i - 1) - C(x, i), 0.0, C(x, i + 1) - C(x, i), C(x, i + 2) - C(x, i), c);
    else first(C(x, i + 1) - C(x, i), c);
  });
  put(B.off_cyL, NY, [&](int j, double* c) {
    if (j > 0) fd3_coeffs(C(y, j - 2) - C(y, j), C(y, j - 1) - C(y, j), 0.0, C(y, j + 1) - C(y, j), c);
    else first(C(y, j) - C(y, j - 1), c);
  });
  put(B.off_cyR3, NY, [&](int j, double* c) {
    if (j + 2 <= NY) fd3_coeffs(C(y, j - 1) - C(y, j), 0.0, C(y, j + 1) - C(y, j), C(y, j + 2) - C(y, j), c);
    else c[0] = c[1] = c[2] = c[3] = 0.0;  // never read (LevelSetOperator.cpp:840-845)
  });
  put(B.off_cyR1, NY, [&](int j, double* c) { first(C(y, j + 1) - C(y, j), c); });
  put(B.off_czL, NZ, [&](int k, double* c) {
    if (k > 0) fd3_coeffs(C(z, k - 2) - C(z, k), C(z, k - 1) - C(z, k), 0.0, C(z, k + 1) - C(z, k), c);
    else first(C(z, k) - C(z, k - 1), c);
  });
  put(B.off_czR, NZ, [&](int k, double* c) {
    if (k < NZ - 1) fd3_coeffs(C(z, k - 1) - C(z, k), 0.0, C(z, k + 1) - C(z, k), C(z, k + 2) - C(z, k), c);
    else first(C(z, k + 1) - C(z, k), c);
  });
  if (hipMalloc(&B.fd3, t.size() * 8) != hipSuccess) return M2CX_E_OOM;
  if (hipMemcpy(B.fd3, t.data(), t.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return M2CX_E_HIP;
  B.nls = sch.n_levelsets;
  B.lsid = sch.ls_materialid;
  for (int q = 0; q < 6; q++) B.lsbc[q] = sch.ls_bc[q];
  B.rc = RiemannCfg{sch.riem_maxits_main, sch.riem_maxits_bracket, sch.riem_maxits_shock,
                    sch.riem_steps_rarefaction, sch.riem_tol_main, sch.riem_tol_shock, sch.riem_tol_rarefaction,
                    sch.riem_min_pressure, sch.riem_failure_threshold, sch.riem_pressure_at_failure};
  B.guess = sch.riem_guess;
  // per-lane integration-path scratch: lanes x 2 paths x cap points x 3
  B.lanes = RIEM_LANES;
  if (const char* e = getenv("M2C_RIEMANN_LANES")) {  // launch-shape override: a multiple of 64
    const long v = atol(e);
    if (v < 64 || v > 4L * RIEM_LANES || v % 64) return M2CX_E_ARG;
    B.lanes = (int)v;
  }
  B.pathcap = 512;
  // the integration-path slab is allocated on first use (mp_slab): solvers
  // that run in an ensemble share the ensemble's slab instead
  B.list_cap = (int)std::max<long>(1 << 16, std::min<long>(nin, 1L << 22));
  if (hipMalloc(&B.remaining, B.list_cap * 4) || hipMalloc(&B.changed, B.list_cap * 4) ||
      hipMalloc(&B.unresolved, B.list_cap * 4) || hipMalloc(&B.tag, n) || hipMalloc(&B.done, n) ||
      hipMalloc(&B.IDn, n))
    return M2CX_E_OOM;
  if (hipMemsetAsync(B.tag, 0, n, st) || hipMemsetAsync(B.done, 0, n, st) || hipMemsetAsync(B.IDn, 0, n, st))
    return M2CX_E_HIP;
  return 0;
}

inline void mp_destroy(MPBuffers& B) {
  for (int d = 0; d < 3; d++) (void)hipFree(B.fidx[d]);
  if (B.fv.lsk) (void)hipFree((void*)B.fv.lsk);
  for (void* p : {(void*)B.L.cell, (void*)B.L.dir, (void*)B.L.f1, (void*)B.L.f2, (void*)B.L.vsm, (void*)B.L.vsp,
                  (void*)B.L.err, (void*)B.L.stats, (void*)B.L.hist, (void*)B.phi, (void*)B.phi1, (void*)B.phin, (void*)B.rls,
                  (void*)B.phitmp, (void*)B.fd3, (void*)B.slab, (void*)B.remaining, (void*)B.changed,
                  (void*)B.unresolved, (void*)B.tag, (void*)B.done, (void*)B.IDn})
    (void)hipFree(p);
}

// ---------------------------------------------------------------------------
// interface classification
// ---------------------------------------------------------------------------
__global__ void k_if_reset(DevCounters* cnt) { cnt->nif = 0; }

__global__ void k_classify(DevGrid g, Scheme s, const unsigned char* __restrict__ ID, int* fx, int* fy, int* fz,
                           IfaceList L, DevCounters* cnt) {
  long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    int my = ID[c];
    if (my == s.inactive) continue;
    int a3[3] = {i, j, k};
    long st3[3] = {1, g.SX, g.S};
    int* fi3[3] = {fx, fy, fz};
    for (int d = 0; d < 3; d++) {
      if (a3[d] == 0) continue;  // ghost neighbour: its ID is the image's
      int nb = ID[c - st3[d]];
      if (nb == my || nb == s.inactive) continue;
      int slot = atomicAdd(&cnt->nif, 1);
      if (slot < L.cap) {
        L.cell[slot] = (int)c;
        L.dir[slot] = d;
        fi3[d][c] = slot;
      } else {
        atomicOr(&cnt->error, 4);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// exact Riemann over the interface list
// ---------------------------------------------------------------------------
__device__ __forceinline__ void if_states(const DevGrid& g, const Scheme& s, const DevMats& M, const CF5& V,
                                          const unsigned char* ID, int D, long c, double* Vm, double* Vp) {
  const long st = axis_stride(g, D);
  if (s.mp_recon == M2CX_MPREC_CONSTANT) {
    load5(V, c - st, Vm);
    load5(V, c, Vp);
    return;
  }
  long r = c;
  int i = (int)(r % g.SX) - 1;
  r /= g.SX;
  int j = (int)(r % g.SY) - 1;
  int k = (int)(r / g.SY) - 1;
  int im = (D == 0) ? i - 1 : i, jm = (D == 1) ? j - 1 : j, km = (D == 2) ? k - 1 : k;
  bool f, u;
  face_state(g, s, M, V, ID, D, im, jm, km, 1, Vm, f, u);
  face_state(g, s, M, V, ID, D, i, j, k, 0, Vp, f, u);
}

// CentralDifferenceLocal (SpaceOperator.h:223-228)
__device__ __forceinline__ double cdiff3(double phi0, double phi1, double phi2, double x0, double x1, double x2) {
  double c0 = -(x2 - x1) / ((x1 - x0) * (x2 - x0));
  double c1 = 1.0 / (x1 - x0) - 1.0 / (x2 - x1);
  double c2 = (x1 - x0) / ((x2 - x0) * (x2 - x1));
  return c0 * phi0 + c1 * phi1 + c2 * phi2;
}

// CalculateGradientAtCellInterface (SpaceOperator.cpp:2274-2415) at the face
// of cell c on axis d (minus side c - stride): grad phi with its normal part
// from the two cells and tangential parts from central differences in both
// columns weighted to the face, normalized.  Interface faces lie between
// interior cells, so the columns' neighbours are interior cells or face
// ghosts (the reference's one-sided branches at ghost rows are not reached).
__device__ __forceinline__ void grad_at_face(const DevGrid& g, const double* __restrict__ phi, int d, long c,
                                             double* dir) {
  dir[0] = dir[1] = dir[2] = 0.0;
  long r = c;
  const int i = (int)(r % g.SX) - 1; r /= g.SX;
  const int j = (int)(r % g.SY) - 1;
  const int k = (int)(r / g.SY) - 1;
  const int ijk[3] = {i, j, k};
  const long st3[3] = {1, g.SX, g.S};
  const double* X = g_axis(g, d);
  const double* DX = g_axis(g, 3 + d);
  const long sd = st3[d];
  const double xc = X[ijk[d] + 1], xm = X[ijk[d]];
  dir[d] = (phi[c] - phi[c - sd]) / (xc - xm);
  const double xh = xc - 0.5 * DX[ijk[d] + 1];
  const double dx = xc - xm;
  const double c1 = (xc - xh) / dx;
  const double c2 = (xh - xm) / dx;
  for (int e = 0; e < 3; e++) {
    if (e == d) continue;
    const double* Y = g_axis(g, e);
    const long se = st3[e];
    const int q = ijk[e];
    const double t1 = cdiff3(phi[c - sd - se], phi[c - sd], phi[c - sd + se], Y[q], Y[q + 1], Y[q + 2]);
    const double t2 = cdiff3(phi[c - se], phi[c], phi[c + se], Y[q], Y[q + 1], Y[q + 2]);
    dir[e] = c1 * t1 + c2 * t2;
  }
  double nrm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  if (nrm != 0.0) {
    const double inv = 1.0 / nrm;  // Vec3D::operator/= multiplies by the reciprocal
    dir[0] *= inv; dir[1] *= inv; dir[2] *= inv;
  }
}

// GetNormalForBimaterialRiemann / CalculateGradPhiAtCellInterface
// (SpaceOperator.cpp:2186-2270) at face c - 1/2 e_d between material idl
// (minus side) and idr: one level set when one side is material 0 (its
// gradient, turned when the minus side is material 0), two otherwise (the
// difference of their unit gradients, normalized); then aligned with +e_d.
// A material no level set tracks (the reference asserts) gives the mesh
// normal.
__device__ __forceinline__ void bimat_normal(const DevGrid& g, int rnormal, const IfaceList& L, int d, long c,
                                             int idl, int idr, double* dir) {
  dir[0] = dir[1] = dir[2] = 0.0;
  if (rnormal == M2CX_RNORMAL_MESH) {
    dir[d] = 1.0;
    return;
  }
  const int my_mat = (idr == 0) ? idl : idr;
  const int nb_mat = (idl == 0) ? idr : idl;
  int my_ls = -1, nb_ls = -1;
  for (int q = 0; q < L.nls; q++) {
    if (my_ls < 0 && L.lsmat[q] == my_mat) my_ls = q;
    if (nb_ls < 0 && L.lsmat[q] == nb_mat) nb_ls = q;
  }
  if (my_ls < 0 || nb_ls < 0) {
    dir[d] = 1.0;
    return;
  }
  if (my_ls == nb_ls) {
    grad_at_face(g, L.lsphi[my_ls], d, c, dir);
    if (idl == 0)
      for (int e = 0; e < 3; e++) dir[e] *= -1.0;
    if (sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) == 0.0) dir[d] = 1.0;
  } else {
    double n1[3], n2[3];
    grad_at_face(g, L.lsphi[my_ls], d, c, n1);
    grad_at_face(g, L.lsphi[nb_ls], d, c, n2);
    for (int e = 0; e < 3; e++) dir[e] = n2[e] - n1[e];
    const double nr = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (nr != 0.0) {
      const double inv = 1.0 / nr;
      dir[0] *= inv; dir[1] *= inv; dir[2] *= inv;
    } else {
      dir[d] = 1.0;
    }
  }
  if (dir[d] < 0.0) dir[d] *= -1.0;
  if (rnormal == M2CX_RNORMAL_AVERAGE) {
    dir[d] += 1.0;
    const double inv = 1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] *= inv; dir[1] *= inv; dir[2] *= inv;
  }
}

// FluxFcnGodunov.h:37-60: the flux through a face with one material on both
// sides is the physical flux of the exact Riemann solution at x/t = 0 along
// the mesh axis.  Uses R's integration-path slot; the caller keeps R's
// statistics apart (they are not interface statistics).  Returns false on a
// solve the reference ends the run on; `ovf` reports an integration path
// longer than the slot holds.
__device__ __forceinline__ bool flux_godunov(const DevMats& M, DevRiemann& R, int d, const double* Vm,
                                             const double* Vp, int id, double* F, bool& ovf) {
  const double dir[3] = {d == 0 ? 1.0 : 0.0, d == 1 ? 1.0 : 0.0, d == 2 ? 1.0 : 0.0};
  double Vmid[5], Vsm[5], Vsp[5];
  int midid = id;
  R.st = RStats{0, 0, false, false, false};
  R.path1.overflow = R.path3.overflow = false;
  R.solve(dir, Vm, id, Vp, id, Vmid, midid, Vsm, Vsp);
  ovf = R.path1.overflow || R.path3.overflow;
  flux_phys(M.m[id], d, Vmid, F);
  return !R.st.fatal;
}

// One interface face: the face states, the exact solve and the two fluxes
// into the list (SpaceOperator.cpp:1613-1690), shared by the single-solver
// and the ensemble kernels.  Returns flag bits: 1 trivial, 2 error, 4
// fallback solver used, 8 learned guess rejected.  GOD: the side fluxes of
// MultiPhase.Flux = Numerical are Godunov fluxes (Flux = Godunov); a separate
// instantiation, so that the common kernel does not carry that code's
// registers.
template <bool GOD = false>
__device__ __forceinline__ int riemann_face(const DevGrid& g, const Scheme& s, const CF5& V,
                                            const unsigned char* __restrict__ ID, const IfaceList& L, int q,
                                            int stage1, const float* guess, DevRiemann& R, int& fatal, int& ovf) {
  const DevMats& M = c_mats[s.mset];
  const long c = L.cell[q];
  const int d = L.dir[q];
  const long st = axis_stride(g, d);
  const int idl = ID[c - st], idr = ID[c];
  double Vm[5], Vp[5];
  if_states(g, s, M, V, ID, d, c, Vm, Vp);
  double dir[3];
  bimat_normal(g, s.rnormal, L, d, c, idl, idr, dir);
  // learned p* guess for this face (m2c_mlp.h), rejected when not a finite
  // positive pressure: the acoustic guess is then used
  double p0g = -1.0;
  int flags = 0;
  if (guess) {
    const float gp = guess[q];
    p0g = (isfinite(gp) && gp > 0.0f) ? (double)gp : -1.0;
    if (p0g <= 0) flags |= 8;
  }
  double Vs[5], Vsm[5], Vsp[5];
  int mid = idl;
  R.st = RStats{0, 0, false, false, false};
  R.path1.overflow = R.path3.overflow = false;
  const int err = R.solve(dir, Vm, idl, Vp, idr, Vs, mid, Vsm, Vsp, p0g);
  flags |= (R.st.trivial ? 1 : 0) | (err != 0 ? 2 : 0) | (R.st.fallback ? 4 : 0);
  fatal |= R.st.fatal;
  ovf |= R.path1.overflow || R.path3.overflow;
  eos_clip(M.m[idl], Vsm);
  eos_clip(M.m[idr], Vsp);
  if (stage1) {
    for (int p = 0; p < 5; p++) {
      L.vsm[(long)p * L.cap + q] = Vsm[p];
      L.vsp[(long)p * L.cap + q] = Vsp[p];
    }
    L.err[q] = err;
  }
  L.stats[q] = R.st.it_main;
  L.stats[L.cap + q] = R.st.fevals;
  L.stats[2L * L.cap + q] = flags;
  atomicAdd(&L.hist[min(R.st.it_main, M2CX_RIEMANN_HIST_BINS - 1)], 1ull);
  atomicAdd(&L.hist[M2CX_RIEMANN_HIST_BINS + min(R.st.fevals, M2CX_RIEMANN_HIST_BINS - 1)], 1ull);
  double f1[5], f2[5];
  int ferr = 0;
  if (s.mp_flux == M2CX_MPFLUX_EXACT) {
    flux_phys(M.m[mid], d, Vs, f1);
    for (int p = 0; p < 5; p++) f2[p] = f1[p];
  } else if (GOD) {
    // the numerical flux of each side against its star state is itself a
    // Godunov flux: two more exact solves on this lane's path slot, whose
    // statistics are not the interface problem's
    const RStats keep = R.st;
    bool o1, o2;
    if (!flux_godunov(M, R, d, Vm, Vsm, idl, f1, o1)) fatal = 1;
    if (!flux_godunov(M, R, d, Vsp, Vp, idr, f2, o2)) fatal = 1;
    ovf |= o1 || o2;
    R.st = keep;
  } else {
    flux_num(M.m[idl], s.flux, s.del, d, Vm, Vsm, f1, ferr);
    flux_num(M.m[idr], s.flux, s.del, d, Vsp, Vp, f2, ferr);
  }
  for (int p = 0; p < 5; p++) {
    L.f1[(long)p * L.cap + q] = f1[p];
    L.f2[(long)p * L.cap + q] = f2[p];
  }
  if (ferr) fatal = 1;
  return flags;
}

// One lane per interface face, the faces dealt across the waves: face q goes
// to wave q % W, lane (q / W) % 64 (W waves of 64 lanes).  The exact solver's
// work per face varies by orders of magnitude and with it its control flow
// (faces in near equilibrium run the secant loop to its iteration limit, the
// rarefaction side being flat within the integration tolerance; strong ones
// spend their time in long Hugoniot and rarefaction solves), and lanes of a
// wave that follow different paths run one after another.  Neighbouring faces
// tend to need the same kind of solve; dealt out this way every wave holds a
// small, even share of the expensive faces.  Each lane owns slot `gid` of the
// integration-path scratch (stride: all lanes of the launch).
// The launch runs beside the first half of the stage kernel on a second
// stream and its slowest wave sets its length; raising the waves' issue
// priority on the SIMDs they share with the stage kernel shortens it (C5 bubble
// cloud: 84.2 -> 80.3 ms of Riemann time per step, results unchanged).
template <bool GOD>
__global__ void __launch_bounds__(64) k_riemann(DevGrid g, Scheme s, RiemannCfg rc, CF5 V,
                                                const unsigned char* __restrict__ ID, IfaceList L, int stage1,
                                                double* slab, int lanes, int pathcap, const float* guess,
                                                DevCounters* cnt) {
  __builtin_amdgcn_s_setprio(3);
  const int W = gridDim.x, wave = blockIdx.x, lane = threadIdx.x;
  const int gid = wave * 64 + lane;
  if (gid >= lanes) return;
  const int nif = min(cnt->nif, L.cap);
  DevRiemann R(c_mats[s.mset], rc, slab + gid, lanes, pathcap);
  unsigned long long calls = 0, triv = 0, errs = 0, its = 0, fev = 0, fb = 0, gfb = 0;
  int fatal = 0, ovf = 0;
  for (long q = (long)lane * W + wave; q < nif; q += 64L * W) {
    const int f = riemann_face<GOD>(g, s, V, ID, L, (int)q, stage1, guess, R, fatal, ovf);
    calls++;
    triv += f & 1;
    errs += (f >> 1) & 1;
    fb += (f >> 2) & 1;
    gfb += (f >> 3) & 1;
    its += R.st.it_main;
    fev += R.st.fevals;
  }
  if (calls) {
    atomicAdd(&cnt->riemann_calls, calls);
    atomicAdd(&cnt->riemann_trivial, triv);
    atomicAdd(&cnt->riemann_errors, errs);
    atomicAdd(&cnt->riemann_iters, its);
    atomicAdd(&cnt->riemann_fevals, fev);
    atomicAdd(&cnt->riemann_fallbacks, fb);
    atomicAdd(&cnt->guess_fallbacks, gfb);
  }
  if (fatal) atomicOr(&cnt->error, 2);
  if (ovf) atomicOr(&cnt->error, 4);
}

// ---------------------------------------------------------------------------
// Ensemble form: the interface faces of many independent solvers in one
// launch.  Each member describes its solver; `off` holds the prefix sums of
// the members' face counts (k_riem_offsets); lanes walk the concatenated
// list, so the latency-bound exact solves of all members overlap.
// ---------------------------------------------------------------------------
struct RiemMember {
  DevGrid g;
  Scheme s;
  RiemannCfg rc;
  CF5 V;
  const unsigned char* ID;
  IfaceList L;
  const float* guess;
  DevCounters* cnt;
};

__global__ void k_riem_offsets(const RiemMember* mem, int nm, int* off) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int acc = 0;
  for (int m = 0; m < nm; m++) {
    off[m] = acc;
    acc += min(mem[m].cnt->nif, mem[m].L.cap);
  }
  off[nm] = acc;
}

// the concatenated list dealt across waves as in k_riemann
__global__ void __launch_bounds__(64) k_riemann_multi(const RiemMember* __restrict__ mem, int nm,
                                                      const int* __restrict__ off, int stage1, double* slab,
                                                      int lanes, int pathcap) {
  __builtin_amdgcn_s_setprio(3);
  const int W = gridDim.x, wave = blockIdx.x, lane = threadIdx.x;
  const int gid = wave * 64 + lane;
  if (gid >= lanes) return;
  const int total = off[nm];
  for (long q = (long)lane * W + wave; q < total; q += 64L * W) {
    int lo = 0, hi = nm - 1;  // member m with off[m] <= q < off[m+1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= q) lo = mid;
      else hi = mid - 1;
    }
    const RiemMember& B = mem[lo];
    DevRiemann R(c_mats[B.s.mset], B.rc, slab + gid, lanes, pathcap);
    int fatal = 0, ovf = 0;
    riemann_face(B.g, B.s, B.V, B.ID, B.L, (int)q - off[lo], stage1, B.guess, R, fatal, ovf);
    if (fatal) atomicOr(&B.cnt->error, 2);
    if (ovf) atomicOr(&B.cnt->error, 4);
  }
}

// per-member counters of an ensemble Riemann launch (one block per member)
__global__ void __launch_bounds__(256) k_riem_count(const RiemMember* mem) {
  const RiemMember& B = mem[blockIdx.x];
  const int nif = min(B.cnt->nif, B.L.cap);
  __shared__ unsigned long long acc[6][256];
  unsigned long long a[6] = {0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < nif; q += blockDim.x) {
    const int f = B.L.stats[2L * B.L.cap + q];
    a[0] += f & 1;
    a[1] += (f >> 1) & 1;
    a[2] += (f >> 2) & 1;
    a[3] += (f >> 3) & 1;
    a[4] += B.L.stats[q];
    a[5] += B.L.stats[B.L.cap + q];
  }
  for (int k = 0; k < 6; k++) acc[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 6; k++) acc[k][threadIdx.x] += acc[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && nif > 0) {
    atomicAdd(&B.cnt->riemann_calls, (unsigned long long)nif);
    atomicAdd(&B.cnt->riemann_trivial, acc[0][0]);
    atomicAdd(&B.cnt->riemann_errors, acc[1][0]);
    atomicAdd(&B.cnt->riemann_fallbacks, acc[2][0]);
    atomicAdd(&B.cnt->guess_fallbacks, acc[3][0]);
    atomicAdd(&B.cnt->riemann_iters, acc[4][0]);
    atomicAdd(&B.cnt->riemann_fevals, acc[5][0]);
  }
}

// point-kernel entry: n problems given explicitly
__global__ void __launch_bounds__(128) k_riemann_batch(int mset, RiemannCfg rc, int n, const double* dir, const double* Vm,
                                                       const int* idm, const double* Vp, const int* idp,
                                                       double* Vs, int* id, double* Vsm, double* Vsp, int* err,
                                                       int* iters, int* fevals, double* slab, int lanes,
                                                       int pathcap, const float* guess, DevCounters* cnt) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= lanes) return;
  DevRiemann R(c_mats[mset], rc, slab + lane, lanes, pathcap);
  for (int q = lane; q < n; q += lanes) {
    R.st = RStats{0, 0, false, false, false};
    R.path1.overflow = R.path3.overflow = false;
    int mid = idm[q];
    double p0g = -1.0;
    if (guess) {
      const float gp = guess[q];
      p0g = (isfinite(gp) && gp > 0.0f) ? (double)gp : -1.0;
    }
    err[q] = R.solve(dir + 3 * q, Vm + 5 * q, idm[q], Vp + 5 * q, idp[q], Vs + 5 * q, mid, Vsm + 5 * q, Vsp + 5 * q,
                     p0g);
    id[q] = mid;
    iters[q] = R.st.it_main;
    fevals[q] = R.st.fevals;
    if (R.st.fatal) {  // the reference exits here; the batch marks the problem instead
      err[q] = -1;
      id[q] = -1;
      for (int p = 0; p < 5; p++) Vs[5 * q + p] = Vsm[5 * q + p] = Vsp[5 * q + p] = __builtin_nan("");
    }
    if (R.path1.overflow || R.path3.overflow) {  // integration path longer than the slab holds
      err[q] = -2;
      id[q] = -1;
      for (int p = 0; p < 5; p++) Vs[5 * q + p] = Vsm[5 * q + p] = Vsp[5 * q + p] = __builtin_nan("");
    }
  }
}

// ---------------------------------------------------------------------------
// Flux = Godunov (FluxFcnGodunov.h:37-60).  Every stage, before the stage
// kernel: one lane per face whose two cells hold the same (active) material,
// faces dealt to the lanes in turn; the face states are the stage kernel's
// (face_pair), the flux goes to gflux[axis][component][plus cell].  The stage
// kernel reads it in place of a numerical flux; interface faces keep their
// multi-material treatment.
// ---------------------------------------------------------------------------
struct GodunovArgs {
  DevGrid g;
  Scheme s;
  RiemannCfg rc;
  CF5 V;
  const unsigned char* ID;
  double* gflux;
  double* slab;
  int lanes, pathcap;
  DevCounters* cnt;
};

__global__ void __launch_bounds__(64) k_godunov(GodunovArgs a) {
  const int lane = blockIdx.x * 64 + threadIdx.x;
  if (lane >= a.lanes) return;
  const DevGrid& g = a.g;
  const DevMats& M = c_mats[a.s.mset];
  DevRiemann R(M, a.rc, a.slab + lane, a.lanes, a.pathcap);
  const long nfx = (long)(g.NX + 1) * g.NY * g.NZ, nfy = (long)g.NX * (g.NY + 1) * g.NZ;
  const long nfz = (long)g.NX * g.NY * (g.NZ + 1);
  const long total = nfx + nfy + nfz;
  int fatal = 0, ovf = 0;
  for (long q = lane; q < total; q += a.lanes) {
    int D, i, j, k;
    long r = q;
    if (r < nfx) {
      D = 0; i = (int)(r % (g.NX + 1)); r /= (g.NX + 1); j = (int)(r % g.NY); k = (int)(r / g.NY);
    } else if ((r -= nfx) < nfy) {
      D = 1; i = (int)(r % g.NX); r /= g.NX; j = (int)(r % (g.NY + 1)); k = (int)(r / (g.NY + 1));
    } else {
      r -= nfy;
      D = 2; i = (int)(r % g.NX); r /= g.NX; j = (int)(r % g.NY); k = (int)(r / g.NY);
    }
    const long c = gidx(g, i, j, k);
    const int myid = a.ID[c], nid = a.ID[c - axis_stride(g, D)];
    if (nid != myid || myid == a.s.inactive) continue;
    double Vm[5], Vp[5], f[5];
    int bm, bp;
    face_pair<true>(g, a.s, M, a.V, a.ID, D, i, j, k, Vm, Vp, bm, bp);
    bool o;
    if (!flux_godunov(M, R, D, Vm, Vp, myid, f, o)) fatal = 1;
    ovf |= o;
    for (int p = 0; p < 5; p++) a.gflux[((long)D * 5 + p) * g.ncells + c] = f[p];
  }
  if (fatal) atomicOr(&a.cnt->error, 2);
  if (ovf) atomicOr(&a.cnt->error, 4);
}

// m2cx_numerical_flux_batch with Flux = Godunov
__global__ void __launch_bounds__(64) k_godunov_batch(int mset, RiemannCfg rc, int n, int d, const double* Vm,
                                                      const double* Vp, const int* id, double* F, double* slab,
                                                      int lanes, int pathcap, DevCounters* cnt) {
  const int lane = blockIdx.x * 64 + threadIdx.x;
  if (lane >= lanes) return;
  const DevMats& M = c_mats[mset];
  DevRiemann R(M, rc, slab + lane, lanes, pathcap);
  int fatal = 0, ovf = 0;
  for (int q = lane; q < n; q += lanes) {
    bool o;
    if (!flux_godunov(M, R, d, Vm + 5 * q, Vp + 5 * q, id[q], F + 5 * q, o)) fatal = 1;
    ovf |= o;
  }
  if (fatal) atomicOr(&cnt->error, 2);
  if (ovf) atomicOr(&cnt->error, 4);
}

// ---------------------------------------------------------------------------
// level set: FD3 residual (LevelSetOperator.cpp:796-870) fused with the RK
// combination of the stage (TimeIntegrator.cpp:376-430, 581-690)
// ---------------------------------------------------------------------------
struct LSArgs {
  DevGrid g;
  LSFvm fv;
  CF5 V;
  const double* phi_st;   // stencil field of this stage
  const double* phi_a;    // phi(n)
  const double* phi_b;    // phi1 (own-cell term)
  double* out;
  double* rls;
  const double* fd3;
  long ocxL, ocxR, ocyL, ocyR3, ocyR1, oczL, oczR;
  int which;
  const DevCounters* cnt;
};

__device__ __forceinline__ double ls_fvm_rhs(const LSArgs& a, const double* __restrict__ P, long c, int i, int j,
                                             int k);

__global__ void k_levelset(LSArgs a) {
  const DevGrid& g = a.g;
  long n = (long)g.NX * g.NY * g.NZ;
  const double dt = a.cnt->dt;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    const double* P = a.phi_st;
    const long sx = 1, sy = g.SX, sz = g.S;
    auto fd = [&](long off, int idx, long st, long c0) {
      const double* cc = a.fd3 + off + 4L * idx;
      return cc[0] * P[c0] + cc[1] * P[c0 + st] + cc[2] * P[c0 + 2 * st] + cc[3] * P[c0 + 3 * st];
    };
    const double* gx = g_axis(g, 0);
    const double* gy = g_axis(g, 1);
    const double* gz = g_axis(g, 2);
    const double xi = gx[i + 1], yj = gy[j + 1], zk = gz[k + 1];
    double r;
    if (a.fv.on) {
      r = ls_fvm_rhs(a, P, c, i, j, k);
    } else {
    double av = (i - 2 >= -1) ? fd(a.ocxL, i, sx, c - 2 * sx) : (P[c] - P[c - sx]) / (xi - gx[i]);
    double bv = (i + 2 <= g.NX) ? fd(a.ocxR, i, sx, c - sx) : (P[c + sx] - P[c]) / (gx[i + 2] - xi);
    double cv = (j - 2 >= -1) ? fd(a.ocyL, j, sy, c - 2 * sy) : (P[c] - P[c - sy]) / (yj - gy[j]);
    double dv = (j + 2 <= g.NY) ? fd((i < g.NY - 1) ? a.ocyR3 : a.ocyR1, j, sy, c - sy)
                                : (P[c + sy] - P[c]) / (gy[j + 2] - yj);
    double ev = (k - 2 >= -1) ? fd(a.oczL, k, sz, c - 2 * sz) : (P[c] - P[c - sz]) / (zk - gz[k]);
    double fv = (k + 2 <= g.NZ) ? fd(a.oczR, k, sz, c - sz) : (P[c + sz] - P[c]) / (gz[k + 2] - zk);
    const double u = a.V(1, c), v = a.V(2, c), w = a.V(3, c);
    r = (u >= 0 ? av * u : bv * u) + (v >= 0 ? cv * v : dv * v) + (w >= 0 ? ev * w : fv * w);
    r = r * -1.0;
    }
    double o;
    switch (a.which) {
      case 0: o = 1.0 * a.phi_a[c] + dt * r; break;                                    // FE
      case 1: o = 1.0 * a.phi_a[c] + dt * r; break;                                    // RK stage 1
      case 2: o = 0.5 * a.phi_a[c] + 0.5 * a.phi_b[c]; o = 1.0 * o + 0.5 * dt * r; break;  // RK2 final
      case 3: o = 0.25 * a.phi_b[c] + 0.75 * a.phi_a[c]; o = 1.0 * o + 0.25 * dt * r; break;  // RK3 stage 2
      case 4: o = 1.0 / 3.0 * a.phi_a[c] + 2.0 / 3.0 * a.phi_b[c]; o = 1.0 * o + 2.0 / 3.0 * dt * r; break;
      default: o = 0.0;
    }
    if (a.which < 0) a.rls[c] = r;
    else a.out[c] = o;
  }
}

// ---------------------------------------------------------------------------
// level-set finite-volume residual (LevelSetOperator::ComputeResidualFVM,
// LevelSetOperator.cpp:966-1300)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ls_sigma(const LSArgs& a, int D, int idx, double dq0, double dq1) {
  if (a.fv.rec != M2CX_REC_LINEAR) return 0.0;
  const double A = g_coefA(a.g, D)[idx], B = g_coefB(a.g, D)[idx];
  if (a.fv.limiter == M2CX_LIM_GENERALIZED_MINMOD) return lim_gmm(A, B, a.fv.gmm, dq0, dq1);
  if (a.fv.limiter == M2CX_LIM_VANALBADA) {
    long off = 0;
    if (D >= 1) off += a.g.NX;
    if (D >= 2) off += a.g.NY;
    return lim_va(A, B, a.fv.lsk[off + idx], dq0, dq1);
  }
  return 0.5 * (dq0 + dq1);
}

// minus and plus face values of an interior cell along axis D
// (ReconstructIn1D for the velocity component, Reconstruct for phi)
struct LSFaces { double um, up, pm, pp, slope; };
__device__ __forceinline__ LSFaces ls_faces(const LSArgs& a, const double* __restrict__ P, int D, long q, int idx) {
  const long st = axis_stride(a.g, D);
  LSFaces f;
  const double u0 = a.V(1 + D, q - st), u1 = a.V(1 + D, q), u2 = a.V(1 + D, q + st);
  const double su = ls_sigma(a, D, idx, u1 - u0, u2 - u1);
  f.um = u1 - 0.5 * su;
  f.up = u1 + 0.5 * su;
  f.slope = su / g_axis(a.g, 3 + D)[idx + 1];
  const double p0 = P[q - st], p1 = P[q], p2 = P[q + st];
  const double sp = ls_sigma(a, D, idx, p1 - p0, p2 - p1);
  f.pm = p1 - 0.5 * sp;
  f.pp = p1 + 0.5 * sp;
  return f;
}

// the face of a ghost cell towards the domain, by the mesh boundary
// condition (Reconstructor.cpp:807-880, 1048-1140): Dirichlet sides take the
// ghost value, outlets the image's face on the boundary, walls and symmetry
// planes that face with the velocity negated; `img` is the image's face
__device__ __forceinline__ void ls_ghost_face(const LSArgs& a, const double* __restrict__ P, int side, long qg,
                                              double img_u, double img_p, double& u, double& p) {
  const int D = side / 2;
  switch (a.g.bc[side]) {
    case M2CX_BC_INLET: case M2CX_BC_INLET2: case M2CX_BC_OVERSET:
      u = a.V(1 + D, qg);
      p = P[qg];
      break;
    case M2CX_BC_OUTLET:
      u = img_u;
      p = img_p;
      break;
    default:
      u = -img_u;
      p = img_p;
      break;
  }
}

// ComputeLocalAdvectionFlux (LevelSetOperator.cpp:1229-1262)
__device__ __forceinline__ double ls_face_flux(const LSArgs& a, double phim, double phip, double um, double up) {
  if (a.fv.flux == M2CX_LSFLUX_UPWIND) {
    const double vmid = 0.5 * (um + up);
    if (vmid > 0) return phim * um;
    if (vmid < 0) return phip * up;
    return 0.5 * (phim * um + phip * up);
  }
  const double fu = fabs(um), fp = fabs(up);
  const double mx = (fu < fp) ? fp : fu;  // std::max
  if (a.fv.flux == M2CX_LSFLUX_LLF) return 0.5 * (phim * um + phip * up) - 0.5 * mx * (phip - phim);
  double lam = fabs(0.5 * (um + up));
  const double tol = mx * a.fv.delta;
  if (lam < tol) lam = (lam * lam + tol * tol) / (2 * tol);
  return 0.5 * (phim * um + phip * up) - 0.5 * lam * (phip - phim);
}

// Right-hand side at one interior cell: its six face fluxes added in the
// order the reference's face loop accumulates them (own minus faces x, y, z,
// then the plus faces as the loop reaches the neighbours), then the source
// phi (u_x + v_y + w_z) of AddSourceTerm (:1266-1285).
__device__ __forceinline__ double ls_fvm_rhs(const LSArgs& a, const double* __restrict__ P, long c, int i, int j,
                                             int k) {
  const int ijk[3] = {i, j, k};
  const int Ns[3] = {a.g.NX, a.g.NY, a.g.NZ};
  double fm[3], fp[3], sl[3];
  for (int D = 0; D < 3; D++) {
    const long st = axis_stride(a.g, D);
    const double* dxs = g_axis(a.g, 3 + D);
    const int q = ijk[D];
    const LSFaces me = ls_faces(a, P, D, c, q);
    sl[D] = me.slope;
    double um, pm;  // plus face of the minus neighbour
    if (q == 0) ls_ghost_face(a, P, 2 * D, c - st, me.um, me.pm, um, pm);
    else {
      const LSFaces nb = ls_faces(a, P, D, c - st, q - 1);
      um = nb.up;
      pm = nb.pp;
    }
    fm[D] = ls_face_flux(a, pm, me.pm, um, me.um) / dxs[q + 1];
    double up, pp;  // minus face of the plus neighbour
    if (q == Ns[D] - 1) ls_ghost_face(a, P, 2 * D + 1, c + st, me.up, me.pp, up, pp);
    else {
      const LSFaces nb = ls_faces(a, P, D, c + st, q + 1);
      up = nb.um;
      pp = nb.pm;
    }
    fp[D] = ls_face_flux(a, me.pp, pp, me.up, up) / dxs[q + 2];
  }
  double r = 0.0;
  r += fm[0];
  r += fm[1];
  r += fm[2];
  r -= fp[0];
  r -= fp[1];
  r -= fp[2];
  r += P[c] * (sl[0] + sl[1] + sl[2]);
  return r;
}

// LevelSetOperator::ApplyBoundaryConditions (LevelSetOperator.cpp:408-508);
// LevelSetReinitializer::ApplyBoundaryConditions (LevelSetReinitializer.cpp:
// 1249-1360) fills the same face ghosts by the same rules, with the narrow
// band only those `useful` marks.
__device__ void ls_bc_faces(const DevGrid& g, double* phi, const int* bcs,
                            const unsigned char* __restrict__ useful = nullptr) {
  long nfaces = 2L * ((long)g.NY * g.NZ + (long)g.NX * g.NZ + (long)g.NX * g.NY);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nfaces; q += (long)gridDim.x * blockDim.x) {
    long r = q;
    int side, i, j, k;
    long nyz = (long)g.NY * g.NZ, nxz = (long)g.NX * g.NZ, nxy = (long)g.NX * g.NY;
    if (r < 2 * nyz) { side = (int)(r / nyz); r %= nyz; j = r % g.NY; k = r / g.NY; i = side == 0 ? -1 : g.NX; }
    else if ((r -= 2 * nyz) < 2 * nxz) { side = 2 + (int)(r / nxz); r %= nxz; i = r % g.NX; k = r / g.NX; j = side == 2 ? -1 : g.NY; }
    else { r -= 2 * nxz; side = 4 + (int)(r / nxy); r %= nxy; i = r % g.NX; j = r / g.NX; k = side == 4 ? -1 : g.NZ; }
    int ii = i, jj = j, kk = k;
    if (side == 0) ii = 0; else if (side == 1) ii = g.NX - 1;
    else if (side == 2) jj = 0; else if (side == 3) jj = g.NY - 1;
    else if (side == 4) kk = 0; else kk = g.NZ - 1;
    const int bc = bcs[side];
    long c = gidx(g, i, j, k);
    if (useful && !useful[c]) continue;
    double im = phi[gidx(g, ii, jj, kk)];
    double ph = phi[c];
    if (bc == M2CX_LSBC_ZERO_NEUMANN) {
      ph = im;
    } else if (bc == M2CX_LSBC_LINEAR_EXTRAPOLATION || bc == M2CX_LSBC_NON_NEGATIVE) {
      const int ax = side / 2;
      const int N = ax == 0 ? g.NX : (ax == 1 ? g.NY : g.NZ);
      const int a = ax == 0 ? i : (ax == 1 ? j : k);
      const int dirn = (side % 2 == 0) ? 1 : -1;
      const double* X = g_axis(g, ax);
      const long st = ax == 0 ? 1 : (ax == 1 ? g.SX : g.S);
      bool ok = (side % 2 == 0) ? (a + 2 < N) : (a - 2 >= 0);
      if (ok) {
        double rr = X[a + 1], r1 = X[a + dirn + 1], r2 = X[a + 2 * dirn + 1];
        double f1 = phi[c + dirn * st], f2 = phi[c + 2 * dirn * st];
        ph = f1 + (f2 - f1) / (r2 - r1) * (rr - r1);
      } else {
        ph = im;
      }
      if (bc == M2CX_LSBC_NON_NEGATIVE) {
        const int ai = ax == 0 ? ii : (ax == 1 ? jj : kk);
        double d2wall = 0.5 * fabs(X[ai + 1] - X[a + 1]);
        double m = -d2wall;
        ph = (m < ph) ? ph : m;  // std::max(-d2wall, phi)
      }
    }
    phi[c] = ph;
  }
}

__global__ void k_ls_bc(DevGrid g, double* phi, int b0, int b1, int b2, int b3, int b4, int b5) {
  const int bcs[6] = {b0, b1, b2, b3, b4, b5};
  ls_bc_faces(g, phi, bcs);
}

// ---------------------------------------------------------------------------
// end of step
// ---------------------------------------------------------------------------
__global__ void k_end_reset(DevCounters* cnt) {
  cnt->n_remaining = 0;
  cnt->n_changed = 0;
  cnt->n_unresolved = 0;
}

// every level set of a solver, for the routines that look at all of them
struct LSPtrs {
  double* phi[M2CX_MAX_LEVELSETS];
  const double* phin[M2CX_MAX_LEVELSETS];  // Phi_tmp: the level sets at the start of the step
  int mat[M2CX_MAX_LEVELSETS];
  int n;
  unsigned tracked;  // bit m: material m has a level set
};

// UpdateMaterialIDByLevelSet, sign-change part (MultiPhaseOperator.cpp:177-221):
// a node whose phi changed sign or reached zero for some level sets takes
// the material of the one it is now inside (two: error bit 16, the
// reference exits); if it is inside none it is left to the vote
__global__ void k_id_sweep(DevGrid g, Scheme s, LSPtrs P, unsigned char* ID, unsigned char* tag, int* remaining,
                           int cap, DevCounters* cnt) {
  long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    if (ID[c] == s.inactive) continue;
    bool swept = false;
    int myls = -1, inside = 0;
    for (int l = 0; l < P.n; l++) {
      const double a = P.phi[l][c], b = P.phin[l][c];
      if (a * b > 0.0) continue;
      swept = true;
      if (a < 0) {
        inside++;
        myls = l;
      }
    }
    if (!swept) continue;
    if (inside > 1) {
      atomicOr(&cnt->error, 16);
      continue;
    }
    if (myls >= 0) {
      ID[c] = (unsigned char)P.mat[myls];
    } else {
      tag[c] = 1;
      int slot = atomicAdd(&cnt->n_remaining, 1);
      if (slot < cap) remaining[slot] = (int)c;
      else atomicOr(&cnt->error, 4);
    }
  }
}

// Neighbour vote for swept cells outside every level-set subdomain
// (MultiPhaseOperator.cpp:236-338), Jacobi sweeps in one work-group.  With a
// single material not tracked by a level set every candidate set is that
// material or empty, so the result is the same as the reference's in-order
// sweep; with several untracked materials a node decided in the same sweep
// is seen by its neighbours one sweep later than in the reference.
__global__ void __launch_bounds__(1024) k_id_vote(DevGrid g, Scheme s, unsigned char* ID, unsigned char* tag,
                                                  const int* remaining, int cap, unsigned tracked,
                                                  DevCounters* cnt) {
  __shared__ int progress, left;
  const int nrem = min(cnt->n_remaining, cap);
  for (int iter = 0; iter < 100; iter++) {
    if (threadIdx.x == 0) { progress = 0; left = 0; }
    __syncthreads();
    // phase 1: find candidates (read only)
    for (int q = threadIdx.x; q < nrem; q += blockDim.x) {
      long c = remaining[q];
      if (!tag[c]) continue;
      atomicAdd(&left, 1);
    }
    __syncthreads();
    if (left == 0) break;
    // candidate of each remaining node into the high bits of its tag
    for (int q = threadIdx.x; q < nrem; q += blockDim.x) {
      long c = remaining[q];
      if (!(tag[c] & 1)) continue;
      long r = c;
      int i = (int)(r % g.SX) - 1; r /= g.SX;
      int j = (int)(r % g.SY) - 1;
      int k = (int)(r / g.SY) - 1;
      int best = 256;
      for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
          for (int di = -1; di <= 1; di++) {
            int ni = i + di, nj = j + dj, nk = k + dk;
            if (ni < 0 || ni >= g.NX || nj < 0 || nj >= g.NY || nk < 0 || nk >= g.NZ) continue;
            long nc = gidx(g, ni, nj, nk);
            if (tag[nc] & 1) continue;
            int nid = ID[nc];
            if (nid == s.inactive || ((tracked >> nid) & 1u)) continue;
            if (nid < best) best = nid;
          }
      if (best < 256) tag[c] = (unsigned char)(1 | 2);  // decided this sweep
      if (best < 256) {
        // remember the candidate: the node's own ID is not read by others
        // while it is tagged, so it can hold it now
        ID[c] = (unsigned char)best;
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nrem; q += blockDim.x) {
      long c = remaining[q];
      if (tag[c] == 3) {
        tag[c] = 0;
        atomicAdd(&progress, 1);
      }
    }
    __syncthreads();
    if (progress == 0) {
      for (int q = threadIdx.x; q < nrem; q += blockDim.x) {
        long c = remaining[q];
        if (tag[c] & 1) ID[c] = 0;  // orphans (MultiPhaseOperator.cpp:323-333)
      }
      break;
    }
    __syncthreads();
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nrem; q += blockDim.x) tag[remaining[q]] = 0;
}

// The same vote in the reference's order (MultiPhaseOperator.cpp:236-338),
// for when two or more materials have no level set and the order can change
// the result: the remaining nodes sorted as its std::set<Int3> (i, then j,
// then k), taken by one thread; a node decided in a sweep is seen by the
// nodes after it in that sweep.
__global__ void k_id_vote_seq(DevGrid g, Scheme s, unsigned char* ID, unsigned char* tag, const int* __restrict__ list,
                              int n, unsigned tracked) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int iter = 0; iter < 100; iter++) {
    int left = 0;
    for (int t = 0; t < n; t++) left += tag[list[t]] & 1;
    if (left == 0) break;
    int progress = 0;
    for (int t = 0; t < n; t++) {
      const long c = list[t];
      if (!(tag[c] & 1)) continue;
      long r = c;
      const int i = (int)(r % g.SX) - 1;
      r /= g.SX;
      const int j = (int)(r % g.SY) - 1;
      const int k = (int)(r / g.SY) - 1;
      int best = 256;
      for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
          for (int di = -1; di <= 1; di++) {
            const int ni = i + di, nj = j + dj, nk = k + dk;
            if (ni < 0 || ni >= g.NX || nj < 0 || nj >= g.NY || nk < 0 || nk >= g.NZ) continue;
            const long nc = gidx(g, ni, nj, nk);
            if (tag[nc] & 1) continue;
            const int nid = ID[nc];
            if (nid == s.inactive || ((tracked >> nid) & 1u)) continue;
            if (nid < best) best = nid;
          }
      if (best == 256) continue;
      ID[c] = (unsigned char)best;
      tag[c] = 0;
      progress++;
    }
    if (progress == 0) {
      for (int t = 0; t < n; t++)
        if (tag[list[t]] & 1) ID[list[t]] = 0;  // orphans (MultiPhaseOperator.cpp:323-333)
      break;
    }
  }
  for (int t = 0; t < n; t++) tag[list[t]] = 0;
}

// ResolveConflictsInLevelSets, part I (MultiPhaseOperator.cpp:1705-1762).
// Only a node inside two or more subdomains can be changed, so those are
// collected first (any order; the host sorts them); one thread then takes
// them in the reference's (k, j, i) order, since a node's update changes the
// sign tests of its neighbours, which the reference sees in that order.
__global__ void k_lsr_reset(DevCounters* cnt) {
  cnt->n_lsr = 0;
  cnt->n_resolved = 0;
}

__global__ void k_lsr_mark(DevGrid g, LSPtrs P, int* list, int cap, DevCounters* cnt) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < g.ncells; q += (long)gridDim.x * blockDim.x) {
    int neg = 0;
    for (int l = 0; l < P.n; l++) neg += P.phi[l][q] < 0;
    if (neg < 2) continue;
    const int slot = atomicAdd(&cnt->n_lsr, 1);
    if (slot < cap) list[slot] = (int)q;
  }
}

// Part I at one node of the ghosted brick; returns 1 when it resolved a node
// inside several subdomains (counted).  With the optional corrections (opt)
// a node next to two or more boundaries inside one subdomain, or on the
// boundary (phi = 0) of some, has all of those level sets set to
// +-mean|phi| over them, uncounted (MultiPhaseOperator.cpp:1766-1799).
__device__ int lsr_part1_node(const DevGrid& g, const LSPtrs& P, long c, int i, int j, int k, bool opt) {
  int bnd[M2CX_MAX_LEVELSETS];
  int nb = 0, no = 0, ni = 0, owner0 = -1, inter0 = -1;
  for (int l = 0; l < P.n; l++) {
    const double* ph = P.phi[l];
    const double p = ph[c];
    const bool b = (i - 1 >= -1 && p * ph[c - 1] <= 0) || (i + 1 <= g.NX && p * ph[c + 1] <= 0) ||
                   (j - 1 >= -1 && p * ph[c - g.SX] <= 0) || (j + 1 <= g.NY && p * ph[c + g.SX] <= 0) ||
                   (k - 1 >= -1 && p * ph[c - g.S] <= 0) || (k + 1 <= g.NZ && p * ph[c + g.S] <= 0);
    if (!b) continue;
    bnd[nb++] = l;
    if (p < 0) {
      if (no++ == 0) owner0 = l;
    } else if (p == 0) {
      if (ni++ == 0) inter0 = l;
    }
  }
  if (nb <= 1 || (no == 0 && ni == 0)) return 0;
  if (no <= 1) {
    if (!opt) return 0;
    double new_phi = 0.0;
    for (int w = 0; w < nb; w++) new_phi += fabs(P.phi[bnd[w]][c]);
    new_phi /= (double)nb;
    for (int w = 0; w < nb; w++) {
      double* ph = P.phi[bnd[w]];
      const bool own = (no == 0) ? bnd[w] == inter0 : ph[c] < 0;
      ph[c] = own ? -new_phi : new_phi;
    }
    return 0;
  }
  // several owners: the largest |phi| (the first on ties) keeps the node
  int new_owner = owner0;
  double max_phi = fabs(P.phi[owner0][c]);
  double new_phi = 0.0;
  for (int w = 0; w < nb; w++) {
    const double p = P.phi[bnd[w]][c];
    if (!(p < 0)) continue;
    if (fabs(p) > max_phi) {
      new_owner = bnd[w];
      max_phi = fabs(p);
    }
  }
  for (int w = 0; w < nb; w++) {
    const double p = P.phi[bnd[w]][c];
    if (p < 0) new_phi += fabs(p);
  }
  new_phi /= (double)no;
  for (int w = 0; w < nb; w++) {
    double* ph = P.phi[bnd[w]];
    if (ph[c] < 0) ph[c] = (bnd[w] == new_owner) ? -new_phi : new_phi;
  }
  return 1;
}

__global__ void k_lsr_resolve(DevGrid g, LSPtrs P, const int* __restrict__ list, int n, DevCounters* cnt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int resolved = 0;
  for (int t = 0; t < n; t++) {
    const long c = list[t];
    const unsigned u = (unsigned)c, sx = (unsigned)g.SX, sy = (unsigned)g.SY;
    const unsigned rr = u / sx;
    const int i = (int)(u - rr * sx) - 1;
    const int k = (int)(rr / sy) - 1;
    const int j = (int)(rr - (unsigned)(k + 1) * sy) - 1;
    resolved += lsr_part1_node(g, P, c, i, j, k, false);
  }
  cnt->n_resolved = resolved;
}

// The optional corrections in the reference's (k, j, i) order, one launch per
// plane i + j + k = t: a node reads only itself and its six face neighbours;
// those before it in (k, j, i) order lie on plane t - 1 (updated by the
// previous launch), those after it on plane t + 1 (not yet), and no two nodes
// of one plane are face neighbours, so a plane's nodes run in parallel with
// the results of the sequential sweep.  Part I over the ghosted brick:
__global__ void k_lsr_plane1(DevGrid g, LSPtrs P, int t, DevCounters* cnt) {
  const long m = (long)(g.NY + 2) * (g.NZ + 2);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < m; q += (long)gridDim.x * blockDim.x) {
    const int j = (int)(q % (g.NY + 2)) - 1, k = (int)(q / (g.NY + 2)) - 1;
    const int i = t - j - k;
    if (i < -1 || i > g.NX) continue;
    if (lsr_part1_node(g, P, gidx(g, i, j, k), i, j, k, true)) atomicAdd(&cnt->n_resolved, 1);
  }
}

// Part II over the interior (MultiPhaseOperator.cpp:1836-1968): a background
// node (phi >= 0 for every level set) with at most one background face
// neighbour inside the domain goes to the level set with the smallest phi
// (the first on ties), which becomes -phi.
__device__ __forceinline__ bool lsr_background(const LSPtrs& P, long c) {
  for (int l = 0; l < P.n; l++)
    if (P.phi[l][c] < 0) return false;
  return true;
}

__global__ void k_lsr_plane2(DevGrid g, LSPtrs P, int t, DevCounters* cnt) {
  const long m = (long)g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < m; q += (long)gridDim.x * blockDim.x) {
    const int j = (int)(q % g.NY), k = (int)(q / g.NY);
    const int i = t - j - k;
    if (i < 0 || i >= g.NX) continue;
    const long c = gidx(g, i, j, k);
    if (!lsr_background(P, c)) continue;
    int qi = 0;
    if (i - 1 >= 0 && lsr_background(P, c - 1)) qi++;
    if (i + 1 < g.NX && lsr_background(P, c + 1)) qi++;
    if (j - 1 >= 0 && lsr_background(P, c - g.SX)) qi++;
    if (j + 1 < g.NY && lsr_background(P, c + g.SX)) qi++;
    if (k - 1 >= 0 && lsr_background(P, c - g.S)) qi++;
    if (k + 1 < g.NZ && lsr_background(P, c + g.S)) qi++;
    if (qi >= 2) continue;
    double min_phi = DBL_MAX;
    int new_owner = -1;
    for (int l = 0; l < P.n; l++)
      if (P.phi[l][c] < min_phi) {
        new_owner = l;
        min_phi = P.phi[l][c];
      }
    if (new_owner < 0) continue;
    P.phi[new_owner][c] = -min_phi;
    atomicAdd(&cnt->n_resolved, 1);
  }
}

// UpdateMaterialIDAtGhostNodes (MultiPhaseOperator.cpp:919-942)
__global__ void k_ghost_ids(DevGrid g, unsigned char* ID) {
  long nyz = (long)g.NY * g.NZ, nxz = (long)g.NX * g.NZ, nxy = (long)g.NX * g.NY;
  long nfaces = 2 * (nyz + nxz + nxy);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nfaces; q += (long)gridDim.x * blockDim.x) {
    long r = q;
    int i, j, k, ii, jj, kk;
    if (r < 2 * nyz) { int sd = (int)(r / nyz); r %= nyz; j = r % g.NY; k = r / g.NY; i = sd ? g.NX : -1; ii = sd ? g.NX - 1 : 0; jj = j; kk = k; }
    else if ((r -= 2 * nyz) < 2 * nxz) { int sd = (int)(r / nxz); r %= nxz; i = r % g.NX; k = r / g.NX; j = sd ? g.NY : -1; jj = sd ? g.NY - 1 : 0; ii = i; kk = k; }
    else { r -= 2 * nxz; int sd = (int)(r / nxy); r %= nxy; i = r % g.NX; j = r / g.NX; k = sd ? g.NZ : -1; kk = sd ? g.NZ - 1 : 0; ii = i; jj = j; }
    ID[gidx(g, i, j, k)] = ID[gidx(g, ii, jj, kk)];
  }
}

__global__ void k_changed(DevGrid g, Scheme s, const unsigned char* __restrict__ ID,
                          const unsigned char* __restrict__ IDn, int* changed, int cap, DevCounters* cnt) {
  long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    if (ID[c] == IDn[c] || ID[c] == s.inactive) continue;
    int slot = atomicAdd(&cnt->n_changed, 1);
    if (slot < cap) changed[slot] = (int)c;
    else atomicOr(&cnt->error, 4);
  }
}

struct UpdArgs {
  DevGrid g;
  Scheme s;
  F5 V;
  const unsigned char* ID;
  const unsigned char* IDn;
  const int* fidx[3];
  IfaceList L;
  const int* changed;
  int* unresolved;
  unsigned char* done;
  int cap;
  DevCounters* cnt;
};

// MultiPhaseOperator::LocalUpdateByRiemannSolutions (:1039-1150)
__device__ bool local_update(const UpdArgs& a, long c, int id, bool upwind) {
  const DevGrid& g = a.g;
  const long st[3] = {1, g.SX, g.S};
  int counter = 0;
  double sum_weight = 0.0;
  double v[5];
  for (int p = 0; p < 5; p++) v[p] = a.V(p, c);
  long r = c;
  const int i = (int)(r % g.SX) - 1; r /= g.SX;
  const int j = (int)(r % g.SY) - 1;
  const int k = (int)(r / g.SY) - 1;
  const int ijk[3] = {i, j, k};
  const int N[3] = {g.NX, g.NY, g.NZ};
  const bool llf = a.s.mp_flux == M2CX_MPFLUX_LLF;
  for (int q = 0; q < 6; q++) {
    const int d = q / 2;
    const bool plus = q & 1;          // left/bottom/back: minus side of c; right/top/front: plus side
    const long nb = plus ? c + st[d] : c - st[d];
    const long fc = plus ? nb : c;    // plus-side cell of the face (owner of the slot)
    const int fa = plus ? ijk[d] + 1 : ijk[d];
    if (llf || fa < 1 || fa > N[d] - 1) continue;   // interior faces only hold Riemann solutions
    const int idm = a.IDn[fc - st[d]], idp = a.IDn[fc];
    if (idm == idp || idm == a.s.inactive || idp == a.s.inactive) continue;
    const int slot = a.fidx[d][fc];
    if (a.L.err[slot] != 0) continue;
    const int sid = plus ? idp : idm;                // ID stored with the star state
    if (sid != id) continue;
    const double vn = a.V(d + 1, nb);
    if (upwind && !(plus ? (vn < 0) : (vn > 0))) continue;
    const double* sv = plus ? a.L.vsp : a.L.vsm;
    double weight = 1.0;
    if (upwind) {
      const double u1 = a.V(1, nb), u2 = a.V(2, nb), u3 = a.V(3, nb);
      const double nrm = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
      weight = (plus ? -vn : vn) / nrm;
    }
    sum_weight += weight;
    if (counter == 0)
      for (int p = 0; p < 5; p++) v[p] = weight * sv[(long)p * a.L.cap + slot];
    else
      for (int p = 0; p < 5; p++) v[p] += weight * sv[(long)p * a.L.cap + slot];
    counter++;
  }
  if (sum_weight > 0.0) {
    const double cst = 1.0 / sum_weight;
    for (int p = 0; p < 5; p++) v[p] *= cst;
  }
  if (counter > 0)
    for (int p = 0; p < 5; p++) a.V(p, c) = v[p];
  return counter > 0;
}

// UpdateStateVariablesByRiemannSolutions (:980-1034) in dependency rounds:
// the reference sweeps cells in (k,j,i) order updating V in place, so a cell
// sees new values only at its i-1, j-1, k-1 neighbours.  A cell is processed
// once those of them that also changed are done; every other neighbour it
// reads is still unmodified, exactly as in the sweep.
__global__ void __launch_bounds__(1024) k_state_update(UpdArgs a) {
  __shared__ int left;
  const DevGrid& g = a.g;
  const int nch = min(a.cnt->n_changed, a.cap);
  const long st[3] = {1, g.SX, g.S};
  for (int round = 0; round < 1 << 20; round++) {
    if (threadIdx.x == 0) left = 0;
    __syncthreads();
    // phase 1: who is ready (reads done flags only)
    constexpr int PER = 16;
    int ready[PER];
    int nready = 0;
    for (int q = threadIdx.x; q < nch; q += blockDim.x) {
      long c = a.changed[q];
      if (a.done[c]) continue;
      atomicAdd(&left, 1);
      bool ok = true;
      for (int d = 0; d < 3; d++) {
        long nb = c - st[d];
        if (a.ID[nb] != a.IDn[nb] && a.ID[nb] != a.s.inactive && !a.done[nb]) {
          // only interior neighbours are swept by the reference's loop
          long r = nb;
          int ii = (int)(r % g.SX) - 1; r /= g.SX;
          int jj = (int)(r % g.SY) - 1;
          int kk = (int)(r / g.SY) - 1;
          if (ii >= 0 && jj >= 0 && kk >= 0) ok = false;
        }
      }
      if (ok && nready < PER) ready[nready++] = q;
    }
    __syncthreads();
    if (left == 0) break;
    // phase 2: process the ready ones
    for (int r = 0; r < nready; r++) {
      long c = a.changed[ready[r]];
      int id = a.ID[c];
      bool res = local_update(a, c, id, true);
      if (!res) res = local_update(a, c, id, false);
      if (!res) {
        int slot = atomicAdd(&a.cnt->n_unresolved, 1);
        if (slot < a.cap) a.unresolved[slot] = (int)c;
      }
      a.done[c] = 1;
    }
    __threadfence_block();
    __syncthreads();
  }
}

// UpdateStateVariablesByExtrapolation (:1155-1272), upwind weights: a cell
// that changed material takes the average of its 3x3x3 neighbours of the same
// material that kept it, each weighted by how much its flow points at the
// cell (max(0, unit offset . unit velocity)); the neighbours never change in
// this pass, so the cells are independent.  Cells with no weight go to the
// unresolved list for FixUnresolvedNodes.
__global__ void k_extrapolate(UpdArgs a) {
  const DevGrid& g = a.g;
  const int nch = min(a.cnt->n_changed, a.cap);
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nch; q += gridDim.x * blockDim.x) {
    const long c = a.changed[q];
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    const int myid = a.ID[c];
    const double x0[3] = {X[i + 1], Y[j + 1], Z[k + 1]};
    double sw = 0.0, vs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int nk = k - 1; nk <= k + 1; nk++)
      for (int nj = j - 1; nj <= j + 1; nj++)
        for (int ni = i - 1; ni <= i + 1; ni++) {
          const long nb = gidx(g, ni, nj, nk);
          if (a.ID[nb] != myid) continue;
          if (a.ID[nb] != a.IDn[nb]) continue;
          if (ni < 0 || ni >= g.NX || nj < 0 || nj >= g.NY || nk < 0 || nk >= g.NZ) continue;
          double w;
          if (a.s.phase_dir == M2CX_PHASEDIR_ALL) {  // inverse distance (:1220-1221)
            const double d[3] = {x0[0] - X[ni + 1], x0[1] - Y[nj + 1], x0[2] - Z[nk + 1]};
            w = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          } else {
            double v1[3] = {a.V(1, nb), a.V(2, nb), a.V(3, nb)};
            const double v1n = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
            if (v1n != 0) {
              const double inv = 1.0 / v1n;  // Vec3D::operator/= multiplies by the reciprocal
              v1[0] *= inv; v1[1] *= inv; v1[2] *= inv;
            }
            double d[3] = {x0[0] - X[ni + 1], x0[1] - Y[nj + 1], x0[2] - Z[nk + 1]};
            const double inv = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            d[0] *= inv; d[1] *= inv; d[2] *= inv;
            const double dot = d[0] * v1[0] + d[1] * v1[1] + d[2] * v1[2];
            w = (0.0 < dot) ? dot : 0.0;  // std::max(0.0, dot)
          }
          if (w > 0) {
            sw += w;
            for (int p = 0; p < 5; p++) vs[p] += w * a.V(p, nb);
          }
        }
    if (sw == 0) {
      const int slot = atomicAdd(&a.cnt->n_unresolved, 1);
      if (slot < a.cap) a.unresolved[slot] = (int)c;
      else atomicOr(&a.cnt->error, 4);
    } else {
      for (int p = 0; p < 5; p++) a.V(p, c) = vs[p] / sw;
    }
  }
}

// FixUnresolvedNodes (:1278-1525) with apply_failsafe_density = On; one
// lane, nodes in (k,j,i) order, with the reference's `reset` flag shared by
// all nodes.  With PhaseChangeDirection = All the upwind accumulation still
// writes the node and sets `reset`, but the inverse-distance average is
// taken (:1396).
__global__ void k_fix_unresolved(UpdArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nu = min(a.cnt->n_unresolved, a.cap);
  if (nu == 0) return;
  a.cnt->unresolved_total += (unsigned long long)nu;
  const DevGrid& g = a.g;
  int* U = a.unresolved;
  for (int x = 1; x < nu; x++) {  // insertion sort by linear index = (k,j,i) order
    int v = U[x], y = x - 1;
    while (y >= 0 && U[y] > v) { U[y + 1] = U[y]; y--; }
    U[y + 1] = v;
  }
  auto unres = [&](long c) {
    int lo = 0, hi = nu - 1;
    while (lo <= hi) { int m = (lo + hi) / 2; if (U[m] == c) return true; if (U[m] < c) lo = m + 1; else hi = m - 1; }
    return false;
  };
  bool reset = false;
  for (int u = 0; u < nu; u++) {
    long c = U[u];
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    const int myid = a.ID[c];
    const double* gx = g_axis(g, 0);
    const double* gy = g_axis(g, 1);
    const double* gz = g_axis(g, 2);
    const double x0[3] = {gx[i + 1], gy[j + 1], gz[k + 1]};
    double sum_weight = 0.0, sum_weight2 = 0.0, vtmp[5] = {0, 0, 0, 0, 0};
    double v[5];
    for (int p = 0; p < 5; p++) v[p] = a.V(p, c);
    for (int nk = k - 1; nk <= k + 1; nk++)
      for (int nj = j - 1; nj <= j + 1; nj++)
        for (int ni = i - 1; ni <= i + 1; ni++) {
          if (ni < 0 || ni >= g.NX || nj < 0 || nj >= g.NY || nk < 0 || nk >= g.NZ) continue;
          long nc = gidx(g, ni, nj, nk);
          if (a.ID[nc] != myid) continue;
          if (nc == c) continue;
          if (unres(nc)) continue;
          double vn[5];
          for (int p = 0; p < 5; p++) vn[p] = a.V(p, nc);
          double v1[3] = {vn[1], vn[2], vn[3]};
          double v1n = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
          if (v1n != 0) { double cs = 1.0 / v1n; v1[0] *= cs; v1[1] *= cs; v1[2] *= cs; }
          double xx[3] = {x0[0] - gx[ni + 1], x0[1] - gy[nj + 1], x0[2] - gz[nk + 1]};
          double xn = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
          double cs = 1.0 / xn;
          xx[0] *= cs; xx[1] *= cs; xx[2] *= cs;
          double dot = xx[0] * v1[0] + xx[1] * v1[1] + xx[2] * v1[2];
          double weight = (0.0 < dot) ? dot : 0.0;  // std::max(0.0, dot)
          if (weight > 0) {
            sum_weight += weight;
            if (reset) for (int p = 0; p < 5; p++) v[p] += weight * vn[p];
            else { for (int p = 0; p < 5; p++) v[p] = weight * vn[p]; reset = true; }
          }
          for (int p = 0; p < 5; p++) vtmp[p] += 1.0 / xn * vn[p];
          sum_weight2 += 1.0 / xn;
        }
    bool fixed = false;
    if (a.s.phase_dir == M2CX_PHASEDIR_UPWIND && sum_weight > 0) {
      double cs = 1.0 / sum_weight;
      for (int p = 0; p < 5; p++) v[p] *= cs;
      fixed = true;
    } else if (sum_weight2 > 0) {
      for (int p = 0; p < 5; p++) v[p] = vtmp[p] / sum_weight2;
      fixed = true;
    }
    if (!fixed) {
      double density = 0.0;
      for (int layer = 1; layer <= 10; layer++) {
        for (int nk = k - layer; nk <= k + layer; nk++)
          for (int nj = j - layer; nj <= j + layer; nj++)
            for (int ni = i - layer; ni <= i + layer; ni++) {
              if (ni < 0 || ni >= g.NX || nj < 0 || nj >= g.NY || nk < 0 || nk >= g.NZ) continue;
              long nc = gidx(g, ni, nj, nk);
              if (a.ID[nc] != myid) continue;
              if (nc == c) continue;
              if (unres(nc)) continue;
              double dx0 = gx[ni + 1] - x0[0], dx1 = gy[nj + 1] - x0[1], dx2 = gz[nk + 1] - x0[2];
              double dist = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
              sum_weight += 1.0 / dist;
              density += (1.0 / dist) * a.V(0, nc);
            }
        if (sum_weight > 0) { v[0] = density / sum_weight; break; }
      }
      if (sum_weight == 0) v[0] = c_mats[a.s.mset].m[myid].failsafe;
    }
    for (int p = 0; p < 5; p++) a.V(p, c) = v[p];
  }
}

// ClipDensityAndPressure + ApplyBoundaryConditions after the update
// (TimeIntegrator.cpp:802-803), on the cells that changed (all others are
// unchanged since their last clip), and clearing of the done flags.
__global__ void k_post_update(UpdArgs a) {
  const int nch = min(a.cnt->n_changed, a.cap);
  const DevGrid& g = a.g;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nch; q += gridDim.x * blockDim.x) {
    long c = a.changed[q];
    a.done[c] = 0;
    double v[5];
    for (int p = 0; p < 5; p++) v[p] = a.V(p, c);
    const DevEos& E = c_mats[a.s.mset].m[a.ID[c]];
    if (eos_clip(E, v)) {
      atomicAdd(&a.cnt->clipped, 1ull);
      for (int p = 0; p < 5; p++) a.V(p, c) = v[p];
    }
    if (eos_check5(E, v)) atomicOr(&a.cnt->error, 1);
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    write_ghosts(g, a.V, i, j, k, v);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.cnt->id_changed, (unsigned long long)nch);
}

// ---------------------------------------------------------------------------
// Initial condition from geometric entities (SpaceInitializer.cpp:237-482 for
// the state and material ID, :1390-1514 for the level set): the default state
// everywhere, then every geometry in order overrides the interior cells it
// covers; phi starts at the domain diagonal and takes the signed distance of
// the nearest surface of the level-set material's geometries.
// ---------------------------------------------------------------------------
struct InitGeom {
  int kind, interior, id, pad;  // kind 0 plane, 3 sphere
  double x0[3], n[3], R;        // plane: point and unit normal; sphere: centre and radius
  double V[5];
};

// state: write V and ID too (0: only phi, for level sets after the first)
__global__ void k_initialize(DevGrid g, F5 V, unsigned char* ID, double* phi, const double* __restrict__ V0,
                             int id0, const InitGeom* __restrict__ geo, int ngeo, int lsid, double diag,
                             int state = 1) {
  const long n = (long)g.NX * g.NY * g.NZ;
  const double* gx = g_axis(g, 0);
  const double* gy = g_axis(g, 1);
  const double* gz = g_axis(g, 2);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    const double x = gx[i + 1], y = gy[j + 1], z = gz[k + 1];
    double v[5] = {V0[0], V0[1], V0[2], V0[3], V0[4]};
    int id = id0;
    double ph = diag;
    for (int m = 0; m < ngeo; m++) {
      const InitGeom& G = geo[m];
      double dist;
      bool inside;
      if (G.kind == 0) {
        dist = (x - G.x0[0]) * G.n[0] + (y - G.x0[1]) * G.n[1] + (z - G.x0[2]) * G.n[2];
        inside = dist > 0;
      } else {
        const double a = x - G.x0[0], b = y - G.x0[1], e = z - G.x0[2];
        dist = sqrt(a * a + b * b + e * e) - G.R;
        inside = G.interior ? dist < 0 : dist > 0;
      }
      if (inside) {
        for (int p = 0; p < 5; p++) v[p] = G.V[p];
        id = G.id;
      }
      if (phi && G.id == lsid && fabs(dist) < fabs(ph)) {
        if (G.kind == 0) ph = -dist;
        else ph = G.interior ? dist : -dist;
      }
    }
    if (state) {
      store5(V, c, v);
      ID[c] = (unsigned char)id;
    }
    if (phi) phi[c] = ph;
  }
}

// ---------------------------------------------------------------------------
// host-side orchestration
// ---------------------------------------------------------------------------
static inline int blocks_for(long n, int bs = 256) { return (int)std::min<long>((n + bs - 1) / bs, 256L * 64L); }

inline void mp_classify(MPBuffers& B, const DevGrid& g, const Scheme& s, const unsigned char* ID, DevCounters* cnt,
                        hipStream_t st) {
  k_if_reset<<<1, 1, 0, st>>>(cnt);
  k_classify<<<blocks_for((long)g.NX * g.NY * g.NZ), 256, 0, st>>>(g, s, ID, B.fidx[0], B.fidx[1], B.fidx[2], B.L,
                                                                     cnt);
}

inline int mp_slab(MPBuffers& B) {
  if (B.slab) return 0;
  return hipMalloc(&B.slab, (long)B.lanes * 2 * B.pathcap * 3 * 8) == hipSuccess ? 0 : M2CX_E_OOM;
}

inline void mp_riemann(MPBuffers& B, const DevGrid& g, const Scheme& s, CF5 V, const unsigned char* ID, int stage1,
                       DevCounters* cnt, hipStream_t st) {
  if (s.mp_flux == M2CX_MPFLUX_LLF) return;
  auto* kern = (s.flux == M2CX_FLUX_GODUNOV && s.mp_flux == M2CX_MPFLUX_NUMERICAL) ? k_riemann<true> : k_riemann<false>;
  kern<<<B.lanes / 64, 64, 0, st>>>(g, s, B.rc, V, ID, B.L, stage1, B.slab, B.lanes, B.pathcap,
                                         B.guess ? B.pguess : nullptr, cnt);
}

// Tiles of the stage kernel (TX x TY x TZ cells, numbered as k_stage numbers
// them) that compute a flux through an interface face: the tiles of both
// cells of every listed face.  The other tiles need no Riemann result and can
// run while the interface problems are solved.
__global__ void k_tile_flags(DevGrid g, IfaceList L, const DevCounters* cnt, unsigned char* tflag, int TX, int TY,
                             int TZ) {
  const int nif = min(cnt->nif, L.cap);
  const int ntx = (g.NX + TX - 1) / TX, nty = (g.NY + TY - 1) / TY;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nif; q += gridDim.x * blockDim.x) {
    long r = L.cell[q];
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    const int d = L.dir[q];
    const int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
    tflag[(i / TX) + ntx * ((j / TY) + nty * (k / TZ))] = 1;
    tflag[(im / TX) + ntx * ((jm / TY) + nty * (km / TZ))] = 1;
  }
}

__global__ void k_clear_u8(unsigned char* p, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) p[c] = 0;
}

// plain copies as kernels (no memcpy nodes when a step is captured in a graph)
__global__ void k_copy_u8(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) dst[c] = src[c];
}
__global__ void k_copy_f64(double* __restrict__ dst, const double* __restrict__ src, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) dst[c] = src[c];
}

// copy_phi: Phi_tmp = Phi over the ghosted brick (the narrow band copies
// its useful nodes itself after the first step)
inline void mp_begin_step(MPBuffers& B, const DevGrid& g, const Scheme& s, const unsigned char* ID, DevCounters* cnt,
                          hipStream_t st, bool copy_phi = true) {
  mp_classify(B, g, s, ID, cnt, st);
  k_copy_u8<<<blocks_for(g.ncells), 256, 0, st>>>(B.IDn, ID, g.ncells);
  if (B.nls && copy_phi) k_copy_f64<<<blocks_for(g.ncells), 256, 0, st>>>(B.phin, B.phi, g.ncells);
}

// the level-set kernel's arguments; the stencil field of stage `which` is
// phi (stages -1, 0, 1) or phi1 (2, 3, 4)
inline LSArgs mp_ls_args(const MPBuffers& B, const DevGrid& g, CF5 V, int which, DevCounters* cnt) {
  LSArgs a{};
  a.g = g;
  a.fv = B.fv;
  a.V = V;
  a.fd3 = B.fd3;
  a.ocxL = B.off_cxL; a.ocxR = B.off_cxR; a.ocyL = B.off_cyL; a.ocyR3 = B.off_cyR3; a.ocyR1 = B.off_cyR1;
  a.oczL = B.off_czL; a.oczR = B.off_czR;
  a.which = which;
  a.cnt = cnt;
  a.rls = B.rls;
  a.phi_st = (which <= 1) ? B.phi : B.phi1;
  return a;
}

// which: -1 residual only, 0 FE, 1 RK stage 1, 2 RK2 final, 3 RK3 stage 2, 4 RK3 final
inline void mp_levelset(MPBuffers& B, const DevGrid& g, CF5 V, int which, DevCounters* cnt, hipStream_t st) {
  if (!B.nls) return;
  LSArgs a = mp_ls_args(B, g, V, which, cnt);
  double* written = nullptr;
  switch (which) {
    case -1: a.phi_st = B.phi; break;
    case 0: a.phi_st = B.phi; a.phi_a = B.phi; a.out = B.phitmp; written = B.phitmp; break;
    case 1: a.phi_st = B.phi; a.phi_a = B.phi; a.out = B.phi1; written = B.phi1; break;
    case 2: a.phi_st = B.phi1; a.phi_a = B.phi; a.phi_b = B.phi1; a.out = B.phi; written = B.phi; break;
    case 3: a.phi_st = B.phi1; a.phi_a = B.phi; a.phi_b = B.phi1; a.out = B.phitmp; written = B.phitmp; break;
    case 4: a.phi_st = B.phi1; a.phi_a = B.phi; a.phi_b = B.phi1; a.out = B.phi; written = B.phi; break;
  }
  k_levelset<<<blocks_for((long)g.NX * g.NY * g.NZ), 256, 0, st>>>(a);
  if (which == 0) {
    std::swap(B.phi, B.phitmp);
  }
  if (which == 3) std::swap(B.phi1, B.phitmp);
  if (written) {
    long nf = 2L * ((long)g.NY * g.NZ + (long)g.NX * g.NZ + (long)g.NX * g.NY);
    double* tgt = (which == 0) ? B.phi : (which == 3 ? B.phi1 : written);
    k_ls_bc<<<blocks_for(nf), 256, 0, st>>>(g, tgt, B.lsbc[0], B.lsbc[1], B.lsbc[2], B.lsbc[3], B.lsbc[4],
                                           B.lsbc[5]);
  }
}

// UpdateMaterialIDByLevelSet(phin, phi) with the ghost IDs
inline void mp_end_ids(MPBuffers& B, const DevGrid& g, const Scheme& s, const LSPtrs& P, unsigned char* ID,
                       DevCounters* cnt, hipStream_t st) {
  const long nin = (long)g.NX * g.NY * g.NZ;
  k_end_reset<<<1, 1, 0, st>>>(cnt);
  k_id_sweep<<<blocks_for(nin), 256, 0, st>>>(g, s, P, ID, B.tag, B.remaining, B.list_cap, cnt);
  k_id_vote<<<1, 1024, 0, st>>>(g, s, ID, B.tag, B.remaining, B.list_cap, P.tracked, cnt);
  long nf = 2L * ((long)g.NY * g.NZ + (long)g.NX * g.NZ + (long)g.NX * g.NY);
  k_ghost_ids<<<blocks_for(nf), 256, 0, st>>>(g, ID);
}

inline UpdArgs mp_upd_args(MPBuffers& B, const DevGrid& g, const Scheme& s, F5 V, unsigned char* ID,
                           DevCounters* cnt) {
  UpdArgs a{};
  a.g = g;
  a.s = s;
  a.V = V;
  a.ID = ID;
  a.IDn = B.IDn;
  for (int d = 0; d < 3; d++) a.fidx[d] = B.fidx[d];
  a.L = B.L;
  a.changed = B.changed;
  a.unresolved = B.unresolved;
  a.done = B.done;
  a.cap = B.list_cap;
  a.cnt = cnt;
  return a;
}

// UpdateStateVariablesAfterInterfaceMotion up to the list of unresolved cells
inline void mp_end_states(MPBuffers& B, const DevGrid& g, const Scheme& s, F5 V, unsigned char* ID,
                          DevCounters* cnt, hipStream_t st) {
  const long nin = (long)g.NX * g.NY * g.NZ;
  k_changed<<<blocks_for(nin), 256, 0, st>>>(g, s, ID, B.IDn, B.changed, B.list_cap, cnt);
  if (s.phase_change == M2CX_PHASECHANGE_EXTRAPOLATION)
    k_extrapolate<<<256, 256, 0, st>>>(mp_upd_args(B, g, s, V, ID, cnt));
  else
    k_state_update<<<1, 1024, 0, st>>>(mp_upd_args(B, g, s, V, ID, cnt));
}

// FixUnresolvedNodes, then clipping and the boundary fill of V
inline void mp_end_fix(MPBuffers& B, const DevGrid& g, const Scheme& s, F5 V, unsigned char* ID, DevCounters* cnt,
                       hipStream_t st) {
  UpdArgs a = mp_upd_args(B, g, s, V, ID, cnt);
  k_fix_unresolved<<<1, 64, 0, st>>>(a);
  k_post_update<<<256, 256, 0, st>>>(a);
}

// the end-of-step update without a reinitializer
// the level set of B alone (one level set per solver)
inline LSPtrs mp_ls_single(const MPBuffers& B) {
  LSPtrs P{};
  P.phi[0] = B.phi;
  P.phin[0] = B.phin;
  P.mat[0] = B.lsid;
  P.n = 1;
  P.tracked = 1u << B.lsid;
  return P;
}

inline void mp_end_step(MPBuffers& B, const DevGrid& g, const Scheme& s, F5 V, unsigned char* ID, DevCounters* cnt,
                        hipStream_t st) {
  mp_end_ids(B, g, s, mp_ls_single(B), ID, cnt, st);
  mp_end_states(B, g, s, V, ID, cnt, st);
  mp_end_fix(B, g, s, V, ID, cnt, st);
}

}  // namespace m2c
