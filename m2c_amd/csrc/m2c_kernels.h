// Kernels of the explicit step (gfx950).  See m2c_grid.h for the layout.
#pragma once
#include "m2c_grid.h"

namespace m2c {

// Device-side counters and flags of one solver.
struct DevCounters {
  unsigned long long clipped;      // ClipDensityAndPressure hits (all stages)
  unsigned long long stage_clipped;  // hits in the current intermediate stage
  unsigned long long riemann_calls, riemann_trivial, riemann_errors, riemann_iters, riemann_fevals;
  unsigned long long riemann_fallbacks, guess_fallbacks, id_changed, steps;
  int error;          // bit 0 hyperbolicity (abort), bit 1 Riemann fatal, bit 2 capacity, bit 3 unresolved,
                      // bit 4 a node inside two level-set subdomains
  int done;           // time loop reached MaxTime
  unsigned int dt_ticket;
  int nif;            // interface faces in the current list
  int n_remaining, n_changed, n_unresolved;
  int n_resolved;     // nodes given to one subdomain by ResolveConflictsInLevelSets this step
  int n_lsr;          // nodes inside two or more level-set subdomains (its candidates)
  double t, dt, cfl, dts, tmax;
  int time_step, pad2;
  unsigned long long unresolved_total;  // cells left unresolved by the state update, summed
};

enum StageMode { MODE_RESIDUAL = 0, MODE_STAGE1 = 1, MODE_RK2_FINAL = 2, MODE_RK3_STAGE2 = 3, MODE_RK3_FINAL = 4,
                 MODE_FE = 5 };

// Interface-face list produced once per time step (face between cell
// `cell - e_dir` and `cell`).  Fluxes are refreshed every stage; the star
// states are the stage-1 capture (RiemannSolutions, RiemannSolutions.h:12-30).
struct IfaceList {
  int* cell;      // linear index of the plus-side cell
  int* dir;
  double* f1;     // [5][cap]: flux applied to the minus-side cell (localflux1)
  double* f2;     // [5][cap]: flux applied to the plus-side cell (localflux2)
  double* vsm;    // [5][cap]: captured star state, minus side
  double* vsp;    // [5][cap]
  int* err;       // Riemann error of the capture
  int* stats;     // [3][cap]: main-loop iterations, f(p) evaluations, flags (last solve)
  unsigned long long* hist;  // [2][M2CX_RIEMANN_HIST_BINS]: counts of iterations, of f(p) evaluations
  // the level sets at time n (the Riemann normal of LevelSet / Average) and
  // the material each tracks
  const double* lsphi[M2CX_MAX_LEVELSETS];
  int lsmat[M2CX_MAX_LEVELSETS];
  int nls;
  int cap;
};

struct StageArgs {
  DevGrid g;
  Scheme s;
  CF5 Vin;
  F5 Vout, Un, U1, R;
  const unsigned char* ID;
  const int* fidx[3];
  IfaceList L;
  int mode;
  // 0: every tile; 1: only tiles without interface faces; 2: only tiles with
  // them (tflag per tile, set by k_tile_flags for this tile shape)
  int phase;
  const unsigned char* tflag;
  DevCounters* cnt;
  // Flux = Godunov: the fluxes of the faces with one material on both sides,
  // [axis][component][plus cell], computed by k_godunov before the stage
  const double* gflux;
};

__device__ __forceinline__ void write_ghosts(const DevGrid& g, const F5& V, int i, int j, int k, const double* v) {
  int side = -1;
  for (int pass = 0; pass < 6; pass++) {
    int gi = i, gj = j, gk = k;
    if (pass == 0) { if (i != 0) continue; gi = -1; }
    else if (pass == 1) { if (i != g.NX - 1) continue; gi = g.NX; }
    else if (pass == 2) { if (j != 0) continue; gj = -1; }
    else if (pass == 3) { if (j != g.NY - 1) continue; gj = g.NY; }
    else if (pass == 4) { if (k != 0) continue; gk = -1; }
    else { if (k != g.NZ - 1) continue; gk = g.NZ; }
    side = pass;
    const int bc = g.bc[side];
    double w[5];
    if (bc == M2CX_BC_INLET) { for (int p = 0; p < 5; p++) w[p] = g.inlet[p]; }
    else if (bc == M2CX_BC_INLET2) { for (int p = 0; p < 5; p++) w[p] = g.inlet2[p]; }
    else if (bc == M2CX_BC_SYMMETRY || bc == M2CX_BC_SLIPWALL) {
      for (int p = 0; p < 5; p++) w[p] = v[p];
      w[side / 2 + 1] *= -1.0;
    } else if (bc == M2CX_BC_STICKWALL) {
      w[0] = v[0]; w[1] = -1.0 * v[1]; w[2] = -1.0 * v[2]; w[3] = -1.0 * v[3]; w[4] = v[4];
    } else if (bc == M2CX_BC_OUTLET) {
      for (int p = 0; p < 5; p++) w[p] = v[p];
    } else {
      continue;
    }
    store5(V, gidx(g, gi, gj, gk), w);
  }
}

// Flux through the minus face (axis D) of cell (i,j,k); returns f1 (applied to
// the minus-side cell) and the reconstruction-check bits of the two cells.
template <bool GEN>
__device__ __forceinline__ void face_flux(const StageArgs& a, const DevMats& M, int D, int i, int j, int k,
                                          double* f1, int& bm, int& bp, int& err) {
  const DevGrid& g = a.g;
  const Scheme& s = a.s;
  const long c = gidx(g, i, j, k);
  const long st = axis_stride(g, D);
  const int myid = a.ID[c], nid = a.ID[c - st];
  double Vm[5], Vp[5];
  face_pair<GEN>(g, s, M, a.Vin, a.ID, D, i, j, k, Vm, Vp, bm, bp);
  if (nid == myid) {
    if (myid == s.inactive) {
      for (int p = 0; p < 5; p++) f1[p] = 0.0;
    } else if (s.flux == M2CX_FLUX_GODUNOV) {
      for (int p = 0; p < 5; p++) f1[p] = a.gflux[((long)D * 5 + p) * g.ncells + c];
    } else {
      flux_num(M.m[myid], s.flux, s.del, D, Vm, Vp, f1, err);
    }
  } else if (nid == s.inactive || myid == s.inactive) {
    for (int p = 0; p < 5; p++) f1[p] = 0.0;
  } else if (s.mp_flux == M2CX_MPFLUX_LLF) {
    double vm[5], vp[5];
    if (s.mp_recon == M2CX_MPREC_CONSTANT) { load5(a.Vin, c - st, vm); load5(a.Vin, c, vp); }
    else for (int p = 0; p < 5; p++) { vm[p] = Vm[p]; vp[p] = Vp[p]; }
    flux_llf(M.m[nid], M.m[myid], D, vm, vp, f1, err);
  } else {
    const int* fx = (D == 0) ? a.fidx[0] : (D == 1 ? a.fidx[1] : a.fidx[2]);
    const int slot = fx[c];
    for (int p = 0; p < 5; p++) f1[p] = a.L.f1[(long)p * a.L.cap + slot];
  }
}

constexpr int NXCD = 8;  // XCDs of an MI355X

// LDS slot of the face flux of axis D whose plus cell has tile-local
// position (li, lj, lk) in a BX x BY x BZ tile; x faces: (BX+1) x BY x BZ,
// y: BX x (BY+1) x BZ, z: BX x BY x (BZ+1), stored one after another.
template <int BX, int BY, int BZ>
struct TileShape {
  static constexpr int NT = BX * BY * BZ;
  static constexpr int NFX = (BX + 1) * BY * BZ, NFY = BX * (BY + 1) * BZ, NFZ = BX * BY * (BZ + 1);
  static constexpr int NFACE = NFX + NFY + NFZ;
  // plus faces of the last layer, one per thread: y, then z, then x
  static constexpr int EY = BX * BZ, EZ = BX * BY, EX = BY * BZ;
  static_assert(EY + EZ + EX <= NT, "one extra face per thread at most");
  __device__ static __forceinline__ int slot(int D, int li, int lj, int lk) {
    if (D == 0) return (lk * BY + lj) * (BX + 1) + li;
    if (D == 1) return NFX + (lk * (BY + 1) + lj) * BX + li;
    return NFX + NFY + (lk * BY + lj) * BX + li;
  }
};

// Material tables of the live solvers, one slot per distinct table (solvers
// with equal tables share a slot; Scheme::mset selects it).
constexpr int MAX_MAT_SETS = 32;
__constant__ DevMats c_mats[MAX_MAT_SETS];

// One tile of the fused stage (the body of k_stage and k_stage_multi).
// GEN: reconstruction in any variable type (false: primitive variables only,
// the common case, compiled without the characteristic code)
template <int MODE, int BX, int BY, int BZ, bool GEN>
__device__ __forceinline__ void stage_tile(const StageArgs& a, int tile) {
  using TS = TileShape<BX, BY, BZ>;
  constexpr int TX = BX, TY = BY, TZ = BZ;
  __shared__ double sF[5][TS::NFACE];
  __shared__ int sMask[TS::NT];
  const DevGrid& g = a.g;
  const Scheme& s = a.s;
  const DevMats& M = c_mats[a.s.mset];
  const int tx = threadIdx.x % TX, ty = (threadIdx.x / TX) % TY, tz = threadIdx.x / (TX * TY);
  const int t = threadIdx.x;
  const int ntx = (g.NX + TX - 1) / TX, nty = (g.NY + TY - 1) / TY;
  if (a.phase != 0 && ((a.tflag[tile] != 0) != (a.phase == 2))) return;
  const int ib = (tile % ntx) * TX, jb = ((tile / ntx) % nty) * TY, kb = (tile / (ntx * nty)) * TZ;
  const int i = ib + tx, j = jb + ty, k = kb + tz;
  const bool inside = (i < g.NX) && (j < g.NY) && (k < g.NZ);
  sMask[t] = 0;
  __syncthreads();
  int err = 0;
  // Face fluxes of the tile, one loop body for all of them (a single copy of
  // the reconstruction + flux code keeps the instruction footprint small):
  // jobs 0..2 are the minus faces x, y, z of this thread's position (also at
  // the first position past the domain end, whose minus face closes the last
  // cell); job 3 (threads < 144) is one plus face of the tile's last layer:
  // y faces for t < 64, z faces for t < 128, x faces for t < 144.
  const int njobs = (t < TS::EY + TS::EZ + TS::EX) ? 4 : 3;
#pragma unroll 1
  for (int job = 0; job < njobs; job++) {
    int D, li, lj, lk;  // axis and tile-local position of the face's plus cell
    if (job < 3) {
      D = job; li = tx; lj = ty; lk = tz;
    } else if (t < TS::EY) {
      D = 1; li = t % TX; lj = TY; lk = t / TX;
    } else if (t < TS::EY + TS::EZ) {
      const int e = t - TS::EY;
      D = 2; li = e % TX; lj = e / TX; lk = TZ;
    } else {
      const int e = t - TS::EY - TS::EZ;
      D = 0; li = TX; lj = e % TY; lk = e / TY;
    }
    const int fi = ib + li, fj = jb + lj, fk = kb + lk;
    const bool ok = (D == 0) ? (fi <= g.NX && fj < g.NY && fk < g.NZ)
                  : (D == 1) ? (fi < g.NX && fj <= g.NY && fk < g.NZ)
                             : (fi < g.NX && fj < g.NY && fk <= g.NZ);
    if (!ok) continue;
    double f[5];
    int bm, bp;
    face_flux<GEN>(a, M, D, fi, fj, fk, f, bm, bp, err);
    const int slot = TS::slot(D, li, lj, lk);
#pragma unroll
    for (int p = 0; p < 5; p++) sF[p][slot] = f[p];
    // check bits: the plus cell's minus face (slot 2D), the minus cell's plus
    // face (slot 2D+1), for the cells inside the tile
    const int pl = li + TX * (lj + TY * lk);
    if (bp && li < TX && lj < TY && lk < TZ) atomicOr(&sMask[pl], bp << (4 * D));
    const int mi = li - (D == 0), mj = lj - (D == 1), mk = lk - (D == 2);
    if (bm && mi >= 0 && mj >= 0 && mk >= 0 && mi < TX && mj < TY && mk < TZ)
      atomicOr(&sMask[mi + TX * (mj + TY * mk)], bm << (4 * D + 2));
  }
  __syncthreads();
  if (err) atomicOr(&a.cnt->error, 1);
  if (!inside) return;

  // reconstruction check of this cell's six faces (abort condition)
  {
    const int m = sMask[t];
    int fails = 0, unf = 0;
    for (int q = 0; q < 6; q++) {
      if (m & (1 << (2 * q))) fails |= 1 << q;
      if (m & (2 << (2 * q))) unf |= 1 << q;
    }
    if (fails) {
      const int rest = fails & (fails - 1);
      if (rest & unf) atomicOr(&a.cnt->error, 1);
    }
  }

  const long c = gidx(g, i, j, k);
  const int myid = a.ID[c];
  const double dxi = g_axis(g, 3)[i + 1], dyj = g_axis(g, 4)[j + 1], dzk = g_axis(g, 5)[k + 1];
  const double ax = dyj * dzk;
  const double ay = dxi * dzk;
  const double az = dxi * dyj;
  const int fxm = TS::slot(0, tx, ty, tz), fxp = TS::slot(0, tx + 1, ty, tz);
  const int fym = TS::slot(1, tx, ty, tz), fyp = TS::slot(1, tx, ty + 1, tz);
  const int fzm = TS::slot(2, tx, ty, tz), fzp = TS::slot(2, tx, ty, tz + 1);
  // minus faces of an interface take localflux2 from the Riemann list
  const bool mp = s.mp_flux != M2CX_MPFLUX_LLF;
  const int idx_m = a.ID[c - 1], idy_m = a.ID[c - g.SX], idz_m = a.ID[c - g.S];
  const bool ifx = mp && idx_m != myid && idx_m != s.inactive && myid != s.inactive;
  const bool ify = mp && idy_m != myid && idy_m != s.inactive && myid != s.inactive;
  const bool ifz = mp && idz_m != myid && idz_m != s.inactive && myid != s.inactive;
  const int slx = ifx ? a.fidx[0][c] : 0, sly = ify ? a.fidx[1][c] : 0, slz = ifz ? a.fidx[2][c] : 0;
  double r[5];
  const double cst = 1.0 / (-(dxi * dyj * dzk));
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const double Fm = ifx ? a.L.f2[(long)p * a.L.cap + slx] : sF[p][fxm];
    const double Gm = ify ? a.L.f2[(long)p * a.L.cap + sly] : sF[p][fym];
    const double Hm = ifz ? a.L.f2[(long)p * a.L.cap + slz] : sF[p][fzm];
    double acc = 0.0 - ax * Fm;
    acc = acc - ay * Gm;
    acc = acc - az * Hm;
    acc = acc + ax * sF[p][fxp];
    acc = acc + ay * sF[p][fyp];
    acc = acc + az * sF[p][fzp];
    r[p] = acc * cst;
  }

  if (MODE == MODE_RESIDUAL) {
    store5(a.R, c, r);
    return;
  }
  const DevEos& E = M.m[myid];
  const double dt = a.cnt->dt;
  double u[5], v[5];
  if (MODE == MODE_STAGE1 || MODE == MODE_FE) {
    double v0[5], un[5];
    load5(a.Vin, c, v0);
    eos_p2c(E, v0, un);
    if (MODE == MODE_STAGE1) store5(a.Un, c, un);
#pragma unroll
    for (int p = 0; p < 5; p++) u[p] = un[p] + dt * r[p];
  } else {
    double u1[5], un[5];
#pragma unroll
    for (int p = 0; p < 5; p++) { u1[p] = a.U1(p, c); un[p] = a.Un(p, c); }
    if (MODE == MODE_RK2_FINAL) {
      const double h = 0.5 * dt;
#pragma unroll
      for (int p = 0; p < 5; p++) { u[p] = 0.5 * u1[p] + 0.5 * un[p]; u[p] = u[p] + h * r[p]; }
    } else if (MODE == MODE_RK3_STAGE2) {
      const double h = 0.25 * dt;
#pragma unroll
      for (int p = 0; p < 5; p++) { u[p] = 0.25 * u1[p] + 0.75 * un[p]; u[p] = u[p] + h * r[p]; }
    } else {
      const double h = 2.0 / 3.0 * dt;
#pragma unroll
      for (int p = 0; p < 5; p++) { u[p] = 2.0 / 3.0 * u1[p] + 1.0 / 3.0 * un[p]; u[p] = u[p] + h * r[p]; }
    }
  }
  eos_c2p(E, u, v);
  const bool clipped = eos_clip(E, v);
  if (clipped) {
    atomicAdd(&a.cnt->clipped, 1ull);
    if (MODE == MODE_STAGE1 || MODE == MODE_RK3_STAGE2) atomicAdd(&a.cnt->stage_clipped, 1ull);
  }
  if (eos_check5(E, v)) atomicOr(&a.cnt->error, 1);
  if (MODE == MODE_STAGE1 || MODE == MODE_RK3_STAGE2) store5(a.U1, c, u);
  store5(a.Vout, c, v);
  write_ghosts(g, a.Vout, i, j, k, v);
}

// XCD-aware tile order: work-groups are dealt round-robin to the 8 XCDs,
// each with its own L2; work-group w runs tile (w % 8) * per + w / 8, so
// every XCD sweeps one contiguous run of tiles (x fastest) and the halos its
// neighbouring tiles share stay in its L2.
__device__ __forceinline__ int xcd_tile(int total) {
  const int per = (total + NXCD - 1) / NXCD;
  return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
}

template <int MODE, int BX, int BY, int BZ, bool GEN>
__global__ void __launch_bounds__(BX * BY * BZ) k_stage(StageArgs a) {
  // `a` is read in place, at the start of the kernel-argument segment (the
  // first explicit argument sits at offset 0).  Handing the parameter itself
  // to the tile body by reference makes the compiler build a private copy of
  // it, whose run-time-selected fields (per-axis sizes, boundary conditions,
  // face-index arrays) are then loaded from scratch memory: 213 VGPRs and
  // 488 B of scratch per lane against 166 VGPRs and none this way.
  const StageArgs& ar = *(const StageArgs*)__builtin_amdgcn_kernarg_segment_ptr();
  const DevGrid& g = ar.g;
  const int ntiles = ((g.NX + BX - 1) / BX) * ((g.NY + BY - 1) / BY) * ((g.NZ + BZ - 1) / BZ);
  const int tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;
  stage_tile<MODE, BX, BY, BZ, GEN>(ar, tile);
}

// Ensemble form: the tiles of many solvers (one table entry each, the same
// stage mode) in one launch.  toff holds the prefix sums of the members'
// tile counts; the concatenated tile list is dealt to the XCDs as above, so
// each XCD sweeps a contiguous run of it.  Every tile runs the code it runs
// in k_stage on the same inputs, so each member's result is the same.
template <int MODE, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX * BY * BZ) k_stage_multi(const StageArgs* __restrict__ tab,
                                                               const int* __restrict__ toff, int nm) {
  const int total = toff[nm];
  const int t = xcd_tile(total);
  if (t >= total) return;
  int lo = 0, hi = nm - 1;  // member m with toff[m] <= t < toff[m+1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  stage_tile<MODE, BX, BY, BZ, false>(tab[lo], t - toff[lo]);
}

// Cells whose reconstruction in the scheme's variable type fails CheckState
// on one of their six faces; they are reconstructed in primitive variables
// instead (Reconstructor.cpp:722-778).  c0^2 < 0 in a characteristic
// transform is the reference's abort (FluxFcnBase.h:315-320).
__global__ void k_recon_flags(DevGrid g, Scheme s, CF5 V, const unsigned char* __restrict__ ID,
                              unsigned char* __restrict__ flag, DevCounters* cnt) {
  const DevMats& M = c_mats[s.mset];
  const long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    const int myid = ID[c];
    if (myid == s.inactive) {
      flag[c] = 0;
      continue;
    }
    const DevEos& E = M.m[myid];
    double v[5];
    load5(V, c, v);
    bool fails = false, ok = true;
#pragma unroll 1
    for (int D = 0; D < 3; D++) {
      double q0[5], sig[5], f[5];
      ok = recon_slope(g, s, M, V, ID, s.var_type, D, i, j, k, v, q0, sig) && ok;
      recon_face(E, s.var_type, v, q0, sig, 0, f);
      fails = fails || eos_check5(E, f);
      recon_face(E, s.var_type, v, q0, sig, 1, f);
      fails = fails || eos_check5(E, f);
    }
    flag[c] = fails ? 1 : 0;
    if (!ok) atomicOr(&cnt->error, 1);
  }
}

// U = P2C(V) on interior cells, only if the preceding stage clipped a cell
// (TimeIntegrator.cpp:569-572, 622-625).
__global__ void k_resync(DevGrid g, int mset, CF5 V, F5 U, const unsigned char* __restrict__ ID, const DevCounters* cnt) {
  if (cnt->stage_clipped == 0) return;
  long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    double v[5], u[5];
    load5(V, c, v);
    eos_p2c(c_mats[mset].m[ID[c]], v, u);
    store5(U, c, u);
  }
}

__global__ void k_reset_stage(DevCounters* cnt) { cnt->stage_clipped = 0; }
__global__ void k_reset_error(DevCounters* cnt) { cnt->error = 0; }

// SpaceOperator::ApplyBoundaryConditions on a state field (face ghosts)
__global__ void k_apply_bc(DevGrid g, F5 V) {
  long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    if (i != 0 && i != g.NX - 1 && j != 0 && j != g.NY - 1 && k != 0 && k != g.NZ - 1) continue;
    double v[5];
    CF5 cv{V.b, V.cs};
    load5(cv, gidx(g, i, j, k), v);
    write_ghosts(g, V, i, j, k, v);
  }
}

// ClipDensityAndPressure on interior cells (SpaceOperator.cpp:646-693)
__global__ void k_clip(DevGrid g, int mset, F5 V, const unsigned char* __restrict__ ID, DevCounters* cnt) {
  long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    double v[5];
    for (int p = 0; p < 5; p++) v[p] = V(p, c);
    const DevEos& E = c_mats[mset].m[ID[c]];
    if (eos_clip(E, v)) {
      atomicAdd(&cnt->clipped, 1ull);
      atomicAdd(&cnt->stage_clipped, 1ull);
      for (int p = 0; p < 5; p++) V(p, c) = v[p];
    }
    if (eos_check5(E, v)) atomicOr(&cnt->error, 1);
  }
}

// ---------------------------------------------------------------------------
// Time step (SpaceOperator.cpp:1223-1330 and the loop of Main.cpp:653-712):
// dt = cfl * min over cells and active axes of d_a/(|u_a|+c).  Block minima
// go to `part`; the last block to finish reduces them and, when `advance`,
// clamps dt at MaxTime and moves t, like the driver loop.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double nmin(double a, double b) { return (b < a) ? b : a; }  // std::min(a,b)

__device__ __forceinline__ double wave_min(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = nmin(m, __shfl_xor(m, off, 64));
  return m;
}

// Launched with DT_BLOCKS blocks of 256 (grid-stride over the cells): few
// enough partial minima that the last block reduces them in one pass.
constexpr int DT_BLOCKS = 2048;
__global__ void __launch_bounds__(256) k_dt(DevGrid g, Scheme s, CF5 V, const unsigned char* __restrict__ ID,
                                            double* part, DevCounters* cnt, int advance) {
  __shared__ double red[4];
  __shared__ bool last;
  double m = DBL_MAX;
  const long n = (long)g.NX * g.NY * g.NZ;
  const double* dxs = g_axis(g, 3);
  const double* dys = g_axis(g, 4);
  const double* dzs = g_axis(g, 5);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    long c = gidx(g, i, j, k);
    int id = ID[c];
    if (id == s.inactive) continue;
    const DevEos& E = c_mats[s.mset].m[id];
    double v[5];
    load5(V, c, v);
    double e = eos_e(E, v[0], v[4]);
    double c2 = eos_c2(E, v[0], e);
    if (c2 < 0) atomicOr(&cnt->error, 1);
    double cc = sqrt(c2);
    double lf = fabs(v[1]) + cc, lg = fabs(v[2]) + cc, lh = fabs(v[3]) + cc;
    if (g.NX > 1) m = nmin(m, dxs[i + 1] / lf);
    if (g.NY > 1) m = nmin(m, dys[j + 1] / lg);
    if (g.NZ > 1) m = nmin(m, dzs[k + 1] / lh);
  }
  // min is exact and order-independent: any reduction order gives the same dt
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x] = nmin(nmin(red[0], red[1]), nmin(red[2], red[3]));
    __threadfence();
    unsigned int tk = atomicAdd(&cnt->dt_ticket, 1u);
    last = (tk == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double mm = DBL_MAX;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) mm = nmin(mm, ((volatile double*)part)[b]);
  mm = wave_min(mm);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mm;
  __syncthreads();
  if (threadIdx.x == 0) red[0] = nmin(nmin(red[0], red[1]), nmin(red[2], red[3]));
  if (threadIdx.x == 0) {
    cnt->dt_ticket = 0;
    double cfl, dt;
    if (s.time_step > 0) {  // Time.TimeStep (SpaceOperator.cpp:1315-1321)
      dt = s.time_step;
      cfl = dt / red[0];
    } else {
      cfl = s.cfl;
      dt = cfl * red[0];
    }
    if (advance) {
      double t = cnt->t, tmax = cnt->tmax;
      if (t + dt > tmax) {
        cfl *= (tmax - t) / dt;
        dt = tmax - t;
      }
      cnt->dts = dt;
      cnt->t = t + dt;
      cnt->time_step += 1;
      cnt->steps += 1;
    }
    cnt->dt = dt;
    cnt->cfl = cfl;
  }
}

// ---------------------------------------------------------------------------
// AoS (reference layout) <-> SoA transposes for upload/download
// ---------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double* __restrict__ src, F5 dst, int dof, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
    for (int p = 0; p < dof; p++) dst(p, c) = src[c * dof + p];
}
__global__ void k_soa_to_aos(CF5 src, double* __restrict__ dst, int dof, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
    for (int p = 0; p < dof; p++) dst[c * dof + p] = src(p, c);
}
__global__ void k_i32_to_u8(const int* __restrict__ src, unsigned char* __restrict__ dst, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
    dst[c] = (unsigned char)src[c];
}
__global__ void k_u8_to_i32(const unsigned char* __restrict__ src, int* __restrict__ dst, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
    dst[c] = src[c];
}

// Batched numerical flux (FluxFcnBase::ComputeNumericalFluxAtCellInterface)
__global__ void k_flux_batch(Scheme s, int n, int d, const double* Vm, const double* Vp, const int* id, double* F,
                             DevCounters* cnt) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int err = 0;
  flux_num(c_mats[s.mset].m[id[q]], s.flux, s.del, d, Vm + 5 * q, Vp + 5 * q, F + 5 * q, err);
  if (err) atomicOr(&cnt->error, 1);
}

}  // namespace m2c
