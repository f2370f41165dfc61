// Grid layout and stencil kernels of the explicit step for gfx950.
//
// Layout: every field is structure-of-arrays FP64 (one array per component)
// over the ghosted brick (nx+2) x (ny+2) x (nz+2), i fastest; one ghost layer
// suffices because face states of ghost cells come from the boundary
// conditions, never from a limiter (Reconstructor.cpp:807-889).
//
// The stage kernel fuses, per cell, what the reference does in separate full
// passes for one RK stage: MUSCL reconstruction (Reconstructor.cpp:335-918),
// the reconstructed-state check (SpaceOperator.cpp:2420-2569), the face-flux
// loop (SpaceOperator.cpp:1432-2076), R /= -vol (:2628-2632), the RK
// combination (TimeIntegrator.cpp:319-715), C2P + clipping and the boundary
// fill of the new state (SpaceOperator.cpp:597-821).
#pragma once
#include "m2c_physics.h"

namespace m2c {


struct DevGrid {
  int NX, NY, NZ, SX, SY, SZ;
  long S;        // plane stride SX*SY
  long ncells;   // SX*SY*SZ: also the component stride of every 5-field
  // One geometry table (see g_axis / g_coefA / g_coefB): coordinates x, y, z
  // and spacings dx, dy, dz with ghosts (index i+1), then the Reconstructor
  // CoeffA and CoeffB of interior cells per axis (index i).
  const double* geo;
  const int* kt;  // CoeffK per axis (index i)
  int bc[6];
  double inlet[5], inlet2[5];
};

// coordinates (q = 0..2) or spacings (q = 3..5) along axis q % 3, ghosted
__device__ __host__ __forceinline__ const double* g_axis(const DevGrid& g, int q) {
  const int a = q % 3;
  long off = (q >= 3) ? (long)g.SX + g.SY + g.SZ : 0;
  if (a >= 1) off += g.SX;
  if (a >= 2) off += g.SY;
  return g.geo + off;
}
__device__ __forceinline__ const double* g_coefA(const DevGrid& g, int D) {
  long off = 2L * (g.SX + g.SY + g.SZ);
  if (D >= 1) off += g.NX;
  if (D >= 2) off += g.NY;
  return g.geo + off;
}
__device__ __forceinline__ const double* g_coefB(const DevGrid& g, int D) {
  return g_coefA(g, D) + (g.NX + g.NY + g.NZ);
}
__device__ __forceinline__ const int* g_coefK(const DevGrid& g, int D) {
  long off = 0;
  if (D >= 1) off += g.NX;
  if (D >= 2) off += g.NY;
  return g.kt + off;
}

struct Scheme {
  int flux, rec_type, limiter, slope_zero, integrator;
  int mp_flux, mp_recon, nmat, inactive;
  int var_type;                 // M2CX_VAR_*
  int mset;                     // slot of the solver's material table in c_mats
  double gmm, del, cfl;
  double time_step;             // Time.TimeStep: fixed dt when > 0
  int rnormal;                  // MultiPhase.RiemannNormal, M2CX_RNORMAL_*
  int phase_change;             // MultiPhase.PhaseChange, M2CX_PHASECHANGE_*
  int phase_dir;                // MultiPhase.PhaseChangeDirection, M2CX_PHASEDIR_*
  // per cell (ghosted index): 1 where a reconstruction in var_type fails
  // CheckState on some face and the cell falls back to primitive variables
  // (Reconstructor.cpp:722-778); written by k_recon_flags each stage
  const unsigned char* rflag;
};

__device__ __forceinline__ long gidx(const DevGrid& g, int i, int j, int k) {
  return ((long)(k + 1) * g.SY + (j + 1)) * g.SX + (i + 1);
}
// (i, j, k) of interior cell number q (i fastest), by 32-bit division: the
// interior cell count is checked below 2^31 when a solver is created
__device__ __forceinline__ void cell_ijk(const DevGrid& g, long q, int& i, int& j, int& k) {
  const unsigned u = (unsigned)q, nx = (unsigned)g.NX, ny = (unsigned)g.NY;
  const unsigned r = u / nx;
  i = (int)(u - r * nx);
  k = (int)(r / ny);
  j = (int)(r - (unsigned)k * ny);
}
__device__ __forceinline__ long axis_stride(const DevGrid& g, int D) { return D == 0 ? 1 : (D == 1 ? g.SX : g.S); }

// A 5-component field: components stored one after another, `cs` apart.
struct F5 {
  double* b;
  long cs;
  __device__ __forceinline__ double& operator()(int p, long c) const { return b[p * cs + c]; }
};
struct CF5 {
  const double* b;
  long cs;
  __device__ __forceinline__ double operator()(int p, long c) const { return b[p * cs + c]; }
};

__device__ __forceinline__ void load5(const CF5& f, long c, double* v) {
#pragma unroll
  for (int p = 0; p < 5; p++) v[p] = f(p, c);
}
__device__ __forceinline__ void store5(const F5& f, long c, const double* v) {
#pragma unroll
  for (int p = 0; p < 5; p++) f(p, c) = v[p];
}

// ---------------------------------------------------------------------------
// Limited slope of interior cell (i,j,k) along axis D for the reconstruction
// variables `vtype` (Reconstructor.cpp:472-648): one-sided differences of V
// or U = P2C(V) (each neighbour with its own material, :400-409), zero next
// to an inactive cell; for the characteristic types they are taken to
// characteristic variables about the cell state, limited component by
// component, and brought back.  q0 receives the cell value in the
// reconstructed variables (V or U).  Returns false when c0^2 < 0 at the cell
// (the reference aborts).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool recon_slope(const DevGrid& g, const Scheme& s, const DevMats& M, const CF5& V,
                                            const unsigned char* __restrict__ ID, int vtype, int D, int i, int j,
                                            int k, const double* v, double* q0, double* sig) {
  const long c = gidx(g, i, j, k);
  const long st = axis_stride(g, D);
  const int myid = ID[c], idm = ID[c - st], idp = ID[c + st];
  const DevEos& E = M.m[myid];
  double qm[5], qp[5];
  load5(V, c - st, qm);
  load5(V, c + st, qp);
  const bool cons = (vtype == M2CX_VAR_CONSERVATIVE || vtype == M2CX_VAR_CONSERVATIVE_CHARACTERISTIC);
  const bool chr = (vtype == M2CX_VAR_PRIMITIVE_CHARACTERISTIC || vtype == M2CX_VAR_CONSERVATIVE_CHARACTERISTIC);
  if (cons) {
    double t[5];
    eos_p2c(M.m[idm], qm, t);
#pragma unroll
    for (int p = 0; p < 5; p++) qm[p] = t[p];
    eos_p2c(M.m[idp], qp, t);
#pragma unroll
    for (int p = 0; p < 5; p++) qp[p] = t[p];
    eos_p2c(E, v, q0);
  } else {
#pragma unroll
    for (int p = 0; p < 5; p++) q0[p] = v[p];
  }
  double d0[5], d1[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    d0[p] = (idm == s.inactive) ? 0.0 : q0[p] - qm[p];
    d1[p] = (idp == s.inactive) ? 0.0 : qp[p] - q0[p];
  }
  CharBasis b;
  bool ok = true;
  if (chr) {
    ok = char_setup(E, v, D, b);
    double w[5];
    if (vtype == M2CX_VAR_PRIMITIVE_CHARACTERISTIC) {
      char_from_prim(b, v, d0, w);
#pragma unroll
      for (int p = 0; p < 5; p++) d0[p] = w[p];
      char_from_prim(b, v, d1, w);
#pragma unroll
      for (int p = 0; p < 5; p++) d1[p] = w[p];
    } else {
      char_left(b, v, d0, w);
#pragma unroll
      for (int p = 0; p < 5; p++) d0[p] = w[p];
      char_left(b, v, d1, w);
#pragma unroll
      for (int p = 0; p < 5; p++) d1[p] = w[p];
    }
  }
  const int a = (D == 0) ? i : (D == 1 ? j : k);
  const double A = g_coefA(g, D)[a];
  const double B = g_coefB(g, D)[a];
  const int K = g_coefK(g, D)[a];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    if (s.limiter == M2CX_LIM_GENERALIZED_MINMOD) sig[p] = lim_gmm(A, B, s.gmm, d0[p], d1[p]);
    else if (s.limiter == M2CX_LIM_VANALBADA) sig[p] = lim_va(A, B, K, d0[p], d1[p]);
    else sig[p] = 0.5 * (d0[p] + d1[p]);
  }
  if (chr) {
    double w[5];
    if (vtype == M2CX_VAR_PRIMITIVE_CHARACTERISTIC) char_to_prim(b, v, sig, w);
    else char_right(b, v, sig, w);
#pragma unroll
    for (int p = 0; p < 5; p++) sig[p] = w[p];
  }
  if (s.slope_zero && (myid != idm || myid != idp)) {
#pragma unroll
    for (int p = 0; p < 5; p++) sig[p] = 0.0;
  }
  return ok;
}

// One face (side 0 minus, 1 plus) from recon_slope's output (:665-714).
__device__ __forceinline__ void recon_face(const DevEos& E, int vtype, const double* v, const double* q0,
                                           const double* sig, int side, double* out) {
  if (vtype == M2CX_VAR_CONSERVATIVE || vtype == M2CX_VAR_CONSERVATIVE_CHARACTERISTIC) {
    double u2[5];
#pragma unroll
    for (int p = 0; p < 5; p++) u2[p] = side ? q0[p] + 0.5 * sig[p] : q0[p] - 0.5 * sig[p];
    eos_c2p(E, u2, out);
  } else {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = side ? v[p] + 0.5 * sig[p] : v[p] - 0.5 * sig[p];
  }
}

// ---------------------------------------------------------------------------
// Reconstructed face state of an interior cell q along axis D (side 0: the
// minus face, vl/vb/vk; side 1: the plus face, vr/vt/vf).
//
// The reference replaces only the FIRST failing face of a cell by the cell
// value (Reconstructor.cpp:722-778) and then clips every face, reverting
// clipped ones to the cell value and aborting on any remaining failure
// (SpaceOperator.cpp:2420-2569).  In every run that does not abort, this is
// the same as: a face reverts to the cell value iff it fails CheckState or
// clipping would modify it.  `fail`/`fixable` report the two conditions so
// the owning cell can detect the abort case.
//
// For the other variable types a failing face sends the whole cell back to
// primitive variables (the goto RETRY of :722-778); s.rflag marks those cells
// (k_recon_flags).  The faces of an unmarked cell all pass CheckState, so
// only clipping can revert them.  GEN = false compiles the primitive path
// alone.
// ---------------------------------------------------------------------------
template <bool GEN = true>
__device__ __forceinline__ void face_state(const DevGrid& g, const Scheme& s, const DevMats& M, const CF5& V,
                                           const unsigned char* __restrict__ ID, int D, int i, int j, int k,
                                           int side, double* out, bool& fail, bool& unfixable) {
  const long c = gidx(g, i, j, k);
  const long st = axis_stride(g, D);
  double v[5];
  load5(V, c, v);
  fail = false;
  unfixable = false;
  const int myid = ID[c];
  if (s.rec_type == M2CX_REC_CONSTANT || myid == s.inactive) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = v[p];
    if (s.rec_type == M2CX_REC_CONSTANT && myid != s.inactive) {
      const DevEos& E = M.m[myid];
      bool f = eos_check5(E, out), w = eos_would_clip(E, out);
      fail = f;
      unfixable = f && !w;
    }
    return;
  }
  if (GEN && s.var_type != M2CX_VAR_PRIMITIVE && s.rflag[c] == 0) {
    const DevEos& E = M.m[myid];
    double q0[5], sig[5];
    recon_slope(g, s, M, V, ID, s.var_type, D, i, j, k, v, q0, sig);
    recon_face(E, s.var_type, v, q0, sig, side, out);
    if (eos_would_clip(E, out)) {
#pragma unroll
      for (int p = 0; p < 5; p++) out[p] = v[p];
    }
    return;
  }
  const int idm = ID[c - st], idp = ID[c + st];
  double vm[5], vp[5];
  load5(V, c - st, vm);
  load5(V, c + st, vp);
  const int a = (D == 0) ? i : (D == 1 ? j : k);
  const double A = g_coefA(g, D)[a];
  const double B = g_coefB(g, D)[a];
  const int K = g_coefK(g, D)[a];
  const bool zero = s.slope_zero && (myid != idm || myid != idp);
#pragma unroll
  for (int p = 0; p < 5; p++) {
    double d0 = (idm == s.inactive) ? 0.0 : v[p] - vm[p];
    double d1 = (idp == s.inactive) ? 0.0 : vp[p] - v[p];
    double sig;
    if (s.limiter == M2CX_LIM_GENERALIZED_MINMOD) sig = lim_gmm(A, B, s.gmm, d0, d1);
    else if (s.limiter == M2CX_LIM_VANALBADA) sig = lim_va(A, B, K, d0, d1);
    else sig = 0.5 * (d0 + d1);
    if (zero) sig = 0.0;
    out[p] = side ? v[p] + 0.5 * sig : v[p] - 0.5 * sig;
  }
  const DevEos& E = M.m[myid];
  bool f = eos_check5(E, out);
  bool w = eos_would_clip(E, out);
  fail = f;
  unfixable = f && !w;
  if (f || w) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = v[p];
  }
}

// Face state on the ghost side of a domain boundary (Reconstructor.cpp:807-889):
// INLET/INLET2/OVERSET take the ghost value, OUTLET copies the image's face,
// SYMMETRY/SLIPWALL flip the normal velocity, STICKWALL flips all velocities.
__device__ __forceinline__ void ghost_face(int bc, int D, const double* vghost, const double* fimage, double* out) {
  if (bc == M2CX_BC_INLET || bc == M2CX_BC_INLET2 || bc == M2CX_BC_OVERSET) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = vghost[p];
  } else if (bc == M2CX_BC_SYMMETRY || bc == M2CX_BC_SLIPWALL) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = (p == D + 1) ? -fimage[p] : fimage[p];
  } else if (bc == M2CX_BC_STICKWALL) {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = (p >= 1 && p <= 3) ? -fimage[p] : fimage[p];
  } else {
#pragma unroll
    for (int p = 0; p < 5; p++) out[p] = fimage[p];
  }
}

// Both states of the face between cell (i,j,k) - e_D and (i,j,k), i.e. the
// minus face of (i,j,k).  `bits_m`/`bits_p` return fail/unfixable bits for the
// two owning cells (bit 0 fail, bit 1 unfixable).
template <bool GEN = true>
__device__ __forceinline__ void face_pair(const DevGrid& g, const Scheme& s, const DevMats& M, const CF5& V,
                                          const unsigned char* __restrict__ ID, int D, int i, int j, int k,
                                          double* Vm, double* Vp, int& bits_m, int& bits_p) {
  const int im = (D == 0) ? i - 1 : i, jm = (D == 1) ? j - 1 : j, km = (D == 2) ? k - 1 : k;
  const int N = (D == 0) ? g.NX : (D == 1 ? g.NY : g.NZ);
  const int a = (D == 0) ? i : (D == 1 ? j : k);
  bool f, u;
  bits_m = bits_p = 0;
  if (a == 0) {
    // ghost on the minus side
    double fp[5], vg[5];
    face_state<GEN>(g, s, M, V, ID, D, i, j, k, 0, Vp, f, u);
    bits_p = (f ? 1 : 0) | (u ? 2 : 0);
    load5(V, gidx(g, im, jm, km), vg);
    for (int p = 0; p < 5; p++) fp[p] = Vp[p];
    ghost_face(g.bc[2 * D], D, vg, fp, Vm);
  } else if (a == N) {
    // ghost on the plus side
    double fm[5], vg[5];
    face_state<GEN>(g, s, M, V, ID, D, im, jm, km, 1, Vm, f, u);
    bits_m = (f ? 1 : 0) | (u ? 2 : 0);
    load5(V, gidx(g, i, j, k), vg);
    for (int p = 0; p < 5; p++) fm[p] = Vm[p];
    ghost_face(g.bc[2 * D + 1], D, vg, fm, Vp);
  } else {
    face_state<GEN>(g, s, M, V, ID, D, im, jm, km, 1, Vm, f, u);
    bits_m = (f ? 1 : 0) | (u ? 2 : 0);
    face_state<GEN>(g, s, M, V, ID, D, i, j, k, 0, Vp, f, u);
    bits_p = (f ? 1 : 0) | (u ? 2 : 0);
  }
}

}  // namespace m2c
