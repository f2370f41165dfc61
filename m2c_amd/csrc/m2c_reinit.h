// Level-set reinitialization over the full domain on gfx950
// (LevelSetReinitializer::ReinitializeFullDomain, LevelSetReinitializer.cpp:
// 76-176): phi is driven towards a signed distance by pseudo-time RK3
// iterations of phi_t = -sign(phi0) (|grad phi| - 1) with Godunov upwinding
// (:960-1041), the first layer of nodes around the interface fixed, reset by
// the CR-1 / CR-2 averaging (:650-873) or forced each iteration by HCR-1 /
// HCR-2 (:1364-1455).
//
// Layout: the first layer is a compact list (cell index, neighbour set S as a
// 6-bit mask, |S|, the ratios r and r0, the correction f, the initial
// normal); every per-node step reads its neighbours from full fields and
// writes only its own node, so the list order does not matter.  The
// iteration loop runs on the device behind a done flag: each kernel of an
// iteration returns at once when the iterations have converged (or, after
// the first stage, when the stage-1 residual is below the tolerance), so the
// host launches chunks of iterations and reads the flag between chunks.
#pragma once
#include "m2c_multiphase.h"

namespace m2c {

struct ReinitState {
  unsigned long long res_bits, dphi_bits;  // running maxima (non-negative doubles as bits)
  int done, iters, nfl, pad;
  double last_dphi;
};

struct RBuf {
  unsigned char* tag = nullptr;
  double *sign = nullptr, *phibk = nullptr, *phi0 = nullptr, *pa = nullptr, *pb = nullptr, *normal = nullptr;
  int* fl_c = nullptr;               // ghosted cell index
  unsigned char* fl_s = nullptr;     // S as bits l, r, b, t, k, f
  int* fl_ns = nullptr;
  double *fl_r = nullptr, *fl_r0 = nullptr, *fl_f = nullptr, *fl_n = nullptr;  // r: 6, n: 3 per node
  int cap = 0;
  ReinitState* st = nullptr;
};

__device__ __forceinline__ double bits_d(unsigned long long b) { return __longlong_as_double((long long)b); }

__device__ __forceinline__ double wave_max(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(m, off, 64);
    m = (m < o) ? o : m;  // std::max(m, o)
  }
  return m;
}

// maximum of v >= 0 over the work-group (256 threads), then one atomic per
// work-group: atomics on one address serialize, so one per wave would cost
// more than the pass itself
__device__ __forceinline__ void block_max_atomic(unsigned long long* dst, double v) {
  __shared__ double red[4];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = (m < red[w]) ? red[w] : m;
    atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// Phi of the reference never holds edge or corner ghosts (cells outside the
// domain in two or more directions): they stay zero there.
__device__ __forceinline__ double phi_or_zero(const DevGrid& g, const double* phi, int i, int j, int k) {
  int out = (i < 0 || i >= g.NX) + (j < 0 || j >= g.NY) + (k < 0 || k >= g.NZ);
  return out >= 2 ? 0.0 : phi[gidx(g, i, j, k)];
}

// CentralDifferenceLocal (LevelSetReinitializer.h:118-123)
__device__ __forceinline__ double central_diff(double phi0, double phi1, double phi2, double x0, double x1,
                                               double x2) {
  double c0 = -(x2 - x1) / ((x1 - x0) * (x2 - x0));
  double c1 = 1.0 / (x1 - x0) - 1.0 / (x2 - x1);
  double c2 = (x1 - x0) / ((x2 - x0) * (x2 - x1));
  return c0 * phi0 + c1 * phi1 + c2 * phi2;
}

__global__ void k_rstate_reset(ReinitState* st, int all) {
  if (all) st->nfl = 0;
  st->res_bits = 0;
  st->dphi_bits = 0;
  st->done = 0;
  st->iters = 0;
  st->last_dphi = 0.0;
}

// TagFirstLayerNodes (:304-436) over the ghosted domain, edges and corners
// excluded; interior tagged nodes go to the list with S, r and r0.
__global__ void k_rtag(DevGrid g, const double* __restrict__ phi, RBuf b) {
  const long n = g.ncells;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const unsigned u = (unsigned)q, sx = (unsigned)g.SX, sy = (unsigned)g.SY;
    const unsigned rr = u / sx;
    const int i = (int)(u - rr * sx) - 1;
    const int k = (int)(rr / sy) - 1;
    const int j = (int)(rr - (unsigned)(k + 1) * sy) - 1;
    b.tag[q] = 0;
    const int out = (i < 0 || i >= g.NX) + (j < 0 || j >= g.NY) + (k < 0 || k >= g.NZ);
    if (out >= 2) continue;
    const double p = phi[q];
    bool hit = (i - 1 >= -1 && p * phi_or_zero(g, phi, i - 1, j, k) <= 0) ||
               (i + 1 <= g.NX && p * phi_or_zero(g, phi, i + 1, j, k) <= 0) ||
               (j - 1 >= -1 && p * phi_or_zero(g, phi, i, j - 1, k) <= 0) ||
               (j + 1 <= g.NY && p * phi_or_zero(g, phi, i, j + 1, k) <= 0) ||
               (k - 1 >= -1 && p * phi_or_zero(g, phi, i, j, k - 1) <= 0) ||
               (k + 1 <= g.NZ && p * phi_or_zero(g, phi, i, j, k + 1) <= 0);
    if (!hit) continue;
    b.tag[q] = 1;
    if (out) continue;
    const int slot = atomicAdd(&b.st->nfl, 1);
    if (slot >= b.cap) continue;
    const long nb[6] = {q - 1, q + 1, q - g.SX, q + g.SX, q - g.S, q + g.S};
    unsigned char s = 0;
    int ns = 0;
    double sum = 0.0;
    for (int d = 0; d < 6; d++) {
      const double pn = phi[nb[d]];
      double r = 0.0;
      if (p * pn < 0) {
        s |= (unsigned char)(1 << d);
        r = p / pn;
        ns++;
        sum += pn;
      }
      b.fl_r[6L * slot + d] = r;
    }
    b.fl_c[slot] = (int)q;
    b.fl_s[slot] = s;
    b.fl_ns[slot] = ns;
    b.fl_r0[slot] = (sum != 0) ? p / sum : 0.0;
  }
}

// EvaluateSignFunctionFullDomain (:601-622) with eps = 1, interior cells
// (the only ones the residual reads)
__global__ void k_rsign(DevGrid g, const double* __restrict__ phi, double* __restrict__ sign) {
  const long n = (long)g.NX * g.NY * g.NZ;
  const double* dx = g_axis(g, 3);
  const double* dy = g_axis(g, 4);
  const double* dz = g_axis(g, 5);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    const double dyz = (dz[k + 1] < dy[j + 1]) ? dz[k + 1] : dy[j + 1];
    const double factor = 1.0 * ((dyz < dx[i + 1]) ? dyz : dx[i + 1]);
    sign[c] = phi[c] / sqrt(phi[c] * phi[c] + factor * factor);
  }
}

// dst = 0*dst + 1*src over the whole ghosted brick (AXPlusBY with ghosts)
__global__ void k_raxpy_copy(double* __restrict__ dst, const double* __restrict__ src, long n) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
    dst[c] = 0.0 * dst[c] + 1.0 * src[c];
}

// DifferentiateInFirstLayer (:906-955)
__device__ double diff_first_layer(double x0, double x1, double x2, double tag0, double tag2, double phi0,
                                   double phi1, double phi2, double phi00, double phi3, double eps) {
  bool u0 = tag0 != 0, u2 = tag2 != 0;
  if (!u0 && !u2) return 0.0;
  const double d0 = phi1 - phi0, d1 = phi2 - phi1;
  const bool condB = d0 * d1 < 0 || phi0 * phi00 < 0 || phi2 * phi3 < 0;
  if (condB) {
    if (u0 && (phi0 * phi2 < 0) && fabs(d0 + eps) < fabs(d1)) u0 = false;
    if (u2 && (phi0 * phi2 < 0) && fabs(d1 + eps) < fabs(d0)) u2 = false;
  }
  if (!u0 && !u2) return 0.0;
  if (u0) {
    if (u2) return central_diff(phi0, phi1, phi2, x0, x1, x2);
    return (phi1 - phi0) / (x1 - x0);
  }
  return (phi2 - phi1) / (x2 - x1);
}

// ReinitializeFirstLayerNodes, first part (:660-719): phi = phig / |grad phig|
// with differences chosen by the tags.  phig is the copy of phi with ghosts;
// its second ghost layer (never filled in the reference) reads as zero, and
// the y difference takes j+2 at the top boundary as the reference does.
__global__ void k_rfirst(DevGrid g, RBuf b, double* __restrict__ phi, const double* __restrict__ phig) {
  const int n = min(b.st->nfl, b.cap);
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  const double* dx = g_axis(g, 3);
  const double* dy = g_axis(g, 4);
  const double* dz = g_axis(g, 5);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const long c = b.fl_c[q];
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    auto G = [&](int ii, int jj, int kk) -> double {
      if (ii < -1 || ii > g.NX || jj < -1 || jj > g.NY || kk < -1 || kk > g.NZ) return 0.0;
      return phig[gidx(g, ii, jj, kk)];
    };
    auto T = [&](int ii, int jj, int kk) -> double { return (double)b.tag[gidx(g, ii, jj, kk)]; };
    const double epsx = 1.0e-3 * dx[i + 1], epsy = 1.0e-3 * dy[j + 1], epsz = 1.0e-3 * dz[k + 1];
    double gr[3];
    gr[0] = diff_first_layer(X[i], X[i + 1], X[i + 2], T(i - 1, j, k), T(i + 1, j, k), G(i - 1, j, k), G(i, j, k),
                             G(i + 1, j, k), (i - 2 >= -1) ? G(i - 2, j, k) : G(i - 1, j, k),
                             (i + 2 <= g.NX) ? G(i + 2, j, k) : G(i + 1, j, k), epsx);
    gr[1] = diff_first_layer(Y[j], Y[j + 1], Y[j + 2], T(i, j - 1, k), T(i, j + 1, k), G(i, j - 1, k), G(i, j, k),
                             G(i, j + 1, k), (j - 2 >= -1) ? G(i, j - 2, k) : G(i, j - 1, k), G(i, j + 2, k), epsy);
    gr[2] = diff_first_layer(Z[k], Z[k + 1], Z[k + 2], T(i, j, k - 1), T(i, j, k + 1), G(i, j, k - 1), G(i, j, k),
                             G(i, j, k + 1), (k - 2 >= -1) ? G(i, j, k - 2) : G(i, j, k - 1),
                             (k + 2 <= g.NZ) ? G(i, j, k + 2) : G(i, j, k + 1), epsz);
    const double nrm = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    for (int d = 0; d < 3; d++) b.fl_n[3L * q + d] = gr[d] / nrm;
    phi[c] = (nrm == 0) ? phig[c] : phig[c] / nrm;
  }
}

// ComputeNormalDirectionBeyondFirstLayer, full-domain branch (:1136-1245):
// interior nodes off the first layer, then the first layer's own normals
__global__ void k_rnormal(DevGrid g, RBuf b, const double* __restrict__ phi) {
  const long n = (long)g.NX * g.NY * g.NZ;
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    if (b.tag[c] != 0) continue;
    double nv[3];
    nv[0] = central_diff(phi[c - 1], phi[c], phi[c + 1], X[i], X[i + 1], X[i + 2]);
    nv[1] = central_diff(phi[c - g.SX], phi[c], phi[c + g.SX], Y[j], Y[j + 1], Y[j + 2]);
    nv[2] = central_diff(phi[c - g.S], phi[c], phi[c + g.S], Z[k], Z[k + 1], Z[k + 2]);
    const double nr = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    if (nr != 0.0) {
      const double inv = 1.0 / nr;  // Vec3D::operator/= multiplies by the reciprocal
      nv[0] *= inv; nv[1] *= inv; nv[2] *= inv;
    }
    for (int d = 0; d < 3; d++) b.normal[3 * c + d] = nv[d];
  }
}

__global__ void k_rnormal_fl(RBuf b) {
  const int n = min(b.st->nfl, b.cap);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const long c = b.fl_c[q];
    for (int d = 0; d < 3; d++) b.normal[3 * c + d] = b.fl_n[3L * q + d];
  }
}

// ReinitializeFirstLayerNodes, steps 1-2 (:730-826): the correction set by the
// sign of the curvature, and the CR-1 / CR-2 averages; f = DBL_MAX outside it
__global__ void k_rcurv(DevGrid g, RBuf b, const double* __restrict__ phi, const double* __restrict__ phig,
                        int flt) {
  const int n = min(b.st->nfl, b.cap);
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const long c = b.fl_c[q];
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    const double* N = b.normal;
    const long sx = 1, sy = g.SX, sz = g.S;
    double kx, ky, kz;
    if (i - 1 >= 0 && i + 1 < g.NX)
      kx = central_diff(N[3 * (c - sx)], N[3 * c], N[3 * (c + sx)], X[i], X[i + 1], X[i + 2]);
    else if (i - 1 >= 0)
      kx = (N[3 * c] - N[3 * (c - sx)]) / (X[i + 1] - X[i]);
    else if (i + 1 < g.NX)
      kx = (N[3 * (c + sx)] - N[3 * c]) / (X[i + 2] - X[i + 1]);
    else
      kx = 0.0;
    if (j - 1 >= 0 && j + 1 < g.NY)
      ky = central_diff(N[3 * (c - sy) + 1], N[3 * c + 1], N[3 * (c + sy) + 1], Y[j], Y[j + 1], Y[j + 2]);
    else if (j - 1 >= 0)
      ky = (N[3 * c + 1] - N[3 * (c - sy) + 1]) / (Y[j + 1] - Y[j]);
    else if (j + 1 < g.NY)
      ky = (N[3 * (c + sy) + 1] - N[3 * c + 1]) / (Y[j + 2] - Y[j + 1]);
    else
      ky = 0.0;
    if (k - 1 >= 0 && k + 1 < g.NZ)
      kz = central_diff(N[3 * (c - sz) + 2], N[3 * c + 2], N[3 * (c + sz) + 2], Z[k], Z[k + 1], Z[k + 2]);
    else if (k - 1 >= 0)
      kz = (N[3 * c + 2] - N[3 * (c - sz) + 2]) / (Z[k + 1] - Z[k]);
    else if (k + 1 < g.NZ)
      kz = (N[3 * (c + sz) + 2] - N[3 * c + 2]) / (Z[k + 2] - Z[k + 1]);
    else
      kz = 0.0;
    const double curv = kx + ky + kz;
    const double pg = phig[c];
    if (!(curv * pg < 0 || (curv == 0.0 && pg < 0))) {
      b.fl_f[q] = DBL_MAX;
      continue;
    }
    const long nb[6] = {c - sx, c + sx, c - sy, c + sy, c - sz, c + sz};
    const unsigned s = b.fl_s[q];
    double f;
    if (flt == M2CX_REINIT_CONSTRAINED1) {
      double sum = 0.0;
      for (int d = 0; d < 6; d++)
        if (s & (1u << d)) sum += phi[nb[d]] / phig[nb[d]];
      f = (b.fl_ns[q] != 0) ? pg * sum / (double)b.fl_ns[q] : 0.0;
    } else {
      double s1 = 0.0, s2 = 0.0;
      for (int d = 0; d < 6; d++)
        if (s & (1u << d)) { s1 += phi[nb[d]]; s2 += phig[nb[d]]; }
      f = (s2 != 0) ? pg * s1 / s2 : 0.0;
    }
    b.fl_f[q] = f;
  }
}

// step 3 (:829-838)
__global__ void k_rapply(RBuf b, double* __restrict__ phi) {
  const int n = min(b.st->nfl, b.cap);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const double f = b.fl_f[q];
    if (f == DBL_MAX) continue;
    phi[b.fl_c[q]] = f;
  }
}

// ---------------------------------------------------------------------------
// pseudo-time iterations
// ---------------------------------------------------------------------------

// ComputeResidualFullDomain (:960-1041) at one interior cell; |local residual|
// for the maximum goes to *absr (negative when the node does not count)
__device__ __forceinline__ double reinit_residual(const DevGrid& g, const RBuf& b, const double* __restrict__ phi,
                                                  long c, int i, int j, int k, bool fix, double cfl,
                                                  double* absr) {
  *absr = -1.0;
  if (fix && b.tag[c] != 0) return 0.0;
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  const double* dxs = g_axis(g, 3);
  const double* dys = g_axis(g, 4);
  const double* dzs = g_axis(g, 5);
  const double dyz = (dzs[k + 1] < dys[j + 1]) ? dzs[k + 1] : dys[j + 1];
  const double dx = (dyz < dxs[i + 1]) ? dyz : dxs[i + 1];
  const double dt = cfl * dx;
  const double p = phi[c];
  const double a = (p - phi[c - 1]) / (X[i + 1] - X[i]);
  const double bb = (phi[c + 1] - p) / (X[i + 2] - X[i + 1]);
  const double cc = (p - phi[c - g.SX]) / (Y[j + 1] - Y[j]);
  const double d = (phi[c + g.SX] - p) / (Y[j + 2] - Y[j + 1]);
  const double e = (p - phi[c - g.S]) / (Z[k + 1] - Z[k]);
  const double f = (phi[c + g.S] - p) / (Z[k + 2] - Z[k + 1]);
  auto mx = [](double x, double y) { return (x < y) ? y : x; };  // std::max
  auto mn = [](double x, double y) { return (y < x) ? y : x; };  // std::min
  const double ap = mx(a, 0.0), am = mn(a, 0.0);
  const double bp = mx(bb, 0.0), bm = mn(bb, 0.0);
  const double cp = mx(cc, 0.0), cm = mn(cc, 0.0);
  const double dp = mx(d, 0.0), dm = mn(d, 0.0);
  const double ep = mx(e, 0.0), em = mn(e, 0.0);
  const double fp = mx(f, 0.0), fm = mn(f, 0.0);
  double lr;
  if (p >= 0)
    lr = sqrt(mx(ap * ap, bm * bm) + mx(cp * cp, dm * dm) + mx(ep * ep, fm * fm)) - 1.0;
  else
    lr = sqrt(mx(am * am, bp * bp) + mx(cm * cm, dp * dp) + mx(em * em, fp * fp)) - 1.0;
  double res;
  if (fabs(lr) <= 1.0)
    res = -dt * b.sign[c] * lr;
  else
    res = -dt * b.sign[c] * lr / fabs(lr);
  if (b.tag[c] == 0) *absr = fabs(lr);
  return res;
}

__device__ __forceinline__ bool reinit_skip(const ReinitState* st, double tol, bool after_stage1) {
  if (st->done) return true;
  return after_stage1 && bits_d(st->res_bits) < tol;
}

// RK3 stage 1 (:116-133): Phi0 = Phi, R(Phi) with its maximum, Phi1 = Phi + R
__global__ void k_rstage1(DevGrid g, RBuf b, const double* __restrict__ phi, int fix, double cfl) {
  if (reinit_skip(b.st, 0.0, false)) return;
  const long n = (long)g.NX * g.NY * g.NZ;
  double m = 0.0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    b.phi0[c] = 0.0 * b.phi0[c] + 1.0 * phi[c];
    double ar;
    const double r = reinit_residual(g, b, phi, c, i, j, k, fix != 0, cfl, &ar);
    m = (m < ar) ? ar : m;
    double o = 0.0 * b.pa[c] + 1.0 * phi[c];
    b.pa[c] = 1.0 * o + 1.0 * r;
  }
  block_max_atomic(&b.st->res_bits, m);
}

// stage 2 (:136-142): Phi2 = 0.25 Phi1 + 0.75 Phi + 0.25 R(Phi1)
__global__ void k_rstage2(DevGrid g, RBuf b, const double* __restrict__ phi, int fix, double cfl, double tol) {
  if (reinit_skip(b.st, tol, true)) return;
  const long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    double ar;
    const double r = reinit_residual(g, b, b.pa, c, i, j, k, fix != 0, cfl, &ar);
    double o = 0.25 * b.pa[c] + 0.75 * phi[c];
    b.pb[c] = 1.0 * o + 0.25 * r;
  }
}

// relative change of one node (CalculateMaximumRelativeErrorFullDomain,
// :1966-1994)
__device__ __forceinline__ double rel_change(double phi, double phi0) {
  double e = fabs(phi - phi0);
  if (phi != 0) e /= fabs(phi);
  return e;
}

// stage 3 (:145-150): Phi = 1/3 Phi + 2/3 Phi2 + 2/3 R(Phi2), in place, with
// the relative change of the nodes that keep this value: all of them, or
// with HCR forcing the nodes off the first layer (k_rcorr_apply takes the
// first layer after its correction)
__global__ void k_rstage3(DevGrid g, RBuf b, double* __restrict__ phi, int fix, int hcr, double cfl, double tol) {
  if (reinit_skip(b.st, tol, true)) return;
  const long n = (long)g.NX * g.NY * g.NZ;
  double m = 0.0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    const long c = gidx(g, i, j, k);
    double ar;
    const double r = reinit_residual(g, b, b.pb, c, i, j, k, fix != 0, cfl, &ar);
    double o = 1.0 / 3.0 * phi[c] + 2.0 / 3.0 * b.pb[c];
    o = 1.0 * o + 2.0 / 3.0 * r;
    phi[c] = o;
    if (!hcr || b.tag[c] == 0) {
      const double e = rel_change(o, b.phi0[c]);
      m = (m < e) ? e : m;
    }
  }
  block_max_atomic(&b.st->dphi_bits, m);
}

__global__ void k_rbc(DevGrid g, double* phi, int b0, int b1, int b2, int b3, int b4, int b5, const ReinitState* st,
                      double tol) {
  if (reinit_skip(st, tol, true)) return;
  const int bcs[6] = {b0, b1, b2, b3, b4, b5};
  ls_bc_faces(g, phi, bcs);
}

// ApplyCorrectionToFirstLayerNodes (:1364-1455), forcing f for HCR-1 / HCR-2
__global__ void k_rcorr_f(DevGrid g, RBuf b, const double* __restrict__ phi, int flt, double tol) {
  if (reinit_skip(b.st, tol, true)) return;
  const int n = min(b.st->nfl, b.cap);
  const double* dxs = g_axis(g, 3);
  const double* dys = g_axis(g, 4);
  const double* dzs = g_axis(g, 5);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const long c = b.fl_c[q];
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    const long nb[6] = {c - 1, c + 1, c - g.SX, c + g.SX, c - g.S, c + g.S};
    const unsigned s = b.fl_s[q];
    const double p = phi[c];
    double f = 0.0;
    bool out = false;
    for (int d = 0; d < 6; d++)
      if ((s & (1u << d)) && p * phi[nb[d]] >= 0) { out = true; break; }
    if (!out) {
      double sum = 0.0;
      if (flt == M2CX_REINIT_ITERATIVE_CONSTRAINED1) {
        for (int d = 0; d < 6; d++)
          if (s & (1u << d)) sum += b.fl_r[6L * q + d] * phi[nb[d]];
        f = (b.fl_ns[q] != 0) ? sum / (double)b.fl_ns[q] - p : 0.0 - p;
      } else {
        for (int d = 0; d < 6; d++)
          if (s & (1u << d)) sum += phi[nb[d]];
        f = b.fl_r0[q] * sum - p;
      }
      const double dyz = (dys[j + 1] < dzs[k + 1]) ? dzs[k + 1] : dys[j + 1];  // std::max(dy, dz)
      const double dmax = (dxs[i + 1] < dyz) ? dyz : dxs[i + 1];              // std::max(dx, .)
      f /= dmax;
    }
    b.fl_f[q] = f;
  }
}

// acc: fold the relative change of the corrected nodes into the iteration's
// maximum (the full-domain stage 3 leaves the first layer to this kernel)
__global__ void k_rcorr_apply(DevGrid g, RBuf b, double* __restrict__ phi, double cfl, double tol, int acc) {
  if (reinit_skip(b.st, tol, true)) return;
  const int n = min(b.st->nfl, b.cap);
  double m = 0.0;
  const double* dxs = g_axis(g, 3);
  const double* dys = g_axis(g, 4);
  const double* dzs = g_axis(g, 5);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const long c = b.fl_c[q];
    long r = c;
    const int i = (int)(r % g.SX) - 1; r /= g.SX;
    const int j = (int)(r % g.SY) - 1;
    const int k = (int)(r / g.SY) - 1;
    const double dyz = (dzs[k + 1] < dys[j + 1]) ? dzs[k + 1] : dys[j + 1];
    const double dmin = (dyz < dxs[i + 1]) ? dyz : dxs[i + 1];
    const double dt = cfl * dmin;
    phi[c] += dt * 0.5 * b.fl_f[q];
    const double e = rel_change(phi[c], b.phi0[c]);
    m = (m < e) ? e : m;
  }
  if (acc) block_max_atomic(&b.st->dphi_bits, m);
}

// end of one iteration (:118-128, :157-164): stop on a small residual before
// the update or a small relative change after it
__global__ void k_rcheck(ReinitState* st, double tol) {
  if (st->done) return;
  if (bits_d(st->res_bits) < tol) {
    st->done = 1;
  } else {
    st->iters += 1;
    const double d = bits_d(st->dphi_bits);
    st->last_dphi = d;
    if (d < tol) st->done = 1;
  }
  st->res_bits = 0;
  st->dphi_bits = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline int ri_alloc(RBuf& b, long ncells, int cap, bool normals, hipStream_t st) {
  if (!b.tag) {
    if (hipMalloc(&b.tag, ncells) != hipSuccess) return M2CX_E_OOM;
    for (double** f : {&b.sign, &b.phibk, &b.phi0, &b.pa, &b.pb})
      if (hipMalloc(f, ncells * 8) != hipSuccess) return M2CX_E_OOM;
    if (hipMalloc(&b.fl_c, (long)cap * 4) != hipSuccess || hipMalloc(&b.fl_s, cap) != hipSuccess ||
        hipMalloc(&b.fl_ns, (long)cap * 4) != hipSuccess || hipMalloc(&b.fl_r, (long)cap * 6 * 8) != hipSuccess ||
        hipMalloc(&b.fl_r0, (long)cap * 8) != hipSuccess || hipMalloc(&b.fl_f, (long)cap * 8) != hipSuccess ||
        hipMalloc(&b.fl_n, (long)cap * 3 * 8) != hipSuccess || hipMalloc(&b.st, sizeof(ReinitState)) != hipSuccess)
      return M2CX_E_OOM;
    // zero once: the fields are read by the first 0*x + y copies, and the
    // narrow band keeps the tags of nodes outside it (Tag starts at 0,
    // LevelSetReinitializer.cpp:37)
    for (double* f : {b.sign, b.phibk, b.phi0, b.pa, b.pb})
      if (hipMemsetAsync(f, 0, ncells * 8, st) != hipSuccess) return M2CX_E_HIP;
    if (hipMemsetAsync(b.tag, 0, ncells, st) != hipSuccess) return M2CX_E_HIP;
    b.cap = cap;
  }
  if (normals && !b.normal) {
    if (hipMalloc(&b.normal, ncells * 3 * 8) != hipSuccess) return M2CX_E_OOM;
    if (hipMemsetAsync(b.normal, 0, ncells * 3 * 8, st) != hipSuccess) return M2CX_E_HIP;
  }
  return 0;
}

inline void ri_free(RBuf& b) {
  for (void* p : {(void*)b.tag, (void*)b.sign, (void*)b.phibk, (void*)b.phi0, (void*)b.pa, (void*)b.pb,
                  (void*)b.normal, (void*)b.fl_c, (void*)b.fl_s, (void*)b.fl_ns, (void*)b.fl_r, (void*)b.fl_r0,
                  (void*)b.fl_f, (void*)b.fl_n, (void*)b.st})
    if (p) hipFree(p);
  b = RBuf{};
}

}  // namespace m2c
