// Narrow-band level set on gfx950 (LevelSetOperator.cpp:84-99, 152-196,
// 880-955; LevelSetReinitializer.cpp:181-300, 440-647, 1045-1133,
// 1459-2031): phi is advected and reinitialized only in a band of
// `bandwidth` face-neighbour layers around the interface, and cut off to
// +/- 10 x its extreme in-band value outside it.
//
// Layout: per ghosted cell the band level (int32, 31415927 outside) and
// byte masks `useful` (level <= bandwidth), `active` (level < bandwidth) and
// `plus1` (useful, or a populated face neighbour of a useful node); the band
// is rebuilt by breadth-first dilation, one launch per level, each reading
// only levels set by earlier launches.  The useful and plus-one nodes are
// also kept as compact index lists (stream compaction, order irrelevant:
// every list operation writes only its own node), so the per-iteration work
// of the reinitialization is proportional to the band, not the domain.
// The reinitializer's work fields persist between calls and each kernel
// writes exactly the nodes the reference writes, so values left over from
// earlier calls are read where the reference reads them.
#pragma once
#include <climits>
#include "m2c_reinit.h"

namespace m2c {

constexpr int NB_LARGE = 31415927;  // large_int (LevelSetReinitializer.cpp:14)

struct NBState {
  unsigned long long kmax, kmin;  // order-preserving keys of the running bounds
  double phi_max, phi_min, phi_out_pos, phi_out_neg;
  int n_useful, n_plus1;
};

struct NBBuf {
  int bw = 0;             // Bandwidth; 0: full domain
  bool built = false;
  double diag = 0.0;      // domain_diagonal (Main.cpp:207-209)
  int* level = nullptr;
  unsigned char *useful = nullptr, *active = nullptr, *plus1 = nullptr, *old_useful = nullptr;
  int *ulist = nullptr, *plist = nullptr;  // ghosted cell indices
  double* phig2 = nullptr;                 // PhiG2 of the reinitializer
  double* r = nullptr;                     // R of the reinitializer
  NBState* st = nullptr;
  // finite-volume level set in the band: each useful node's position in the
  // reference's useful-node list (it orders the sums of the face fluxes), and
  // the reconstruction outputs kept between calls (only the active nodes are
  // rewritten; faces of the others read what is left, zero at first)
  bool fv = false;
  int* pos = nullptr;
  int* key = nullptr;    // discovery key of the nodes of the layer being ordered
  int* slot = nullptr;   // scan workspace
  long slot_cap = 0;
  int* bsum = nullptr;   // tile sums of the scan
  long bsum_cap = 0;
  int* total = nullptr;  // scan total
  int n_list = 0;        // useful nodes (host copy, set with the positions)
  double *fvpf = nullptr, *fvuf = nullptr, *fvsl = nullptr;  // phi faces [6][n], velocity faces [6][n], slopes [3][n]
};

// doubles as unsigned keys with the order of their values (-0 below +0)
__device__ __forceinline__ unsigned long long nb_key(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double nb_unkey(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ void ghost_ijk(const DevGrid& g, long q, int& i, int& j, int& k) {
  const unsigned u = (unsigned)q, sx = (unsigned)g.SX, sy = (unsigned)g.SY;
  const unsigned rr = u / sx;
  i = (int)(u - rr * sx) - 1;
  k = (int)(rr / sy) - 1;
  j = (int)(rr - (unsigned)(k + 1) * sy) - 1;
}
__device__ __forceinline__ int out_count(const DevGrid& g, int i, int j, int k) {
  return (i < 0 || i >= g.NX) + (j < 0 || j >= g.NY) + (k < 0 || k >= g.NZ);
}
// inside the ghosted brick and not an edge or corner ghost (never populated)
__device__ __forceinline__ bool populated(const DevGrid& g, int i, int j, int k) {
  return i >= -1 && i <= g.NX && j >= -1 && j <= g.NY && k >= -1 && k <= g.NZ && out_count(g, i, j, k) < 2;
}
// UsefulG2 has two ghost layers; the second is never set
__device__ __forceinline__ bool useful_at(const DevGrid& g, const unsigned char* useful, int i, int j, int k) {
  if (i < -1 || i > g.NX || j < -1 || j > g.NY || k < -1 || k > g.NZ) return false;
  return useful[gidx(g, i, j, k)] != 0;
}

// sign change or zero product towards a populated face neighbour
__device__ __forceinline__ bool nb_crosses(const DevGrid& g, const double* phi, int i, int j, int k, long q) {
  const double p = phi[q];
  return (populated(g, i - 1, j, k) && p * phi[q - 1] <= 0) || (populated(g, i + 1, j, k) && p * phi[q + 1] <= 0) ||
         (populated(g, i, j - 1, k) && p * phi[q - g.SX] <= 0) ||
         (populated(g, i, j + 1, k) && p * phi[q + g.SX] <= 0) ||
         (populated(g, i, j, k - 1) && p * phi[q - g.S] <= 0) || (populated(g, i, j, k + 1) && p * phi[q + g.S] <= 0);
}

#define NB_GHOSTED_LOOP(q) for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < g.ncells; \
                               q += (long)gridDim.x * blockDim.x)

// ConstructNarrowBand, levels 0 and 1 (:1479-1537)
__global__ void k_nb_seed(DevGrid g, const double* __restrict__ phi, NBBuf nb) {
  NB_GHOSTED_LOOP(q) {
    int i, j, k;
    ghost_ijk(g, q, i, j, k);
    int lev = NB_LARGE;
    if (out_count(g, i, j, k) < 2) {
      if (phi[q] == 0)
        lev = 0;
      else if (nb_crosses(g, phi, i, j, k, q))
        lev = 1;
    }
    nb.level[q] = lev;
    nb.useful[q] = lev != NB_LARGE;
    nb.active[q] = lev != NB_LARGE;
  }
}

// UpdateNarrowBand, steps 1-2 (:1789-1821): the old band cleared, the tagged
// nodes of the old band (tags of this reinitialization) are the new levels 0
// and 1
__global__ void k_nb_reseed(DevGrid g, const double* __restrict__ phi, const unsigned char* __restrict__ tag,
                            NBBuf nb) {
  NB_GHOSTED_LOOP(q) {
    const unsigned char ou = nb.useful[q];
    nb.old_useful[q] = ou;
    int lev = NB_LARGE;
    if (ou && tag[q] == 1) lev = (phi[q] == 0) ? 0 : 1;
    nb.level[q] = lev;
    nb.useful[q] = lev != NB_LARGE;
    nb.active[q] = lev != NB_LARGE;
  }
}

// PropagateNarrowBand, one level (:1596-1728): a populated node without a
// level next to a node of a lower level.  Nodes set by this launch have
// level `band`, which the test `< band` does not see, so the launch may
// write and read levels in place.
__global__ void k_nb_grow(DevGrid g, NBBuf nb, int band) {
  NB_GHOSTED_LOOP(q) {
    if (nb.level[q] != NB_LARGE) continue;
    int i, j, k;
    ghost_ijk(g, q, i, j, k);
    if (out_count(g, i, j, k) >= 2) continue;
    const bool hit = (populated(g, i - 1, j, k) && nb.level[q - 1] < band) ||
                     (populated(g, i + 1, j, k) && nb.level[q + 1] < band) ||
                     (populated(g, i, j - 1, k) && nb.level[q - g.SX] < band) ||
                     (populated(g, i, j + 1, k) && nb.level[q + g.SX] < band) ||
                     (populated(g, i, j, k - 1) && nb.level[q - g.S] < band) ||
                     (populated(g, i, j, k + 1) && nb.level[q + g.S] < band);
    if (!hit) continue;
    nb.level[q] = band;
    nb.useful[q] = 1;
    nb.active[q] = band < nb.bw;
  }
}

// CreateUsefulNodesPlusOneLayer (:1890-1933) as a mask
__global__ void k_nb_plus1(DevGrid g, NBBuf nb) {
  NB_GHOSTED_LOOP(q) {
    unsigned char p = nb.useful[q];
    if (!p) {
      int i, j, k;
      ghost_ijk(g, q, i, j, k);
      if (out_count(g, i, j, k) < 2)
        p = (populated(g, i - 1, j, k) && nb.useful[q - 1]) || (populated(g, i + 1, j, k) && nb.useful[q + 1]) ||
            (populated(g, i, j - 1, k) && nb.useful[q - g.SX]) ||
            (populated(g, i, j + 1, k) && nb.useful[q + g.SX]) ||
            (populated(g, i, j, k - 1) && nb.useful[q - g.S]) || (populated(g, i, j, k + 1) && nb.useful[q + g.S]);
    }
    nb.plus1[q] = p;
  }
}

__global__ void k_nb_count_reset(NBState* st) {
  st->n_useful = 0;
  st->n_plus1 = 0;
}

// compact lists of the useful and plus-one nodes, one atomic per wave and list
__global__ void k_nb_lists(DevGrid g, NBBuf nb) {
  const int lane = threadIdx.x & 63;
  for (long base = blockIdx.x * (long)blockDim.x; base < g.ncells; base += (long)gridDim.x * blockDim.x) {
    const long q = base + threadIdx.x;
    const bool u = q < g.ncells && nb.useful[q];
    const bool p = q < g.ncells && nb.plus1[q];
    const unsigned long long mu = __ballot(u), mp = __ballot(p);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int ou = 0, op = 0;
    if (lane == 0) {
      if (mu) ou = atomicAdd(&nb.st->n_useful, __popcll(mu));
      if (mp) op = atomicAdd(&nb.st->n_plus1, __popcll(mp));
    }
    ou = __shfl(ou, 0, 64);
    op = __shfl(op, 0, 64);
    if (u) nb.ulist[ou + __popcll(mu & below)] = (int)q;
    if (p) nb.plist[op + __popcll(mp & below)] = (int)q;
  }
}

// the running bounds: reset to -/+ the diagonal (construction) or kept
__global__ void k_nb_keys(NBBuf nb, int reset) {
  NBState* st = nb.st;
  if (reset) {
    st->kmax = nb_key(-nb.diag);
    st->kmin = nb_key(nb.diag);
  } else {
    st->kmax = nb_key(st->phi_max);
    st->kmin = nb_key(st->phi_min);
  }
}

// max / min of phi over the useful nodes (CutOffPhiOutsideBand, :1741-1749),
// or over those inside the domain (UpdatePhiMaxAndPhiMinInBand, :1459-1476)
__global__ void __launch_bounds__(256) k_nb_minmax(DevGrid g, const double* __restrict__ phi, NBBuf nb,
                                                   int interior) {
  const int n = nb.st->n_useful;
  unsigned long long kx = 0ull, kn = ~0ull;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const long q = nb.ulist[t];
    if (interior) {
      int i, j, k;
      ghost_ijk(g, q, i, j, k);
      if (out_count(g, i, j, k)) continue;
    }
    const unsigned long long key = nb_key(phi[q]);
    kx = (key > kx) ? key : kx;
    kn = (key < kn) ? key : kn;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long ox = __shfl_xor(kx, off, 64), on = __shfl_xor(kn, off, 64);
    kx = (ox > kx) ? ox : kx;
    kn = (on < kn) ? on : kn;
  }
  if ((threadIdx.x & 63) == 0) {
    if (kx) atomicMax(&nb.st->kmax, kx);
    if (kn != ~0ull) atomicMin(&nb.st->kmin, kn);
  }
}

// bounds from the keys; with `cut` also the values outside the band
// (:1751-1757)
__global__ void k_nb_bounds(NBBuf nb, int cut) {
  NBState* st = nb.st;
  st->phi_max = nb_unkey(st->kmax);
  st->phi_min = nb_unkey(st->kmin);
  if (!cut) return;
  double pos = st->phi_max * 10.0, neg = st->phi_min * 10.0;
  if (pos <= 0) pos = nb.diag;  // no interface
  if (neg >= 0) neg = -nb.diag;
  st->phi_out_pos = pos;
  st->phi_out_neg = neg;
}

// CutOffPhiOutsideBand (:1759-1767), every ghosted node outside the band
__global__ void k_nb_cutoff(DevGrid g, double* __restrict__ phi, NBBuf nb) {
  const double pos = nb.st->phi_out_pos, neg = nb.st->phi_out_neg;
  NB_GHOSTED_LOOP(q) {
    if (!nb.useful[q]) phi[q] = phi[q] >= 0 ? pos : neg;
  }
}

// UpdateNarrowBand, steps 4-5 (:1833-1860): nodes that left the band are cut
// off (phi and R), nodes that entered it take the running bounds x 1.1 in
// place of a cut-off value
__global__ void k_nb_leave_enter(DevGrid g, double* __restrict__ phi, NBBuf nb) {
  const NBState* st = nb.st;
  const double pos = st->phi_out_pos, neg = st->phi_out_neg;
  NB_GHOSTED_LOOP(q) {
    if (nb.useful[q]) {
      const double p = phi[q];
      if (p > 0.9 * pos)
        phi[q] = st->phi_max * 1.1;
      else if (p < 0.9 * neg)
        phi[q] = st->phi_min * 1.1;
    } else if (nb.old_useful[q]) {
      phi[q] = (phi[q] >= 0) ? pos : neg;
      nb.r[q] = 0.0;
    }
  }
}

#define NB_LIST_LOOP(t, count) for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < (count); \
                                    t += gridDim.x * blockDim.x)

// TagFirstLayerNodesInBand (:440-598): tags of the useful nodes; the tagged
// nodes inside the domain go to the first-layer list with S, r and r0
__global__ void k_nb_rtag(DevGrid g, const double* __restrict__ phi, RBuf b, NBBuf nb) {
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long q = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, q, i, j, k);
    b.tag[q] = 0;
    if (out_count(g, i, j, k) >= 2 || !nb_crosses(g, phi, i, j, k, q)) continue;
    b.tag[q] = 1;
    if (out_count(g, i, j, k)) continue;
    const int slot = atomicAdd(&b.st->nfl, 1);
    if (slot >= b.cap) continue;
    const double p = phi[q];
    const long nbq[6] = {q - 1, q + 1, q - g.SX, q + g.SX, q - g.S, q + g.S};
    unsigned char s = 0;
    int ns = 0;
    double sum = 0.0;
    for (int d = 0; d < 6; d++) {
      const double pn = phi[nbq[d]];
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

// EvaluateSignFunctionInBand (:627-647), eps = 1
__global__ void k_nb_rsign(DevGrid g, const double* __restrict__ phi, double* __restrict__ sign, NBBuf nb) {
  const double* dx = g_axis(g, 3);
  const double* dy = g_axis(g, 4);
  const double* dz = g_axis(g, 5);
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long q = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, q, i, j, k);
    const double dyz = (dz[k + 1] < dy[j + 1]) ? dz[k + 1] : dy[j + 1];
    const double factor = 1.0 * ((dyz < dx[i + 1]) ? dyz : dx[i + 1]);
    sign[q] = phi[q] / sqrt(phi[q] * phi[q] + factor * factor);
  }
}

// reinitialization kernels after the first residual return at once when
// the iterations have converged (see reinit_skip)
struct NBSkip {
  const ReinitState* st;  // null: never skip
  double tol;
  bool after_stage1;
};
__device__ __forceinline__ bool nb_skip(const NBSkip& s) {
  return s.st && reinit_skip(s.st, s.tol, s.after_stage1);
}

// AXPlusBYInBandPlusOne (:1937-1962), and two of them in a row on the same
// nodes: X = a X + b Y, then X = 1.0 X + c Z when Z is given
__global__ void k_nb_axpy(DevGrid g, NBBuf nb, int ghosts, double* __restrict__ X, double a, const double* Y,
                          double b, const double* Z, double c, NBSkip sk) {
  if (nb_skip(sk)) return;
  const int n = nb.st->n_plus1;
  NB_LIST_LOOP(t, n) {
    const long q = nb.plist[t];
    if (!ghosts) {
      int i, j, k;
      ghost_ijk(g, q, i, j, k);
      if (out_count(g, i, j, k)) continue;
    }
    double o = a * X[q] + b * Y[q];
    if (Z) o = 1.0 * o + c * Z[q];
    X[q] = o;
  }
}

// PopulatePhiG2 over the useful nodes (:877-902)
__global__ void k_nb_phig2(NBBuf nb, const double* __restrict__ src) {
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long q = nb.ulist[t];
    nb.phig2[q] = src[q];
  }
}

// ComputeResidualInBand (:1045-1133): R on the useful nodes inside the
// domain, differences towards nodes outside the band zero; the largest
// |local residual| off the first layer when `want_max`
__global__ void __launch_bounds__(256) k_nb_rres(DevGrid g, RBuf b, NBBuf nb, const double* __restrict__ phi,
                                                 int fix, double cfl, int want_max, NBSkip sk) {
  if (nb_skip(sk)) return;
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  const double* dxs = g_axis(g, 3);
  const double* dys = g_axis(g, 4);
  const double* dzs = g_axis(g, 5);
  const unsigned char* U = nb.useful;
  auto mx = [](double x, double y) { return (x < y) ? y : x; };  // std::max
  auto mn = [](double x, double y) { return (y < x) ? y : x; };  // std::min
  const int n = nb.st->n_useful;
  double m = 0.0;
  NB_LIST_LOOP(t, n) {
    const long c = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    if (out_count(g, i, j, k)) continue;
    if (fix && b.tag[c] != 0) {
      nb.r[c] = 0.0;
      continue;
    }
    const double dyz = (dzs[k + 1] < dys[j + 1]) ? dzs[k + 1] : dys[j + 1];
    const double dx = (dyz < dxs[i + 1]) ? dyz : dxs[i + 1];
    const double dt = cfl * dx;
    const double p = phi[c];
    const double a = U[c - 1] ? (p - phi[c - 1]) / (X[i + 1] - X[i]) : 0.0;
    const double bb = U[c + 1] ? (phi[c + 1] - p) / (X[i + 2] - X[i + 1]) : 0.0;
    const double cc = U[c - g.SX] ? (p - phi[c - g.SX]) / (Y[j + 1] - Y[j]) : 0.0;
    const double d = U[c + g.SX] ? (phi[c + g.SX] - p) / (Y[j + 2] - Y[j + 1]) : 0.0;
    const double e = U[c - g.S] ? (p - phi[c - g.S]) / (Z[k + 1] - Z[k]) : 0.0;
    const double f = U[c + g.S] ? (phi[c + g.S] - p) / (Z[k + 2] - Z[k + 1]) : 0.0;
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
    if (fabs(lr) <= 1.0)
      nb.r[c] = -dt * b.sign[c] * lr;
    else
      nb.r[c] = -dt * b.sign[c] * lr / fabs(lr);
    if (b.tag[c] == 0) m = mx(m, fabs(lr));
  }
  if (want_max) block_max_atomic(&b.st->res_bits, m);
}

// CalculateMaximumRelativeErrorInBand (:1998-2031)
__global__ void __launch_bounds__(256) k_nb_relerr(DevGrid g, RBuf b, NBBuf nb, const double* __restrict__ phi,
                                                   NBSkip sk) {
  if (nb_skip(sk)) return;
  const int n = nb.st->n_useful;
  double m = 0.0;
  NB_LIST_LOOP(t, n) {
    const long c = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    if (out_count(g, i, j, k)) continue;
    const double e = rel_change(phi[c], b.phi0[c]);
    m = (m < e) ? e : m;
  }
  block_max_atomic(&b.st->dphi_bits, m);
}

__global__ void k_nb_bc(DevGrid g, double* phi, int b0, int b1, int b2, int b3, int b4, int b5,
                        const unsigned char* useful, NBSkip sk) {
  if (nb_skip(sk)) return;
  const int bcs[6] = {b0, b1, b2, b3, b4, b5};
  ls_bc_faces(g, phi, bcs, useful);
}

// The unit normal at an interior node of the band (the narrow-band branches
// of ComputeNormalDirectionBeyondFirstLayer, :1155-1208, and of
// LevelSetOperator::ComputeNormalDirectionCentralDifferencing_NarrowBand,
// LevelSetOperator.cpp:1545-1592): central differences where both
// neighbours are useful -- in y the reference tests the upper neighbour
// twice, so the lower one may lie outside the band -- one-sided where one
// is, zero where neither is; divided by its norm when that is nonzero.
__global__ void k_nb_rnormal(DevGrid g, RBuf b, NBBuf nb, const double* __restrict__ phi) {
  const double* X = g_axis(g, 0);
  const double* Y = g_axis(g, 1);
  const double* Z = g_axis(g, 2);
  const unsigned char* U = nb.useful;
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long c = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    if (out_count(g, i, j, k) || b.tag[c] != 0) continue;
    const double p = phi[c];
    double nv[3];
    if (U[c + 1] && U[c - 1])
      nv[0] = central_diff(phi[c - 1], p, phi[c + 1], X[i], X[i + 1], X[i + 2]);
    else if (U[c + 1])
      nv[0] = (phi[c + 1] - p) / (X[i + 2] - X[i + 1]);
    else if (U[c - 1])
      nv[0] = (p - phi[c - 1]) / (X[i + 1] - X[i]);
    else
      nv[0] = 0.0;
    if (U[c + g.SX] && U[c + g.SX])
      nv[1] = central_diff(phi[c - g.SX], p, phi[c + g.SX], Y[j], Y[j + 1], Y[j + 2]);
    else if (U[c + g.SX])
      nv[1] = (phi[c + g.SX] - p) / (Y[j + 2] - Y[j + 1]);
    else if (U[c - g.SX])
      nv[1] = (p - phi[c - g.SX]) / (Y[j + 1] - Y[j]);
    else
      nv[1] = 0.0;
    if (U[c + g.S] && U[c - g.S])
      nv[2] = central_diff(phi[c - g.S], p, phi[c + g.S], Z[k], Z[k + 1], Z[k + 2]);
    else if (U[c + g.S])
      nv[2] = (phi[c + g.S] - p) / (Z[k + 2] - Z[k + 1]);
    else if (U[c - g.S])
      nv[2] = (p - phi[c - g.S]) / (Z[k + 1] - Z[k]);
    else
      nv[2] = 0.0;
    const double nr = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    if (nr != 0.0) {
      const double inv = 1.0 / nr;  // Vec3D::operator/= multiplies by the reciprocal
      nv[0] *= inv; nv[1] *= inv; nv[2] *= inv;
    }
    for (int d = 0; d < 3; d++) b.normal[3 * c + d] = nv[d];
  }
}

// ---------------------------------------------------------------------------
// level-set advection in the band (LevelSetOperator.cpp:880-955, 152-196)
// ---------------------------------------------------------------------------

// ComputeResidualFDM_NarrowBand: R cleared on the useful nodes, computed on
// the active ones inside the domain with FD3 derivatives where the node two
// away is useful and one-sided differences otherwise
__global__ void k_nb_ls_residual(LSArgs a, NBBuf nb) {
  const DevGrid& g = a.g;
  const double* __restrict__ P = a.phi_st;
  const double* gx = g_axis(g, 0);
  const double* gy = g_axis(g, 1);
  const double* gz = g_axis(g, 2);
  const long sx = 1, sy = g.SX, sz = g.S;
  auto fd = [&](long off, int idx, long st, long c0) {
    const double* cc = a.fd3 + off + 4L * idx;
    return cc[0] * P[c0] + cc[1] * P[c0 + st] + cc[2] * P[c0 + 2 * st] + cc[3] * P[c0 + 3 * st];
  };
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long c = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    if (out_count(g, i, j, k) || !nb.active[c]) {
      a.rls[c] = 0.0;
      continue;
    }
    const double xi = gx[i + 1], yj = gy[j + 1], zk = gz[k + 1];
    const double u = a.V(1, c), v = a.V(2, c), w = a.V(3, c);
    double dm, dp;
    dm = useful_at(g, nb.useful, i - 2, j, k) ? fd(a.ocxL, i, sx, c - 2 * sx) : (P[c] - P[c - sx]) / (xi - gx[i]);
    dp = useful_at(g, nb.useful, i + 2, j, k) ? fd(a.ocxR, i, sx, c - sx) : (P[c + sx] - P[c]) / (gx[i + 2] - xi);
    double r = u >= 0 ? dm * u : dp * u;
    dm = useful_at(g, nb.useful, i, j - 2, k) ? fd(a.ocyL, j, sy, c - 2 * sy) : (P[c] - P[c - sy]) / (yj - gy[j]);
    dp = useful_at(g, nb.useful, i, j + 2, k) ? fd((i < g.NY - 1) ? a.ocyR3 : a.ocyR1, j, sy, c - sy)
                                              : (P[c + sy] - P[c]) / (gy[j + 2] - yj);
    r += v >= 0 ? dm * v : dp * v;
    dm = useful_at(g, nb.useful, i, j, k - 2) ? fd(a.oczL, k, sz, c - 2 * sz) : (P[c] - P[c - sz]) / (zk - gz[k]);
    dp = useful_at(g, nb.useful, i, j, k + 2) ? fd(a.oczR, k, sz, c - sz) : (P[c + sz] - P[c]) / (gz[k + 2] - zk);
    r += w >= 0 ? dm * w : dp * w;
    a.rls[c] = r * -1.0;
  }
}

// the RK combinations of TimeIntegrator.cpp on the useful nodes inside the
// domain (which: 0 FE, 1 RK stage 1, 2 RK2 final, 3 RK3 stage 2, 4 RK3 final)
__global__ void k_nb_ls_combine(DevGrid g, NBBuf nb, int which, double* __restrict__ phi,
                                double* __restrict__ phi1, const double* __restrict__ rls, const DevCounters* cnt) {
  const double dt = cnt->dt;
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long c = nb.ulist[t];
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    if (out_count(g, i, j, k)) continue;
    switch (which) {
      case 0: phi[c] = 1.0 * phi[c] + dt * rls[c]; break;
      case 1: {
        double o = 0.0 * phi1[c] + 1.0 * phi[c];
        phi1[c] = 1.0 * o + dt * rls[c];
      } break;
      case 2: {
        const double hdt = 0.5 * dt;
        double o = 0.5 * phi[c] + 0.5 * phi1[c];
        phi[c] = 1.0 * o + hdt * rls[c];
      } break;
      case 3: {
        const double qdt = 0.25 * dt;
        double o = 0.25 * phi1[c] + 0.75 * phi[c];
        phi1[c] = 1.0 * o + qdt * rls[c];
      } break;
      case 4: {
        const double tdt = 2.0 / 3.0 * dt;
        double o = 1.0 / 3.0 * phi[c] + 2.0 / 3.0 * phi1[c];
        phi[c] = 1.0 * o + tdt * rls[c];
      } break;
    }
  }
}

// Phi_tmp = Phi on the useful nodes, ghosts included (TimeIntegrator.cpp:
// 333-339 after the first step)
__global__ void k_nb_copy_useful(NBBuf nb, double* __restrict__ dst, const double* __restrict__ src) {
  const int n = nb.st->n_useful;
  NB_LIST_LOOP(t, n) {
    const long q = nb.ulist[t];
    dst[q] = 0.0 * dst[q] + 1.0 * src[q];
  }
}

// ---------------------------------------------------------------------------
// list order of the useful nodes (finite-volume level set in the band)
// ---------------------------------------------------------------------------
// ConstructNarrowBand pushes levels 0 and 1 in a (k, j, i) sweep of the
// ghosted brick (:1500-1533); UpdateNarrowBand pushes them in the order of the
// old list, which TagFirstLayerNodesInBand walks (:440-647, 1805-1821); then
// PropagateNarrowBand appends each level b in the order its nodes are first
// reached from the list: by the reaching node's position, then the direction
// (-x, +x, -y, +y, -z, +z) (:1596-1728).  On the device a node's position is
// a rank: in index order, in old-position order, or by the smallest key
// (parent position - layer base) * 6 + direction over its parents, each
// taken by an exclusive scan of 0/1 slots.
constexpr int NB_SCAN_TILE = 1024;  // 256 threads x 4 items

__global__ void __launch_bounds__(256) k_scan_tiles(int* d, long n, int* bsum) {
  __shared__ int s[256];
  const long base = (long)blockIdx.x * NB_SCAN_TILE + threadIdx.x * 4;
  int v[4], sum = 0;
  for (int e = 0; e < 4; e++) {
    v[e] = (base + e < n) ? d[base + e] : 0;
    sum += v[e];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int tv = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += tv;
    __syncthreads();
  }
  int run = s[threadIdx.x] - sum;
  for (int e = 0; e < 4; e++) {
    if (base + e < n) d[base + e] = run;
    run += v[e];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

// one work-group: exclusive scan of the tile sums, 256 at a time
__global__ void __launch_bounds__(256) k_scan_sums(int* bsum, long nt, int* total) {
  __shared__ int s[256];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (long b0 = 0; b0 < nt; b0 += 256) {
    const long q = b0 + threadIdx.x;
    const int v = (q < nt) ? bsum[q] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int tv = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += tv;
      __syncthreads();
    }
    if (q < nt) bsum[q] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 255) carry += s[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(int* d, long n, const int* __restrict__ bsum) {
  const long q0 = (long)blockIdx.x * NB_SCAN_TILE;
  const int add = bsum[blockIdx.x];
  for (int e = threadIdx.x; e < NB_SCAN_TILE; e += blockDim.x)
    if (q0 + e < n) d[q0 + e] += add;
}

__global__ void k_pos_flag_first(DevGrid g, NBBuf nb) {
  NB_GHOSTED_LOOP(q) nb.slot[q] = nb.level[q] <= 1;
}
__global__ void k_pos_set_first(DevGrid g, NBBuf nb) {
  NB_GHOSTED_LOOP(q) if (nb.level[q] <= 1) nb.pos[q] = nb.slot[q];
}
// levels 0 and 1 of an update are nodes of the old band (k_nb_reseed), whose
// positions are still in pos
// (positions outside [0, m) cannot occur; the test only keeps a broken
// invariant from writing outside the slots)
__global__ void k_pos_mark_update(DevGrid g, NBBuf nb, long m) {
  NB_GHOSTED_LOOP(q) {
    if (!(nb.old_useful[q] && nb.level[q] <= 1)) continue;
    const int p = nb.pos[q];
    if (p >= 0 && p < m) nb.slot[p] = 1;
  }
}
__global__ void k_pos_set_update(DevGrid g, NBBuf nb, long m) {
  NB_GHOSTED_LOOP(q) {
    if (!(nb.old_useful[q] && nb.level[q] <= 1)) continue;
    const int p = nb.pos[q];
    nb.pos[q] = (p >= 0 && p < m) ? nb.slot[p] : 0;
  }
}
__global__ void k_pos_key_init(DevGrid g, NBBuf nb, int band) {
  NB_GHOSTED_LOOP(q) if (nb.level[q] == band) nb.key[q] = INT_MAX;
}
__global__ void k_pos_keys(DevGrid g, NBBuf nb, int band, int base) {
  NB_GHOSTED_LOOP(q) {
    const int lp = nb.level[q];
    if (!(band == 2 ? lp <= 1 : lp == band - 1)) continue;
    int i, j, k;
    ghost_ijk(g, q, i, j, k);
    const int r = (nb.pos[q] - base) * 6;
    const int ni[6] = {i - 1, i + 1, i, i, i, i}, nj[6] = {j, j, j - 1, j + 1, j, j};
    const int nk[6] = {k, k, k, k, k - 1, k + 1};
    const long off[6] = {-1, 1, -(long)g.SX, (long)g.SX, -g.S, g.S};
    for (int d = 0; d < 6; d++) {
      if (!populated(g, ni[d], nj[d], nk[d])) continue;
      const long m = q + off[d];
      if (nb.level[m] == band) atomicMin(&nb.key[m], r + d);
    }
  }
}
// every node of level b has a parent of level b - 1, so its key is below m
__global__ void k_pos_mark_layer(DevGrid g, NBBuf nb, int band, long m) {
  NB_GHOSTED_LOOP(q) if (nb.level[q] == band && nb.key[q] < m) nb.slot[nb.key[q]] = 1;
}
__global__ void k_pos_set_layer(DevGrid g, NBBuf nb, int band, int base, long m) {
  NB_GHOSTED_LOOP(q) if (nb.level[q] == band) nb.pos[q] = base + ((nb.key[q] < m) ? nb.slot[nb.key[q]] : 0);
}

// ---------------------------------------------------------------------------
// finite-volume level set in the band (LevelSetOperator::ReconstructInBand,
// ComputeAdvectionFluxInBand, AddSourceTermInBand, LevelSetOperator.cpp:
// 1031-1062, 1160-1232, 1288-1306)
// ---------------------------------------------------------------------------
// reconstruction at the active nodes inside the domain (the stencils lie on
// useful nodes, where the reference's scalar field holds the velocity)
__global__ void k_nb_fv_recon(LSArgs a, NBBuf nb) {
  const DevGrid& g = a.g;
  const double* __restrict__ P = a.phi_st;
  const long n = g.ncells;
  NB_LIST_LOOP(t, nb.st->n_useful) {
    const long c = nb.ulist[t];
    if (!nb.active[c]) continue;
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    if (out_count(g, i, j, k)) continue;
    const int ijk[3] = {i, j, k};
    for (int D = 0; D < 3; D++) {
      const LSFaces f = ls_faces(a, P, D, c, ijk[D]);
      nb.fvuf[(2 * D) * n + c] = f.um;
      nb.fvuf[(2 * D + 1) * n + c] = f.up;
      nb.fvpf[(2 * D) * n + c] = f.pm;
      nb.fvpf[(2 * D + 1) * n + c] = f.pp;
      nb.fvsl[D * n + c] = f.slope;
    }
  }
}

// the face of an active face ghost towards the domain, by the mesh boundary
// condition, from its image's kept face (Reconstructor.cpp:807-880, 1048-1140)
__global__ void k_nb_fv_ghost(LSArgs a, NBBuf nb) {
  const DevGrid& g = a.g;
  const long n = g.ncells;
  const long nyz = (long)g.NY * g.NZ, nxz = (long)g.NX * g.NZ, nxy = (long)g.NX * g.NY;
  const long nfaces = 2 * (nyz + nxz + nxy);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nfaces; q += (long)gridDim.x * blockDim.x) {
    long r = q;
    int side, i, j, k;
    if (r < 2 * nyz) { side = (int)(r / nyz); r %= nyz; j = r % g.NY; k = r / g.NY; i = side == 0 ? -1 : g.NX; }
    else if ((r -= 2 * nyz) < 2 * nxz) { side = 2 + (int)(r / nxz); r %= nxz; i = r % g.NX; k = r / g.NX; j = side == 2 ? -1 : g.NY; }
    else { r -= 2 * nxz; side = 4 + (int)(r / nxy); r %= nxy; i = r % g.NX; j = r / g.NX; k = side == 4 ? -1 : g.NZ; }
    const long c = gidx(g, i, j, k);
    if (!nb.active[c]) continue;
    const int D = side / 2;
    const bool lo = (side % 2) == 0;
    const long im = c + (lo ? 1 : -1) * axis_stride(g, D);
    const int gf = lo ? 2 * D + 1 : 2 * D, imf = lo ? 2 * D : 2 * D + 1;
    double u, p;
    ls_ghost_face(a, a.phi_st, side, c, nb.fvuf[imf * n + im], nb.fvpf[imf * n + im], u, p);
    nb.fvuf[gf * n + c] = u;
    nb.fvpf[gf * n + c] = p;
  }
}

// the flux through the minus face of node d along axis D, from the kept faces
__device__ __forceinline__ double nb_fv_flux(const LSArgs& a, const NBBuf& nb, int D, long d, int idx) {
  const long n = a.g.ncells, m = d - axis_stride(a.g, D);
  return ls_face_flux(a, nb.fvpf[(2 * D + 1) * n + m], nb.fvpf[(2 * D) * n + d], nb.fvuf[(2 * D + 1) * n + m],
                      nb.fvuf[(2 * D) * n + d]) /
         g_axis(a.g, 3 + D)[idx + 1];
}

// R zero on the useful nodes; at an active node the fluxes the reference's
// list loop adds to it -- its own minus faces (x, y, z) at its own turn, each
// plus face at the turn of the neighbour across it -- in list order, then
// the source term inside the domain
__global__ void k_nb_fv_flux(LSArgs a, NBBuf nb) {
  const DevGrid& g = a.g;
  const double* __restrict__ P = a.phi_st;
  const long n = g.ncells;
  NB_LIST_LOOP(t, nb.st->n_useful) {
    const long c = nb.ulist[t];
    if (!nb.active[c]) {
      a.rls[c] = 0.0;
      continue;
    }
    int i, j, k;
    ghost_ijk(g, c, i, j, k);
    int tp[4], tk[4], nt = 0;
    if (i != -1 && j != -1 && k != -1) { tp[nt] = nb.pos[c]; tk[nt++] = 0; }
    if (i + 1 <= g.NX && j != -1 && k != -1 && k != g.NZ && j != g.NY && nb.useful[c + 1]) {
      tp[nt] = nb.pos[c + 1]; tk[nt++] = 1;
    }
    if (j + 1 <= g.NY && i != -1 && k != -1 && k != g.NZ && i != g.NX && nb.useful[c + g.SX]) {
      tp[nt] = nb.pos[c + g.SX]; tk[nt++] = 2;
    }
    if (k + 1 <= g.NZ && i != -1 && j != -1 && j != g.NY && i != g.NX && nb.useful[c + g.S]) {
      tp[nt] = nb.pos[c + g.S]; tk[nt++] = 3;
    }
    for (int x = 1; x < nt; x++)  // list order
      for (int y = x; y > 0 && tp[y - 1] > tp[y]; y--) {
        const int a0 = tp[y], a1 = tk[y];
        tp[y] = tp[y - 1]; tk[y] = tk[y - 1];
        tp[y - 1] = a0; tk[y - 1] = a1;
      }
    double r = 0.0;
    for (int x = 0; x < nt; x++) {
      switch (tk[x]) {
        case 0:
          if (k != g.NZ && j != g.NY) r += nb_fv_flux(a, nb, 0, c, i);
          if (k != g.NZ && i != g.NX) r += nb_fv_flux(a, nb, 1, c, j);
          if (j != g.NY && i != g.NX) r += nb_fv_flux(a, nb, 2, c, k);
          break;
        case 1: r -= nb_fv_flux(a, nb, 0, c + 1, i + 1); break;
        case 2: r -= nb_fv_flux(a, nb, 1, c + g.SX, j + 1); break;
        default: r -= nb_fv_flux(a, nb, 2, c + g.S, k + 1); break;
      }
    }
    if (!out_count(g, i, j, k)) r += P[c] * (nb.fvsl[0 * n + c] + nb.fvsl[1 * n + c] + nb.fvsl[2 * n + c]);
    a.rls[c] = r;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline int nb_alloc(NBBuf& nb, long ncells, hipStream_t st) {
  if (nb.level) return 0;
  if (nb.fv) {
    if (hipMalloc(&nb.pos, ncells * 4) != hipSuccess || hipMalloc(&nb.key, ncells * 4) != hipSuccess ||
        hipMalloc(&nb.total, 4) != hipSuccess || hipMalloc(&nb.fvpf, 6 * ncells * 8) != hipSuccess ||
        hipMalloc(&nb.fvuf, 6 * ncells * 8) != hipSuccess || hipMalloc(&nb.fvsl, 3 * ncells * 8) != hipSuccess)
      return M2CX_E_OOM;
    for (void* p : {(void*)nb.fvpf, (void*)nb.fvuf})
      if (hipMemsetAsync(p, 0, 6 * ncells * 8, st) != hipSuccess) return M2CX_E_HIP;
    if (hipMemsetAsync(nb.fvsl, 0, 3 * ncells * 8, st) != hipSuccess) return M2CX_E_HIP;
  }
  if (hipMalloc(&nb.level, ncells * 4) != hipSuccess) return M2CX_E_OOM;
  for (unsigned char** p : {&nb.useful, &nb.active, &nb.plus1, &nb.old_useful})
    if (hipMalloc(p, ncells) != hipSuccess) return M2CX_E_OOM;
  if (hipMalloc(&nb.ulist, ncells * 4) != hipSuccess || hipMalloc(&nb.plist, ncells * 4) != hipSuccess ||
      hipMalloc(&nb.phig2, ncells * 8) != hipSuccess || hipMalloc(&nb.r, ncells * 8) != hipSuccess ||
      hipMalloc(&nb.st, sizeof(NBState)) != hipSuccess)
    return M2CX_E_OOM;
  // zero, as a new SpaceVariable3D is
  for (void* p : {(void*)nb.phig2, (void*)nb.r})
    if (hipMemsetAsync(p, 0, ncells * 8, st) != hipSuccess) return M2CX_E_HIP;
  for (unsigned char* p : {nb.useful, nb.active, nb.plus1, nb.old_useful})
    if (hipMemsetAsync(p, 0, ncells, st) != hipSuccess) return M2CX_E_HIP;
  if (hipMemsetAsync(nb.st, 0, sizeof(NBState), st) != hipSuccess) return M2CX_E_HIP;
  return 0;
}

inline void nb_free(NBBuf& nb) {
  for (void* p : {(void*)nb.level, (void*)nb.useful, (void*)nb.active, (void*)nb.plus1, (void*)nb.old_useful,
                  (void*)nb.ulist, (void*)nb.plist, (void*)nb.phig2, (void*)nb.r, (void*)nb.st, (void*)nb.pos,
                  (void*)nb.key, (void*)nb.slot, (void*)nb.bsum, (void*)nb.total, (void*)nb.fvpf, (void*)nb.fvuf,
                  (void*)nb.fvsl})
    if (p) hipFree(p);
  const int bw = nb.bw;
  const double diag = nb.diag;
  const bool fv = nb.fv;
  nb = NBBuf{};
  nb.bw = bw;
  nb.diag = diag;
  nb.fv = fv;
}

// exclusive scan of d[0, n) in place; *total_host gets the sum
inline int nb_scan(NBBuf& nb, int* d, long n, hipStream_t st, int* total_host) {
  const long tiles = (n + NB_SCAN_TILE - 1) / NB_SCAN_TILE;
  if (tiles == 0) {
    *total_host = 0;
    return 0;
  }
  if (tiles > nb.bsum_cap) {
    if (nb.bsum) hipFree(nb.bsum);
    nb.bsum = nullptr;
    nb.bsum_cap = 0;
    if (hipMalloc(&nb.bsum, tiles * 4) != hipSuccess) return M2CX_E_OOM;
    nb.bsum_cap = tiles;
  }
  k_scan_tiles<<<(unsigned)tiles, 256, 0, st>>>(d, n, nb.bsum);
  k_scan_sums<<<1, 256, 0, st>>>(nb.bsum, tiles, nb.total);
  k_scan_add<<<(unsigned)tiles, 256, 0, st>>>(d, n, nb.bsum);
  if (hipMemcpyAsync(total_host, nb.total, 4, hipMemcpyDeviceToHost, st) != hipSuccess) return M2CX_E_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return M2CX_E_HIP;
  return 0;
}

inline int nb_slots(NBBuf& nb, long m, hipStream_t st) {
  if (m > nb.slot_cap) {
    if (nb.slot) hipFree(nb.slot);
    nb.slot = nullptr;
    nb.slot_cap = 0;
    if (hipMalloc(&nb.slot, (m > 0 ? m : 1) * 4) != hipSuccess) return M2CX_E_OOM;
    nb.slot_cap = m;
  }
  if (m > 0 && hipMemsetAsync(nb.slot, 0, m * 4, st) != hipSuccess) return M2CX_E_HIP;
  return 0;
}

// the useful nodes' positions in the reference's list (after the levels of a
// construction or an update are in place)
inline int nb_positions(NBBuf& nb, const DevGrid& g, bool update, hipStream_t st) {
  const unsigned B = (unsigned)blocks_for(g.ncells);
  int cnt = 0;
  if (!update) {
    if (int rc = nb_slots(nb, g.ncells, st)) return rc;
    k_pos_flag_first<<<B, 256, 0, st>>>(g, nb);
    if (int rc = nb_scan(nb, nb.slot, g.ncells, st, &cnt)) return rc;
    k_pos_set_first<<<B, 256, 0, st>>>(g, nb);
  } else {
    if (int rc = nb_slots(nb, nb.n_list, st)) return rc;
    const long m = nb.n_list;
    k_pos_mark_update<<<B, 256, 0, st>>>(g, nb, m);
    if (int rc = nb_scan(nb, nb.slot, m, st, &cnt)) return rc;
    k_pos_set_update<<<B, 256, 0, st>>>(g, nb, m);
  }
  int base = 0;
  for (int band = 2; band <= nb.bw; band++) {
    const long m = 6L * cnt;
    if (m >= INT_MAX) return M2CX_E_CAPACITY;
    if (int rc = nb_slots(nb, m, st)) return rc;
    k_pos_key_init<<<B, 256, 0, st>>>(g, nb, band);
    k_pos_keys<<<B, 256, 0, st>>>(g, nb, band, base);
    k_pos_mark_layer<<<B, 256, 0, st>>>(g, nb, band, m);
    int next = 0;
    if (int rc = nb_scan(nb, nb.slot, m, st, &next)) return rc;
    k_pos_set_layer<<<B, 256, 0, st>>>(g, nb, band, base + cnt, m);
    base += cnt;
    cnt = next;
  }
  nb.n_list = base + cnt;
  return 0;
}

// the band from levels 0 and 1 already in place: the higher levels, the
// plus-one layer and both lists
inline void nb_propagate(NBBuf& nb, const DevGrid& g, hipStream_t st) {
  for (int band = 2; band <= nb.bw; band++) k_nb_grow<<<blocks_for(g.ncells), 256, 0, st>>>(g, nb, band);
  k_nb_plus1<<<blocks_for(g.ncells), 256, 0, st>>>(g, nb);
  k_nb_count_reset<<<1, 1, 0, st>>>(nb.st);
  k_nb_lists<<<blocks_for(g.ncells), 256, 0, st>>>(g, nb);
}

// ConstructNarrowBand (LevelSetReinitializer.cpp:1479-1592)
inline int nb_construct(NBBuf& nb, const DevGrid& g, double* phi, hipStream_t st) {
  k_nb_seed<<<blocks_for(g.ncells), 256, 0, st>>>(g, phi, nb);
  nb_propagate(nb, g, st);
  if (nb.fv)
    if (int rc = nb_positions(nb, g, false, st)) return rc;
  k_nb_keys<<<1, 1, 0, st>>>(nb, 1);
  k_nb_minmax<<<blocks_for(g.ncells), 256, 0, st>>>(g, phi, nb, 0);
  k_nb_bounds<<<1, 1, 0, st>>>(nb, 1);
  k_nb_cutoff<<<blocks_for(g.ncells), 256, 0, st>>>(g, phi, nb);
  nb.built = true;
  return 0;
}

}  // namespace m2c
