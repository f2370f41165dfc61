// Learned initial guess for the exact two-material Riemann solver: a small
// MLP (8 -> 64 -> 64 -> 64 -> 1, SiLU) predicts p* of an interface problem;
// the guess replaces the acoustic-theory pressure that starts the bracketing
// of ExactRiemannSolverBase::FindInitialFeasiblePointsByAcousticTheory
// (ExactRiemannSolverBase.cpp:1107-1158), and the FP64 iteration that follows
// is unchanged.  A guess that is not a finite positive pressure falls back to
// the acoustic one.
//
// Inputs (nondimensionalised so that stiffened gases with very different p_c
// share one scale): with P = max(pL + pcL, pR + pcR), R = max(rhoL, rhoR) and
// C = sqrt(P / R),
//   x = [log(rhoL/R), log(rhoR/R), unL/C, unR/C, log((pL+pcL)/P),
//        log((pR+pcR)/P), gammaL, gammaR]
// and the output is y = log((p* + pcmin)/P), pcmin = min(pcL, pcR).
//
// Matrix products run on the MFMA units in BF16 with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16, one 16-row tile per wave); bias, SiLU, loss and
// the Adam update are FP32.  Training keeps FP32 master weights; the BF16
// copies are refreshed after every update.
#pragma once
#include <hip/hip_runtime.h>
#include "m2c_kernels.h"

namespace m2c {

constexpr int MLP_IN = 8, MLP_H = 64;
// flat parameter vector: W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
constexpr int MLP_OW1 = 0, MLP_OB1 = 512, MLP_OW2 = 576, MLP_OB2 = 4672, MLP_OW3 = 4736, MLP_OB3 = 8832,
              MLP_OW4 = 8896, MLP_OB4 = 8960, MLP_NPAR = 8961;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Device copies of the weights used by the MFMA products.
struct MLPDev {
  __bf16* wt1;  // [64][32]  W1 transposed, K padded 8 -> 32 with zeros
  __bf16* wt2;  // [64][64]  W2 transposed
  __bf16* wt3;  // [64][64]
  __bf16* wt4;  // [16][64]  W4 transposed, N padded 1 -> 16 with zeros
  __bf16* w2;   // [64][64]  W2 as stored ([in][out]), for dA = dZ W^T
  __bf16* w3;   // [64][64]
  const float* theta;  // FP32 master parameters (biases and W4 are read from here)
};

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l holds
//   A: row l%16, k = 8*(l/16) .. +7          B: col l%16, k = 8*(l/16) .. +7
//   C/D: rows 4*(l/16) .. +3, col l%16
__device__ __forceinline__ bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Features of one interface problem (see the header comment).  Returns false
// when the scales are not positive (no guess).
__device__ __forceinline__ bool mlp_features(const DevMats& M, const double* dir, const double* Vm, int idl,
                                             const double* Vp, int idr, float* x, double& P, double& pcmin) {
  const DevEos& L = M.m[idl];
  const DevEos& Rr = M.m[idr];
  const double pcl = (L.type == M2CX_EOS_SG || L.type == M2CX_EOS_NASG) ? L.pc : 0.0;
  const double pcr = (Rr.type == M2CX_EOS_SG || Rr.type == M2CX_EOS_NASG) ? Rr.pc : 0.0;
  const double gl = (L.type == M2CX_EOS_MG) ? L.Gamma0 + 1.0 : L.gam;
  const double gr = (Rr.type == M2CX_EOS_MG) ? Rr.Gamma0 + 1.0 : Rr.gam;
  const double ul = Vm[1] * dir[0] + Vm[2] * dir[1] + Vm[3] * dir[2];
  const double ur = Vp[1] * dir[0] + Vp[2] * dir[1] + Vp[3] * dir[2];
  const double ql = Vm[4] + pcl, qr = Vp[4] + pcr;
  P = fmax(ql, qr);
  const double R = fmax(Vm[0], Vp[0]);
  pcmin = fmin(pcl, pcr);
  if (!(P > 0) || !(R > 0) || !(ql > 0) || !(qr > 0) || !(Vm[0] > 0) || !(Vp[0] > 0)) return false;
  const double C = sqrt(P / R);
  x[0] = (float)log(Vm[0] / R);
  x[1] = (float)log(Vp[0] / R);
  x[2] = (float)(ul / C);
  x[3] = (float)(ur / C);
  x[4] = (float)log(ql / P);
  x[5] = (float)log(qr / P);
  x[6] = (float)gl;
  x[7] = (float)gr;
  return true;
}

// One dense layer for a 16-row tile held in LDS (bf16, row stride `ldin`):
// acc[nt] += in[16][K] x W[K][16 nt .. 16 nt + 15] for nt < NT, W given
// transposed (wt[n][K], row stride K).
template <int K, int NT>
__device__ __forceinline__ void mlp_tile_gemm(const __bf16* in, int ldin, const __bf16* wt, f32x4* acc) {
  const int l = threadIdx.x & 63;
  const int r = l & 15, kq = 8 * (l >> 4);
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kc = 0; kc < K; kc += 32) {
    const bf16x8 a = load8(in + r * ldin + kc + kq);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const bf16x8 b = load8(wt + (nt * 16 + r) * K + kc + kq);
      acc[nt] = mfma_bf16(a, b, acc[nt]);
    }
  }
}

// Hidden layer epilogue: z = acc + b, a = silu(z); a (bf16) to out[16][64],
// z (fp32, optional) to zs[16][64].
__device__ __forceinline__ void mlp_hidden_epilogue(const f32x4* acc, const float* bias, __bf16* out, float* zs) {
  const int l = threadIdx.x & 63;
  const int c = l & 15, r0 = 4 * (l >> 4);
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int col = nt * 16 + c, row = r0 + i;
      const float z = acc[nt][i] + bias[col];
      if (zs) zs[row * 64 + col] = z;
      out[row * 64 + col] = (__bf16)silu(z);
    }
}

// Forward pass of one 16-row tile: X (bf16 [16][32], zero padded) in x_lds;
// h1/h2 are bf16 [16][64] scratch; returns y for rows 4*(l/16)+i in lanes with
// l%16 == 0.  z1..z3 (fp32 [16][64]) and a1..a3 (bf16) are kept when given.
__device__ __forceinline__ void mlp_forward_tile(const MLPDev& W, const __bf16* x_lds, __bf16* a1, __bf16* a2,
                                                 __bf16* a3, float* z1, float* z2, float* z3, float* y4) {
  f32x4 acc[4];
  mlp_tile_gemm<32, 4>(x_lds, 32, W.wt1, acc);
  mlp_hidden_epilogue(acc, W.theta + MLP_OB1, a1, z1);
  __syncthreads();
  mlp_tile_gemm<64, 4>(a1, 64, W.wt2, acc);
  mlp_hidden_epilogue(acc, W.theta + MLP_OB2, a2, z2);
  __syncthreads();
  mlp_tile_gemm<64, 4>(a2, 64, W.wt3, acc);
  mlp_hidden_epilogue(acc, W.theta + MLP_OB3, a3, z3);
  __syncthreads();
  f32x4 o[1];
  mlp_tile_gemm<64, 1>(a3, 64, W.wt4, o);
  const int l = threadIdx.x & 63;
  if ((l & 15) == 0) {
    const float b4 = W.theta[MLP_OB4];
#pragma unroll
    for (int i = 0; i < 4; i++) y4[i] = o[0][i] + b4;
  }
}

// ---------------------------------------------------------------------------
// Guesses for the interface list of one stage: one wave per 16 faces.  The
// face states and the normal are the ones k_riemann will see (same
// reconstruction, same Riemann normal).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_mlp_guess(DevGrid g, Scheme s, CF5 V, const unsigned char* __restrict__ ID,
                                                  IfaceList L, MLPDev W, float* guess, const DevCounters* cnt) {
  __shared__ __align__(16) __bf16 xs[16 * 32];
  __shared__ __align__(16) __bf16 h1[16 * 64], h2[16 * 64], h3[16 * 64];
  __shared__ double sP[16], sPc[16];
  __shared__ int sOk[16];
  const int nif = min(cnt->nif, L.cap);
  const int l = threadIdx.x;
  for (int base = blockIdx.x * 16; base < nif; base += gridDim.x * 16) {
    for (int q = l; q < 16 * 32; q += 64) xs[q] = (__bf16)0.0f;
    __syncthreads();
    if (l < 16) {
      const int q = base + l;
      int ok = 0;
      double P = 1.0, pc = 0.0;
      if (q < nif) {
        const long c = L.cell[q];
        const int d = L.dir[q];
        const long st = axis_stride(g, d);
        double Vm[5], Vp[5];
        if_states(g, s, c_mats[s.mset], V, ID, d, c, Vm, Vp);
        // the direction the exact solver will solve along (Mesh, LevelSet
        // or Average normal), so the guess is for the same problem
        double dir[3];
        bimat_normal(g, s.rnormal, L, d, c, ID[c - st], ID[c], dir);
        float x[8];
        ok = mlp_features(c_mats[s.mset], dir, Vm, ID[c - st], Vp, ID[c], x, P, pc);
        if (ok)
          for (int f = 0; f < 8; f++) xs[l * 32 + f] = (__bf16)x[f];
      }
      sP[l] = P;
      sPc[l] = pc;
      sOk[l] = ok;
    }
    __syncthreads();
    float y[4];
    mlp_forward_tile(W, xs, h1, h2, h3, nullptr, nullptr, nullptr, y);
    if ((l & 15) == 0) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 4 * (l >> 4) + i, q = base + row;
        if (q < nif) guess[q] = sOk[row] ? (float)(exp((double)y[i]) * sP[row] - sPc[row]) : -1.0f;
      }
    }
    __syncthreads();
  }
}

// Batched features / targets of explicit problems (training data and tests):
// X[n][8], P[n], pcmin[n]; T[n] = log((p* + pcmin)/P) when pstar is given.
__global__ void k_mlp_features(int mset, int n, const double* dir, const double* Vm, const int* idm, const double* Vp,
                               const int* idp, const double* pstar, float* X, double* P, double* pcmin, float* T,
                               int* ok) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float x[8];
  double Pq = 1.0, pc = 0.0;
  int good = mlp_features(c_mats[mset], dir + 3 * q, Vm + 5 * q, idm[q], Vp + 5 * q, idp[q], x, Pq, pc);
  for (int f = 0; f < 8; f++) X[8L * q + f] = good ? x[f] : 0.0f;
  P[q] = Pq;
  pcmin[q] = pc;
  if (pstar) {
    const double a = (pstar[q] + pc) / Pq;
    if (!(a > 0) || !isfinite(a)) good = 0;
    T[q] = good ? (float)log(a) : 0.0f;
  }
  ok[q] = good;
}

// Prediction of explicit feature rows: y[n] (one wave per 16 rows).
__global__ void __launch_bounds__(64) k_mlp_predict(int n, const float* X, MLPDev W, float* y) {
  __shared__ __align__(16) __bf16 xs[16 * 32];
  __shared__ __align__(16) __bf16 h1[16 * 64], h2[16 * 64], h3[16 * 64];
  const int base = blockIdx.x * 16;
  const int l = threadIdx.x;
  for (int q = l; q < 16 * 32; q += 64) {
    const int r = q / 32, f = q % 32;
    xs[q] = (__bf16)((f < 8 && base + r < n) ? X[8L * (base + r) + f] : 0.0f);
  }
  __syncthreads();
  float yy[4];
  mlp_forward_tile(W, xs, h1, h2, h3, nullptr, nullptr, nullptr, yy);
  if ((l & 15) == 0)
    for (int i = 0; i < 4; i++) {
      const int q = base + 4 * (l >> 4) + i;
      if (q < n) y[q] = yy[i];
    }
}

// ---------------------------------------------------------------------------
// Training: one wave per 32 samples (two row tiles) computes the forward
// pass, the MSE loss and its gradient with respect to all parameters; each
// wave writes its partial gradient (fp32, MLP_NPAR values) to gpart[wave].
// ---------------------------------------------------------------------------
struct MLPTrainArgs {
  MLPDev W;
  const float* X;      // [n][8]
  const float* T;      // [n]
  const int* perm;     // sample order of this epoch
  int n, start, count; // this batch: perm[start .. start+count)
  float inv_batch;     // 1 / count
  float* gpart;        // [waves][MLP_NPAR]
  float* lpart;        // [waves] sum of squared errors
};

__global__ void __launch_bounds__(64) k_mlp_train_step(MLPTrainArgs a) {
  __shared__ __align__(16) __bf16 xs[32 * 32];
  __shared__ __align__(16) __bf16 A1[32 * 64], A2[32 * 64], A3[32 * 64], dZ[32 * 64];
  __shared__ float Z1[32 * 64], Z2[32 * 64], Z3[32 * 64];
  __shared__ float ys[32], dys[32];
  const int l = threadIdx.x;
  const int wave = blockIdx.x;
  const int s0 = wave * 32;
  const MLPDev& W = a.W;
  float* G = a.gpart + (long)wave * MLP_NPAR;
  for (int q = l; q < 32 * 32; q += 64) {
    const int r = q / 32, f = q % 32;
    const int s = s0 + r;
    float v = 0.0f;
    if (f < 8 && s < a.count) v = a.X[8L * a.perm[a.start + s] + f];
    xs[q] = (__bf16)v;
  }
  __syncthreads();
  // forward, two row tiles
  for (int rt = 0; rt < 2; rt++) {
    float y[4];
    mlp_forward_tile(W, xs + rt * 16 * 32, A1 + rt * 16 * 64, A2 + rt * 16 * 64, A3 + rt * 16 * 64,
                     Z1 + rt * 16 * 64, Z2 + rt * 16 * 64, Z3 + rt * 16 * 64, y);
    if ((l & 15) == 0)
      for (int i = 0; i < 4; i++) ys[rt * 16 + 4 * (l >> 4) + i] = y[i];
    __syncthreads();
  }
  // loss and dL/dy
  float sq = 0.0f;
  if (l < 32) {
    const int s = s0 + l;
    float d = 0.0f;
    if (s < a.count) {
      const float e = ys[l] - a.T[a.perm[a.start + s]];
      sq = e * e;
      d = 2.0f * e * a.inv_batch;
    }
    dys[l] = d;
  }
  // wave reduction of the loss (lanes >= 32 contribute 0)
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if (l == 0) a.lpart[wave] = sq;
  __syncthreads();
  // layer 4 (N = 1) on the VALU: dW4[k] = sum_s A3[s][k] dy[s]; db4 = sum dy
  {
    float gw = 0.0f;
    for (int s = 0; s < 32; s++) gw += (float)A3[s * 64 + l] * dys[s];
    G[MLP_OW4 + l] = gw;
    if (l == 0) {
      float gb = 0.0f;
      for (int s = 0; s < 32; s++) gb += dys[s];
      G[MLP_OB4] = gb;
    }
    // dZ3 = dy W4^T * silu'(Z3)   (W4 as used in the forward pass: bf16)
    const float w4 = (float)W.wt4[l];  // row 0 of W4^T
    float gb3 = 0.0f;
    for (int s = 0; s < 32; s++) {
      const __bf16 dz = (__bf16)(dys[s] * w4 * dsilu(Z3[s * 64 + l]));
      dZ[s * 64 + l] = dz;
      gb3 += (float)dz;
    }
    G[MLP_OB3 + l] = gb3;
  }
  __syncthreads();
  const int r = l & 15, kq = 8 * (l >> 4);
  // hidden layers 3 and 2: dW = A_prev^T dZ, dA_prev = dZ W^T, dZ_prev = dA_prev * silu'(Z_prev)
  for (int layer = 3; layer >= 2; layer--) {
    const __bf16* Aprev = (layer == 3) ? A2 : A1;
    const float* Zprev = (layer == 3) ? Z2 : Z1;
    const __bf16* Wl = (layer == 3) ? W.w3 : W.w2;
    const int ow = (layer == 3) ? MLP_OW3 : MLP_OW2;
    const int obp = (layer == 3) ? MLP_OB2 : MLP_OB1;
    // dW[m][n] = sum_s Aprev[s][m] dZ[s][n]: M = 64, N = 64, K = 32 samples
    for (int mt = 0; mt < 4; mt++) {
      bf16x8 av;
      for (int j = 0; j < 8; j++) av[j] = Aprev[(kq + j) * 64 + mt * 16 + r];
      for (int nt = 0; nt < 4; nt++) {
        bf16x8 bv;
        for (int j = 0; j < 8; j++) bv[j] = dZ[(kq + j) * 64 + nt * 16 + r];
        f32x4 acc = mfma_bf16(av, bv, f32x4{0.f, 0.f, 0.f, 0.f});
        for (int i = 0; i < 4; i++) G[ow + (mt * 16 + 4 * (l >> 4) + i) * 64 + nt * 16 + r] = acc[i];
      }
    }
    // dA_prev[s][k] = sum_n dZ[s][n] W[k][n]: M = 32 samples, N = 64, K = 64
    float dA[2][4][4];
    for (int rt = 0; rt < 2; rt++) {
      f32x4 acc[4];
      for (int nt = 0; nt < 4; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
      for (int kc = 0; kc < 64; kc += 32) {
        const bf16x8 av = load8(dZ + (rt * 16 + r) * 64 + kc + kq);
        for (int nt = 0; nt < 4; nt++) {
          const bf16x8 bv = load8(Wl + (nt * 16 + r) * 64 + kc + kq);
          acc[nt] = mfma_bf16(av, bv, acc[nt]);
        }
      }
      for (int nt = 0; nt < 4; nt++)
        for (int i = 0; i < 4; i++) dA[rt][nt][i] = acc[nt][i];
    }
    __syncthreads();  // every lane has read dZ
    for (int rt = 0; rt < 2; rt++)
      for (int nt = 0; nt < 4; nt++)
        for (int i = 0; i < 4; i++) {
          const int s = rt * 16 + 4 * (l >> 4) + i, k = nt * 16 + r;
          dZ[s * 64 + k] = (__bf16)(dA[rt][nt][i] * dsilu(Zprev[s * 64 + k]));
        }
    __syncthreads();
    // bias gradient of the previous layer: column sums of dZ_prev (as stored)
    {
      const int k = l;
      float sum = 0.0f;
      for (int s = 0; s < 32; s++) sum += (float)dZ[s * 64 + k];
      G[obp + k] = sum;
    }
    __syncthreads();
  }
  // layer 1: dW1[m][n] = sum_s X[s][m] dZ1[s][n], m < 8 (one 16-row tile, rows 8..15 unused)
  for (int nt = 0; nt < 4; nt++) {
    bf16x8 av, bv;
    for (int j = 0; j < 8; j++) av[j] = xs[(kq + j) * 32 + r];  // r < 8 real features, else padding zeros
    for (int j = 0; j < 8; j++) bv[j] = dZ[(kq + j) * 64 + nt * 16 + r];
    f32x4 acc = mfma_bf16(av, bv, f32x4{0.f, 0.f, 0.f, 0.f});
    for (int i = 0; i < 4; i++) {
      const int m = 4 * (l >> 4) + i;
      if (m < 8) G[MLP_OW1 + m * 64 + nt * 16 + r] = acc[i];
    }
  }
}

// Sum of the partial gradients (fixed order, deterministic) and one Adam step
// on the FP32 master parameters.
__global__ void k_mlp_adam(float* theta, float* m, float* v, const float* gpart, int waves, float lr, float b1,
                           float b2, float eps, float c1, float c2, float* gout, const float* lpart,
                           double* loss_acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && loss_acc) {
    double sum = 0.0;
    for (int w = 0; w < waves; w++) sum += lpart[w];
    *loss_acc += sum;
  }
  if (i >= MLP_NPAR) return;
  float gsum = 0.0f;
  for (int w = 0; w < waves; w++) gsum += gpart[(long)w * MLP_NPAR + i];
  if (gout) gout[i] = gsum;
  if (lr <= 0.0f) return;
  const float mi = b1 * m[i] + (1.0f - b1) * gsum;
  const float vi = b2 * v[i] + (1.0f - b2) * gsum * gsum;
  m[i] = mi;
  v[i] = vi;
  theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// guess = exp(y) P - pcmin for explicit problems (ok == 0: no guess)
__global__ void k_mlp_compose(int n, const float* y, const double* P, const double* pcmin, const int* ok,
                              float* guess) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  guess[q] = ok[q] ? (float)(exp((double)y[q]) * P[q] - pcmin[q]) : -1.0f;
}

// BF16 copies of the weight matrices from the FP32 master parameters.
__global__ void k_mlp_refresh(const float* theta, MLPDev W) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 64 * 32) {  // wt1[n][k], k < 8 real
    const int n = t / 32, k = t % 32;
    W.wt1[t] = (__bf16)(k < 8 ? theta[MLP_OW1 + k * 64 + n] : 0.0f);
  }
  if (t < 64 * 64) {
    const int n = t / 64, k = t % 64;
    W.wt2[t] = (__bf16)theta[MLP_OW2 + k * 64 + n];
    W.wt3[t] = (__bf16)theta[MLP_OW3 + k * 64 + n];
    W.w2[t] = (__bf16)theta[MLP_OW2 + t];
    W.w3[t] = (__bf16)theta[MLP_OW3 + t];
  }
  if (t < 16 * 64) {  // wt4[n][k], n == 0 real
    const int n = t / 64, k = t % 64;
    W.wt4[t] = (__bf16)(n == 0 ? theta[MLP_OW4 + k] : 0.0f);
  }
}

}  // namespace m2c
