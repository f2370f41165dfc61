// Host side of the C ABI declared in include/m2c_amd.h: owns the device
// fields of one solver and orders the kernels of the explicit step on its
// stream.  All compute runs in the HIP kernels of m2c_kernels.h /
// m2c_multiphase.h; there is no CPU fallback.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "m2c_kernels.h"
#include "m2c_multiphase.h"

using namespace m2c;

struct Timer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  double ms = 0.0;
  long long launches = 0;
};

struct m2cx_solver {
  m2cx_grid grid;
  m2cx_scheme sch;
  std::vector<m2cx_material> mats;
  int device = 0;
  hipStream_t stream = nullptr;
  DevGrid g{};
  Scheme s{};
  DevMats M{};
  long ncells = 0;
  // fields
  double* V[3][5] = {};   // V, V1, V2
  double* Un[5] = {};
  double* U1[5] = {};
  double* R[5] = {};
  unsigned char* ID = nullptr;
  unsigned char* IDn = nullptr;
  double* axis[6] = {};   // x, y, z, dx, dy, dz
  double* coefAB[6] = {}; // Ax, Ay, Az, Bx, By, Bz
  int* coefK[3] = {};
  double* fd3 = nullptr;  // level-set FD3 coefficient tables
  double* staging = nullptr;
  int* staging_i = nullptr;
  double* dt_part = nullptr;
  int dt_blocks = 0;
  DevCounters* cnt = nullptr;
  // multi-material
  MPBuffers mp{};
  bool multimat = false;
  // profiling
  bool profile = false;
  std::map<std::string, Timer> timers;
  std::string err;
};

static int fail(m2cx_solver* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, x)                                                                  \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail((h), M2CX_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

static F5 f5(double* const* a) { return F5{{a[0], a[1], a[2], a[3], a[4]}}; }
static CF5 cf5(double* const* a) { return CF5{{a[0], a[1], a[2], a[3], a[4]}}; }

static void tic(m2cx_solver* h, const char* name, hipEvent_t* a) {
  if (!h->profile) return;
  hipEventCreate(a);
  hipEventRecord(*a, h->stream);
  (void)name;
}
static void toc(m2cx_solver* h, const char* name, hipEvent_t a) {
  if (!h->profile) return;
  hipEvent_t b;
  hipEventCreate(&b);
  hipEventRecord(b, h->stream);
  h->timers[name].ev.push_back({a, b});
}

static void flush_timers(m2cx_solver* h) {
  for (auto& kv : h->timers) {
    for (auto& p : kv.second.ev) {
      float ms = 0;
      hipEventSynchronize(p.second);
      hipEventElapsedTime(&ms, p.first, p.second);
      kv.second.ms += ms;
      kv.second.launches++;
      hipEventDestroy(p.first);
      hipEventDestroy(p.second);
    }
    kv.second.ev.clear();
  }
}

static int grid_blocks(long n, int bs = 256) {
  long b = (n + bs - 1) / bs;
  return (int)std::min<long>(b, 256L * 64L);
}

// Reconstructor::CalculateSlopeLimiterCoefficientK (Reconstructor.cpp:103-121)
static int slope_k(int limiter, double A, double B) {
  if (limiter != M2CX_LIM_VANALBADA) return 0;
  int k;
  for (k = 2; k < 1000; k++) {
    double rhs = 2.0 / (1.0 + std::pow((k - 1.0) / k, k - 1) / k) * std::min(A, 1.0);
    if (B <= rhs) break;
  }
  return k;
}

extern "C" {

const char* m2cx_last_error(const m2cx_solver* h) { return h ? h->err.c_str() : "null handle"; }

int m2cx_create(const m2cx_grid* grid, const m2cx_material* mats, int nmat, const m2cx_scheme* sch, int device,
                m2cx_solver** out) {
  if (!grid || !mats || !sch || !out || nmat < 1 || nmat > M2C_MAX_MAT) return M2CX_E_ARG;
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return M2CX_E_ARG;
  if (sch->var_type != M2CX_VAR_PRIMITIVE || sch->n_levelsets > 1) return M2CX_E_UNSUPPORTED;
  auto* h = new m2cx_solver;
  *out = h;
  h->grid = *grid;
  h->sch = *sch;
  h->mats.assign(mats, mats + nmat);
  h->device = device;
  HIPCHK(h, hipSetDevice(device));
  HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const int NX = grid->nx, NY = grid->ny, NZ = grid->nz;
  DevGrid& g = h->g;
  g.NX = NX; g.NY = NY; g.NZ = NZ;
  g.SX = NX + 2; g.SY = NY + 2; g.SZ = NZ + 2;
  g.S = (long)g.SX * g.SY;
  g.ncells = g.S * g.SZ;
  h->ncells = g.ncells;
  for (int q = 0; q < 6; q++) g.bc[q] = grid->bc[q];
  for (int p = 0; p < 5; p++) { g.inlet[p] = grid->inlet[p]; g.inlet2[p] = grid->inlet2[p]; }
  // mesh (MeshGenerator.cpp:161-173, SpaceOperator.cpp:259-327)
  std::vector<double> ax[6];
  auto axis = [](int N, double a0, double a1, std::vector<double>& c, std::vector<double>& d) {
    double hh = (a1 - a0) / N;
    c.assign(N + 2, 0.0);
    d.assign(N + 2, hh);
    for (int i = 0; i < N; i++) c[i + 1] = a0 + 0.5 * hh + i * hh;
    c[0] = c[1] - d[1];
    c[N + 1] = c[N] + d[N];
  };
  axis(NX, grid->x0, grid->xmax, ax[0], ax[3]);
  axis(NY, grid->y0, grid->ymax, ax[1], ax[4]);
  axis(NZ, grid->z0, grid->zmax, ax[2], ax[5]);
  for (int q = 0; q < 6; q++) {
    HIPCHK(h, hipMalloc(&h->axis[q], ax[q].size() * 8));
    HIPCHK(h, hipMemcpy(h->axis[q], ax[q].data(), ax[q].size() * 8, hipMemcpyHostToDevice));
  }
  g.x = h->axis[0]; g.y = h->axis[1]; g.z = h->axis[2];
  g.dx = h->axis[3]; g.dy = h->axis[4]; g.dz = h->axis[5];
  // Reconstructor CoeffA/B/K per axis (Reconstructor.cpp:53-99)
  int Ns[3] = {NX, NY, NZ};
  for (int a = 0; a < 3; a++) {
    const std::vector<double>& d = ax[3 + a];
    std::vector<double> A(Ns[a]), B(Ns[a]);
    std::vector<int> K(Ns[a]);
    for (int i = 0; i < Ns[a]; i++) {
      A[i] = (d[i] + d[i + 1]) / (d[i + 1] + d[i + 2]);
      B[i] = 2.0 * d[i + 1] / (d[i + 1] + d[i + 2]);
      K[i] = slope_k(sch->limiter, A[i], B[i]);
    }
    HIPCHK(h, hipMalloc(&h->coefAB[a], Ns[a] * 8));
    HIPCHK(h, hipMalloc(&h->coefAB[3 + a], Ns[a] * 8));
    HIPCHK(h, hipMalloc(&h->coefK[a], Ns[a] * 4));
    HIPCHK(h, hipMemcpy(h->coefAB[a], A.data(), Ns[a] * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->coefAB[3 + a], B.data(), Ns[a] * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->coefK[a], K.data(), Ns[a] * 4, hipMemcpyHostToDevice));
  }
  g.Ax = h->coefAB[0]; g.Ay = h->coefAB[1]; g.Az = h->coefAB[2];
  g.Bx = h->coefAB[3]; g.By = h->coefAB[4]; g.Bz = h->coefAB[5];
  g.Kx = h->coefK[0]; g.Ky = h->coefK[1]; g.Kz = h->coefK[2];
  // scheme + materials
  Scheme& s = h->s;
  s.flux = sch->flux; s.rec_type = sch->rec_type; s.limiter = sch->limiter;
  s.slope_zero = (sch->slope_near_interface == M2CX_SLOPE_ZERO);
  s.integrator = sch->integrator; s.mp_flux = sch->mp_flux; s.mp_recon = sch->mp_recon;
  s.nmat = nmat; s.inactive = nmat;
  s.gmm = sch->gmm_coeff; s.del = sch->entropy_fix; s.cfl = sch->cfl;
  DevMats& M = h->M;
  M.nmat = nmat;
  for (int q = 0; q < nmat; q++) M.m[q] = make_eos(mats[q]);
  m2cx_material dm{};
  dm.eos = M2CX_EOS_DUMMY;
  M.m[nmat] = make_eos(dm);
  HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_mats), &M, sizeof(DevMats)));
  // fields
  const long n = h->ncells;
  for (int f = 0; f < 3; f++)
    for (int p = 0; p < 5; p++) HIPCHK(h, hipMalloc(&h->V[f][p], n * 8));
  for (int p = 0; p < 5; p++) {
    HIPCHK(h, hipMalloc(&h->Un[p], n * 8));
    HIPCHK(h, hipMalloc(&h->U1[p], n * 8));
    HIPCHK(h, hipMalloc(&h->R[p], n * 8));
    HIPCHK(h, hipMemsetAsync(h->U1[p], 0, n * 8, h->stream));
    HIPCHK(h, hipMemsetAsync(h->R[p], 0, n * 8, h->stream));
    HIPCHK(h, hipMemsetAsync(h->Un[p], 0, n * 8, h->stream));
    for (int f = 0; f < 3; f++) HIPCHK(h, hipMemsetAsync(h->V[f][p], 0, n * 8, h->stream));
  }
  HIPCHK(h, hipMalloc(&h->ID, n));
  HIPCHK(h, hipMalloc(&h->IDn, n));
  HIPCHK(h, hipMemsetAsync(h->ID, 0, n, h->stream));
  HIPCHK(h, hipMemsetAsync(h->IDn, 0, n, h->stream));
  HIPCHK(h, hipMalloc(&h->staging, n * 5 * 8));
  HIPCHK(h, hipMalloc(&h->staging_i, n * 4));
  h->dt_blocks = grid_blocks((long)NX * NY * NZ);
  HIPCHK(h, hipMalloc(&h->dt_part, h->dt_blocks * 8));
  HIPCHK(h, hipMalloc(&h->cnt, sizeof(DevCounters)));
  DevCounters c0{};
  c0.tmax = (sch->max_time > 0) ? sch->max_time : DBL_MAX;
  HIPCHK(h, hipMemcpy(h->cnt, &c0, sizeof c0, hipMemcpyHostToDevice));
  h->multimat = nmat > 1 || sch->n_levelsets > 0;
  if (h->multimat) {
    int rc = mp_create(h->mp, g, *sch, ax, h->stream);
    if (rc) return fail(h, rc, "multi-material buffers: " + std::string(hipGetErrorString(hipGetLastError())));
    
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_destroy(m2cx_solver* h) {
  if (!h) return 0;
  hipSetDevice(h->device);
  hipStreamSynchronize(h->stream);
  flush_timers(h);
  for (int f = 0; f < 3; f++)
    for (int p = 0; p < 5; p++) hipFree(h->V[f][p]);
  for (int p = 0; p < 5; p++) { hipFree(h->Un[p]); hipFree(h->U1[p]); hipFree(h->R[p]); }
  hipFree(h->ID); hipFree(h->IDn);
  for (auto* p : h->axis) hipFree(p);
  for (auto* p : h->coefAB) hipFree(p);
  for (auto* p : h->coefK) hipFree(p);
  hipFree(h->staging); hipFree(h->staging_i); hipFree(h->dt_part); hipFree(h->cnt);
  if (h->multimat) mp_destroy(h->mp);
  hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

static double** field_ptr(m2cx_solver* h, int field) {
  if (field >= M2CX_FIELD_V && field <= M2CX_FIELD_V2) return h->V[field];
  if (field == M2CX_FIELD_R) return h->R;
  return nullptr;
}

int m2cx_upload(m2cx_solver* h, int field, const void* host) {
  if (!h || !host) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const long n = h->ncells;
  if (field == M2CX_FIELD_ID) {
    HIPCHK(h, hipMemcpyAsync(h->staging_i, host, n * 4, hipMemcpyHostToDevice, h->stream));
    k_i32_to_u8<<<grid_blocks(n), 256, 0, h->stream>>>(h->staging_i, h->ID, n);
  } else if (field == M2CX_FIELD_PHI) {
    if (!h->mp.phi) return fail(h, M2CX_E_ARG, "no level set");
    HIPCHK(h, hipMemcpyAsync(h->mp.phi, host, n * 8, hipMemcpyHostToDevice, h->stream));
  } else {
    double** f = field_ptr(h, field);
    if (!f) return fail(h, M2CX_E_ARG, "bad field");
    HIPCHK(h, hipMemcpyAsync(h->staging, host, n * 5 * 8, hipMemcpyHostToDevice, h->stream));
    k_aos_to_soa<<<grid_blocks(n), 256, 0, h->stream>>>(h->staging, f5(f), 5, n);
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_download(m2cx_solver* h, int field, void* host) {
  if (!h || !host) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const long n = h->ncells;
  if (field == M2CX_FIELD_ID) {
    k_u8_to_i32<<<grid_blocks(n), 256, 0, h->stream>>>(h->ID, h->staging_i, n);
    HIPCHK(h, hipMemcpyAsync(host, h->staging_i, n * 4, hipMemcpyDeviceToHost, h->stream));
  } else if (field == M2CX_FIELD_PHI) {
    if (!h->mp.phi) return fail(h, M2CX_E_ARG, "no level set");
    HIPCHK(h, hipMemcpyAsync(host, h->mp.phi, n * 8, hipMemcpyDeviceToHost, h->stream));
  } else {
    double** f = field_ptr(h, field);
    if (!f) return fail(h, M2CX_E_ARG, "bad field");
    k_soa_to_aos<<<grid_blocks(n), 256, 0, h->stream>>>(cf5(f), h->staging, 5, n);
    HIPCHK(h, hipMemcpyAsync(host, h->staging, n * 5 * 8, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

static int check_error(m2cx_solver* h) {
  DevCounters c;
  HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (c.error & 1) return fail(h, M2CX_E_HYPERBOLICITY, "state violates hyperbolicity (reference aborts)");
  if (c.error & 2) return fail(h, M2CX_E_RIEMANN_FATAL, "exact Riemann solver reached a fatal state");
  if (c.error & 4) return fail(h, M2CX_E_CAPACITY, "device work-list capacity exceeded");
  if (c.error & 8) return fail(h, M2CX_E_UNRESOLVED, "unresolved phase-change cells");
  return 0;
}

int m2cx_set_time(m2cx_solver* h, double t, int time_step) {
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  c.t = t;
  c.time_step = time_step;
  HIPCHK(h, hipMemcpy(h->cnt, &c, sizeof c, hipMemcpyHostToDevice));
  return 0;
}

int m2cx_get_time(m2cx_solver* h, double* t, double* dt, int* time_step) {
  DevCounters c;
  HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (t) *t = c.t;
  if (dt) *dt = c.dt;
  if (time_step) *time_step = c.time_step;
  return 0;
}

static StageArgs stage_args(m2cx_solver* h, int vin, int vout, int mode) {
  StageArgs a{};
  a.g = h->g;
  a.s = h->s;
  a.Vin = cf5(h->V[vin]);
  a.Vout = (vout >= 0) ? f5(h->V[vout]) : F5{};
  a.Un = f5(h->Un);
  a.U1 = f5(h->U1);
  a.R = f5(h->R);
  a.ID = h->ID;
  if (h->multimat) {
    for (int d = 0; d < 3; d++) a.fidx[d] = h->mp.fidx[d];
    a.L = h->mp.L;
  }
  a.mode = mode;
  a.cnt = h->cnt;
  return a;
}

static void launch_stage(m2cx_solver* h, int vin, int vout, int mode) {
  StageArgs a = stage_args(h, vin, vout, mode);
  dim3 grid((h->g.NX + 1 + TX - 1) / TX, (h->g.NY + 1 + TY - 1) / TY, (h->g.NZ + 1 + TZ - 1) / TZ);
  // the tile grid must cover index NX (the last face); a tile starting past
  // the domain end has no work and returns after the face phase
  grid.x = (h->g.NX + TX - 1) / TX;
  grid.y = (h->g.NY + TY - 1) / TY;
  grid.z = (h->g.NZ + TZ - 1) / TZ;
  hipEvent_t e0{};
  tic(h, "stage", &e0);
  switch (mode) {
    case MODE_RESIDUAL: k_stage<MODE_RESIDUAL><<<grid, NT, 0, h->stream>>>(a); break;
    case MODE_STAGE1: k_stage<MODE_STAGE1><<<grid, NT, 0, h->stream>>>(a); break;
    case MODE_RK2_FINAL: k_stage<MODE_RK2_FINAL><<<grid, NT, 0, h->stream>>>(a); break;
    case MODE_RK3_STAGE2: k_stage<MODE_RK3_STAGE2><<<grid, NT, 0, h->stream>>>(a); break;
    case MODE_RK3_FINAL: k_stage<MODE_RK3_FINAL><<<grid, NT, 0, h->stream>>>(a); break;
    case MODE_FE: k_stage<MODE_FE><<<grid, NT, 0, h->stream>>>(a); break;
  }
  toc(h, "stage", e0);
}

// Interface Riemann problems of this stage (multi-material only).
static void interface_stage(m2cx_solver* h, int vin, int stage1) {
  if (!h->multimat) return;
  hipEvent_t e0{};
  tic(h, "riemann", &e0);
  mp_riemann(h->mp, h->g, h->s, cf5(h->V[vin]), h->ID, stage1, h->cnt, h->stream);
  toc(h, "riemann", e0);
}

static void levelset_stage(m2cx_solver* h, int vin, int which) {
  if (!h->multimat || h->sch.n_levelsets == 0) return;
  hipEvent_t e0{};
  tic(h, "levelset", &e0);
  mp_levelset(h->mp, h->g, cf5(h->V[vin]), which, h->cnt, h->stream);
  toc(h, "levelset", e0);
}

int m2cx_compute_residual(m2cx_solver* h, int v_field, int capture_riemann) {
  if (!h || v_field < 0 || v_field > 2) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  if (h->multimat) mp_classify(h->mp, h->g, h->s, h->ID, h->cnt, h->stream);
  interface_stage(h, v_field, capture_riemann);
  launch_stage(h, v_field, -1, MODE_RESIDUAL);
  HIPCHK(h, hipGetLastError());
  int rc = check_error(h);
  if (rc) return rc;
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  return (int)c.riemann_errors;
}

int m2cx_compute_time_step(m2cx_solver* h, double* dt, double* cfl) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  k_dt<<<h->dt_blocks, 256, 0, h->stream>>>(h->g, h->s, cf5(h->V[0]), h->ID, h->dt_part, h->cnt, 0);
  HIPCHK(h, hipGetLastError());
  int rc = check_error(h);
  if (rc) return rc;
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  if (dt) *dt = c.dt;
  if (cfl) *cfl = c.cfl;
  return 0;
}

int m2cx_apply_boundary_conditions(m2cx_solver* h, int v_field) {
  double** f = field_ptr(h, v_field);
  if (!f) return M2CX_E_ARG;
  k_apply_bc<<<grid_blocks((long)h->g.NX * h->g.NY * h->g.NZ), 256, 0, h->stream>>>(h->g, f5(f));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_clip(m2cx_solver* h, int v_field) {
  double** f = field_ptr(h, v_field);
  if (!f) return M2CX_E_ARG;
  k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
  k_clip<<<grid_blocks((long)h->g.NX * h->g.NY * h->g.NZ), 256, 0, h->stream>>>(h->g, f5(f), h->ID, h->cnt);
  int rc = check_error(h);
  if (rc) return rc;
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  return (int)c.stage_clipped;
}

// One TimeIntegrator step with dt already in cnt->dt (TimeIntegrator.cpp:167-715)
static void step_kernels(m2cx_solver* h) {
  const int ti = h->sch.integrator;
  const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
  if (h->multimat) mp_begin_step(h->mp, h->g, h->s, h->ID, h->cnt, h->stream);
  if (ti == M2CX_TI_FORWARD_EULER) {
    interface_stage(h, 0, 1);
    launch_stage(h, 0, 1, MODE_FE);
    // V <- new state (pointer swap), then the level set with the new
    // velocities (TimeIntegrator.cpp:199-214)
    for (int p = 0; p < 5; p++) std::swap(h->V[0][p], h->V[1][p]);
    levelset_stage(h, 0, 0);
  } else {
    interface_stage(h, 0, 1);
    k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
    launch_stage(h, 0, 1, MODE_STAGE1);
    k_resync<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, cf5(h->V[1]), f5(h->U1), h->ID, h->cnt);
    levelset_stage(h, 0, 1);
    if (ti == M2CX_TI_RK2) {
      interface_stage(h, 1, 0);
      launch_stage(h, 1, 0, MODE_RK2_FINAL);
      levelset_stage(h, 1, 2);
    } else {
      interface_stage(h, 1, 0);
      k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
      launch_stage(h, 1, 2, MODE_RK3_STAGE2);
      k_resync<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, cf5(h->V[2]), f5(h->U1), h->ID, h->cnt);
      levelset_stage(h, 1, 3);
      interface_stage(h, 2, 0);
      launch_stage(h, 2, 0, MODE_RK3_FINAL);
      levelset_stage(h, 2, 4);
    }
  }
  if (h->multimat && h->sch.n_levelsets > 0) {
    hipEvent_t e0{};
    tic(h, "endstep", &e0);
    mp_end_step(h->mp, h->g, h->s, f5(h->V[0]), h->ID, h->cnt, h->stream);
    toc(h, "endstep", e0);
  }
}

int m2cx_advance_one_time_step(m2cx_solver* h, double time, double dt, int time_step) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  c.dt = dt;
  c.t = time;
  c.time_step = time_step;
  HIPCHK(h, hipMemcpy(h->cnt, &c, sizeof c, hipMemcpyHostToDevice));
  step_kernels(h);
  HIPCHK(h, hipGetLastError());
  return check_error(h);
}

int m2cx_run_steps(m2cx_solver* h, int nsteps) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const bool clamp = h->sch.max_time > 0;
  for (int q = 0; q < nsteps; q++) {
    if (clamp) {
      DevCounters c;
      HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      if (!(c.t < c.tmax - 1.0e-6 * c.dts)) break;
    }
    hipEvent_t e0{};
    tic(h, "dt", &e0);
    k_dt<<<h->dt_blocks, 256, 0, h->stream>>>(h->g, h->s, cf5(h->V[0]), h->ID, h->dt_part, h->cnt, 1);
    toc(h, "dt", e0);
    step_kernels(h);
    HIPCHK(h, hipGetLastError());
  }
  return check_error(h);
}

int m2cx_get_counters(m2cx_solver* h, m2cx_counters* out) {
  DevCounters c;
  HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  out->riemann_calls = c.riemann_calls;
  out->riemann_trivial = c.riemann_trivial;
  out->riemann_errors = c.riemann_errors;
  out->riemann_main_iters = c.riemann_iters;
  out->riemann_fevals = c.riemann_fevals;
  out->riemann_guess_fallbacks = c.guess_fallbacks;
  out->clipped_cells = c.clipped;
  out->id_changed_cells = c.id_changed;
  out->steps = c.steps;
  return 0;
}

int m2cx_reset_counters(m2cx_solver* h) {
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  c.clipped = c.riemann_calls = c.riemann_trivial = c.riemann_errors = c.riemann_iters = c.riemann_fevals = 0;
  c.riemann_fallbacks = c.guess_fallbacks = c.id_changed = c.steps = 0;
  HIPCHK(h, hipMemcpy(h->cnt, &c, sizeof c, hipMemcpyHostToDevice));
  return 0;
}

int m2cx_numerical_flux_batch(m2cx_solver* h, int n, int dir, const double* Vm, const double* Vp,
                              const int32_t* id, double* F) {
  if (!h || n <= 0 || dir < 0 || dir > 2) return M2CX_E_ARG;
  double *dVm, *dVp, *dF;
  int* did;
  HIPCHK(h, hipMalloc(&dVm, n * 40));
  HIPCHK(h, hipMalloc(&dVp, n * 40));
  HIPCHK(h, hipMalloc(&dF, n * 40));
  HIPCHK(h, hipMalloc(&did, n * 4));
  HIPCHK(h, hipMemcpy(dVm, Vm, n * 40, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dVp, Vp, n * 40, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(did, id, n * 4, hipMemcpyHostToDevice));
  k_flux_batch<<<(n + 255) / 256, 256, 0, h->stream>>>(h->s, n, dir, dVm, dVp, did, dF, h->cnt);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(F, dF, n * 40, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  hipFree(dVm); hipFree(dVp); hipFree(dF); hipFree(did);
  return check_error(h);
}

int m2cx_riemann_batch(m2cx_solver* h, int n, const double* dir, const double* Vm, const int32_t* idm,
                       const double* Vp, const int32_t* idp, double* Vs, int32_t* id, double* Vsm, double* Vsp,
                       int32_t* err, int32_t* iters, int32_t* fevals) {
  if (!h || n <= 0) return M2CX_E_ARG;
  return mp_riemann_batch(h->mp, h->s, n, dir, Vm, idm, Vp, idp, Vs, id, Vsm, Vsp, err, iters, fevals, h->cnt,
                          h->stream, &h->err);
}

int m2cx_levelset_residual(m2cx_solver* h, int v_field) {
  if (!h || !h->mp.phi) return M2CX_E_ARG;
  mp_levelset(h->mp, h->g, cf5(h->V[v_field]), -1, h->cnt, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_download_ls_residual(m2cx_solver* h, double* host) {
  if (!h || !h->mp.rls) return M2CX_E_ARG;
  HIPCHK(h, hipMemcpy(host, h->mp.rls, h->ncells * 8, hipMemcpyDeviceToHost));
  return 0;
}

int m2cx_profile(m2cx_solver* h, int enable) {
  h->profile = enable != 0;
  return 0;
}

int m2cx_kernel_time(m2cx_solver* h, const char* name, double* ms, int64_t* launches) {
  HIPCHK(h, hipStreamSynchronize(h->stream));
  flush_timers(h);
  auto it = h->timers.find(name);
  if (it == h->timers.end()) { *ms = 0; *launches = 0; return 0; }
  *ms = it->second.ms;
  *launches = it->second.launches;
  return 0;
}

int m2cx_synchronize(m2cx_solver* h) {
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"
