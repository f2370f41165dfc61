// Host side of the C ABI declared in include/m2c_amd.h: owns the device
// fields of one solver and orders the kernels of the explicit step on its
// stream.  All compute runs in the HIP kernels of m2c_kernels.h /
// m2c_multiphase.h; there is no CPU fallback.
#include <hip/hip_runtime.h>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <string>
#include <vector>
#include "m2c_kernels.h"
#include "m2c_multiphase.h"
#include "m2c_reinit.h"
#include "m2c_band.h"
#include "m2c_mlp.h"
#include <algorithm>
#include <random>

using namespace m2c;

struct Timer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  double ms = 0.0;
  long long launches = 0;
};

// Stage-kernel tile shapes the autotuner chooses from (all 256 threads).
struct TileVariant {
  int bx, by, bz;
};
static const TileVariant kTiles[] = {{16, 4, 4}, {32, 4, 2}, {8, 8, 4}, {16, 8, 2}};

// A level set after the first (m2cx_create_levelsets): the fields, settings
// and reinitializer state the solver's own members hold for level set 0,
// exchanged with them (swap_ls) while that level set's operator runs.
struct LSSlot {
  double *phi = nullptr, *phi1 = nullptr, *phin = nullptr, *rls = nullptr, *phitmp = nullptr;
  LSFvm fv{};
  int lsbc[6] = {1, 1, 1, 1, 1, 1};
  int lsid = 0;
  RBuf rb{};
  NBBuf nb{};
  double reinit_cfl = 0.8;
  m2cx_levelset set{};
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
  int mset = -1;  // slot of M in c_mats
  long ncells = 0;
  // fields
  double* V[3][5] = {};   // V, V1, V2
  double* Un[5] = {};
  double* U1[5] = {};
  double* R[5] = {};
  unsigned char* ID = nullptr;
  unsigned char* IDn = nullptr;
  unsigned char* rflag = nullptr;  // primitive fallback of the reconstruction (Scheme::rflag)
  double* geo = nullptr;  // x, y, z, dx, dy, dz (ghosted), CoeffA, CoeffB per axis
  int* coefK = nullptr;   // CoeffK per axis
  double* fd3 = nullptr;  // level-set FD3 coefficient tables
  double* staging = nullptr;
  int* staging_i = nullptr;
  double* dt_part = nullptr;
  int dt_blocks = 0;
  DevCounters* cnt = nullptr;
  // multi-material
  MPBuffers mp{};
  bool multimat = false;
  // the stage kernel's non-interface tiles run on stream2 while the
  // interface Riemann problems are solved on stream (fork / join events)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  unsigned char* tflag = nullptr;
  long tflag_n = 0;
  bool overlap = false;
  // level-set reinitialization
  RBuf rb{};
  double reinit_cfl = 0.8;  // the reinitializer's CFL, lowered by its failsafe
  NBBuf nb{};               // narrow-band level set (nb.bw == 0: full domain)
  // several level sets: level set 0 in mp / rb / nb / reinit_cfl / the
  // scheme's LevelSet[0] fields, level set q >= 1 in lsx[q - 1]
  int nls = 0;
  std::vector<LSSlot> lsx;
  int ls_bucket = 0;        // TimeIntegratorBase::ls_correction_bucket (TimeIntegrator.h:31)
  int* lsr_list = nullptr;  // candidates of ResolveConflictsInLevelSets
  // number of the step being taken when the host knows it without reading
  // the device (run_steps counts from one read; advance_one_time_step is
  // given it); -1 otherwise
  int host_time_step = -1;
  // Flux = Godunov: face fluxes of the stage, the exact solver's path slots
  double* gflux = nullptr;
  double* gslab = nullptr;
  int glanes = 0;
  RiemannCfg grc{};
  long long reinit_calls = 0, reinit_iters = 0;
  hipStream_t stream_saved = nullptr;  // own stream while an ensemble borrows it
  int tile = -1;                        // stage tile variant (-1: not tuned yet)
  // learned p* guess
  struct {
    bool ready = false;
    float *theta = nullptr, *m = nullptr, *v = nullptr;
    __bf16 *wt1 = nullptr, *wt2 = nullptr, *wt3 = nullptr, *wt4 = nullptr, *w2 = nullptr, *w3 = nullptr;
    float* guess = nullptr;  // per interface slot
    long adam_t = 0;
  } mlp;
  // profiling
  bool profile = false;
  std::map<std::string, Timer> timers;
  std::string err;
};

static int fail(m2cx_solver* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

// a LevelSet[i] block from / into the LevelSet[0] fields of a scheme
static m2cx_levelset ls_from_scheme(const m2cx_scheme& s) {
  m2cx_levelset L{};
  L.materialid = s.ls_materialid;
  for (int q = 0; q < 6; q++) L.bc[q] = s.ls_bc[q];
  L.solver = s.ls_solver;
  L.flux = s.ls_flux;
  L.delta = s.ls_delta;
  L.rec_type = s.ls_rec_type;
  L.limiter = s.ls_limiter;
  L.gmm_coeff = s.ls_gmm_coeff;
  L.reinit_frequency = s.reinit_frequency;
  L.reinit_maxits = s.reinit_maxits;
  L.reinit_frequency_dt = s.reinit_frequency_dt;
  L.reinit_cfl = s.reinit_cfl;
  L.reinit_tol = s.reinit_tol;
  L.reinit_first_layer = s.reinit_first_layer;
  L.bandwidth = s.ls_bandwidth;
  return L;
}

static void scheme_set_ls(m2cx_scheme& s, const m2cx_levelset& L) {
  s.ls_materialid = L.materialid;
  for (int q = 0; q < 6; q++) s.ls_bc[q] = L.bc[q];
  s.ls_solver = L.solver;
  s.ls_flux = L.flux;
  s.ls_delta = L.delta;
  s.ls_rec_type = L.rec_type;
  s.ls_limiter = L.limiter;
  s.ls_gmm_coeff = L.gmm_coeff;
  s.reinit_frequency = L.reinit_frequency;
  s.reinit_maxits = L.reinit_maxits;
  s.reinit_frequency_dt = L.reinit_frequency_dt;
  s.reinit_cfl = L.reinit_cfl;
  s.reinit_tol = L.reinit_tol;
  s.reinit_first_layer = L.reinit_first_layer;
  s.ls_bandwidth = L.bandwidth;
}

// exchanges level set q >= 1 with the solver's level-set members; a second
// call puts it back
static void swap_ls(m2cx_solver* h, int q) {
  if (q <= 0) return;
  LSSlot& S = h->lsx[q - 1];
  MPBuffers& B = h->mp;
  std::swap(B.phi, S.phi);
  std::swap(B.phi1, S.phi1);
  std::swap(B.phin, S.phin);
  std::swap(B.rls, S.rls);
  std::swap(B.phitmp, S.phitmp);
  std::swap(B.fv, S.fv);
  std::swap(B.lsbc, S.lsbc);
  std::swap(B.lsid, S.lsid);
  std::swap(h->rb, S.rb);
  std::swap(h->nb, S.nb);
  std::swap(h->reinit_cfl, S.reinit_cfl);
  const m2cx_levelset mine = ls_from_scheme(h->sch);
  scheme_set_ls(h->sch, S.set);
  S.set = mine;
}

// every level set's fields (for the routines that look at all of them); while
// m2cx_create_levelsets is still allocating the slots, only those that exist
static LSPtrs ls_ptrs(const m2cx_solver* h) {
  LSPtrs P{};
  P.n = std::min(h->nls, 1 + (int)h->lsx.size());
  for (int q = 0; q < P.n; q++) {
    const bool own = q == 0;
    P.phi[q] = own ? h->mp.phi : h->lsx[q - 1].phi;
    P.phin[q] = own ? h->mp.phin : h->lsx[q - 1].phin;
    P.mat[q] = own ? h->mp.lsid : h->lsx[q - 1].lsid;
    P.tracked |= 1u << P.mat[q];
  }
  return P;
}

// the level sets at time n for the Riemann normals of the interface list
static void fill_iface_ls(m2cx_solver* h) {
  const LSPtrs P = ls_ptrs(h);
  IfaceList& L = h->mp.L;
  L.nls = P.n;
  for (int q = 0; q < M2CX_MAX_LEVELSETS; q++) {
    L.lsphi[q] = q < P.n ? P.phi[q] : nullptr;
    L.lsmat[q] = q < P.n ? P.mat[q] : -1;
  }
}

// Slots of c_mats per device: solvers with byte-identical material tables
// share one; a slot is rewritten only when it is free (no live solver uses it).
struct MatSlot {
  DevMats m;
  int refs = 0;
};
static std::mutex g_slot_mu;
static std::map<int, std::vector<MatSlot>> g_slots;

static int mat_slot_acquire(m2cx_solver* h) {
  std::lock_guard<std::mutex> lk(g_slot_mu);
  std::vector<MatSlot>& v = g_slots[h->device];
  int slot = -1;
  for (int i = 0; i < (int)v.size() && slot < 0; i++)
    if (v[i].refs > 0 && std::memcmp(&v[i].m, &h->M, sizeof(DevMats)) == 0) slot = i;
  if (slot < 0)
    for (int i = 0; i < (int)v.size() && slot < 0; i++)
      if (v[i].refs == 0) slot = i;
  if (slot < 0 && (int)v.size() < MAX_MAT_SETS) {
    v.emplace_back();
    slot = (int)v.size() - 1;
  }
  if (slot < 0) return fail(h, M2CX_E_CAPACITY, "too many distinct material tables on one device");
  if (v[slot].refs == 0) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_mats), &h->M, sizeof(DevMats), slot * sizeof(DevMats));
    if (e != hipSuccess) return fail(h, M2CX_E_HIP, std::string("material table upload: ") + hipGetErrorString(e));
    v[slot].m = h->M;
  }
  v[slot].refs++;
  h->mset = slot;
  return 0;
}

static void mat_slot_release(m2cx_solver* h) {
  if (h->mset < 0) return;
  std::lock_guard<std::mutex> lk(g_slot_mu);
  g_slots[h->device][h->mset].refs--;
  h->mset = -1;
}

#define HIPCHK(h, x)                                                                  \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail((h), M2CX_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

// 5-component fields are one allocation each, components n apart; a[p] = a[0] + p*n
static F5 f5(const m2cx_solver* h, double* const* a) { return F5{a[0], h->ncells}; }
static CF5 cf5(const m2cx_solver* h, double* const* a) { return CF5{a[0], h->ncells}; }

static void tic(m2cx_solver* h, const char* name, hipEvent_t* a, hipStream_t st = nullptr) {
  if (!h->profile) return;
  hipEventCreate(a);
  hipEventRecord(*a, st ? st : h->stream);
  (void)name;
}
static void toc(m2cx_solver* h, const char* name, hipEvent_t a, hipStream_t st = nullptr) {
  if (!h->profile) return;
  hipEvent_t b;
  hipEventCreate(&b);
  hipEventRecord(b, st ? st : h->stream);
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
// launch width of the narrow-band list kernels (grid-stride over a count
// that lives on the device)
static int nb_blocks(const m2cx_solver* h) { return (int)std::min<long>(blocks_for(h->ncells), 2048L); }

static int slope_k(int limiter, double A, double B) {
  if (limiter != M2CX_LIM_VANALBADA) return 0;
  int k;
  for (k = 2; k < 1000; k++) {
    double rhs = 2.0 / (1.0 + std::pow((k - 1.0) / k, k - 1) / k) * std::min(A, 1.0);
    if (B <= rhs) break;
  }
  return k;
}

// ---------------------------------------------------------------------------
// Fixed device scratch for m2cx_riemann_batch_noalloc, the one entry point
// that allocates nothing: the integration-path slab of NOALLOC_LANES lanes
// and per-problem staging (learned-guess features, prediction, scales,
// starting pressure; iteration counts when the caller passes none).  One
// copy per GPU, held until the process ends; every call on a GPU shares it,
// so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_LANES = 1024;
constexpr int NOALLOC_PATHCAP = 512;
constexpr long NOALLOC_SLAB_BYTES = (long)NOALLOC_LANES * 2 * NOALLOC_PATHCAP * 3 * 8;
constexpr long NOALLOC_MAXP = M2CX_NOALLOC_MAX_PROBLEMS;
constexpr long NOALLOC_X = NOALLOC_SLAB_BYTES;              // float [max][8]
constexpr long NOALLOC_Y = NOALLOC_X + NOALLOC_MAXP * 32;   // float [max]
constexpr long NOALLOC_P = NOALLOC_Y + NOALLOC_MAXP * 4;    // double [max]
constexpr long NOALLOC_PC = NOALLOC_P + NOALLOC_MAXP * 8;   // double [max]
constexpr long NOALLOC_OK = NOALLOC_PC + NOALLOC_MAXP * 8;  // int [max]
constexpr long NOALLOC_G = NOALLOC_OK + NOALLOC_MAXP * 4;   // float [max]
constexpr long NOALLOC_IT = NOALLOC_G + NOALLOC_MAXP * 4;   // int [max]
constexpr long NOALLOC_FE = NOALLOC_IT + NOALLOC_MAXP * 4;  // int [max]
constexpr long NOALLOC_BYTES = NOALLOC_FE + NOALLOC_MAXP * 4;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, unsigned char*> g_scratch;  // device -> address of `scratch` there

// Looks up, once per device and with that device current, where `scratch`
// lives on it.  Host code uses only this address, never the array's name.
static int scratch_setup(m2cx_solver* h) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch.count(h->device)) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(h, M2CX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[h->device] = static_cast<unsigned char*>(ptr);
  return 0;
}

static unsigned char* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(device);
  return it == g_scratch.end() ? nullptr : it->second;
}

extern "C" {

const char* m2cx_last_error(const m2cx_solver* h) { return h ? h->err.c_str() : "null handle"; }

static int create_impl(const m2cx_grid* grid, const m2cx_material* mats, int nmat, const m2cx_scheme* sch,
                       int device, m2cx_solver** out, bool several);

int m2cx_create(const m2cx_grid* grid, const m2cx_material* mats, int nmat, const m2cx_scheme* sch, int device,
                m2cx_solver** out) {
  return create_impl(grid, mats, nmat, sch, device, out, false);
}

// CoeffK of a level set's own reconstruction on the (uniform) mesh, per axis
static std::vector<int> ls_k_coefs(const m2cx_grid& gr, int limiter) {
  std::vector<int> K;
  const int Ns[3] = {gr.nx, gr.ny, gr.nz};
  const double lo[3] = {gr.x0, gr.y0, gr.z0}, hi[3] = {gr.xmax, gr.ymax, gr.zmax};
  for (int a = 0; a < 3; a++) {
    const double hh = (hi[a] - lo[a]) / Ns[a];
    for (int i = 0; i < Ns[a]; i++) K.push_back(slope_k(limiter, (hh + hh) / (hh + hh), 2.0 * hh / (hh + hh)));
  }
  return K;
}

// Main.cpp:285-330 with each LevelSet[i] block
int m2cx_create_levelsets(const m2cx_grid* grid, const m2cx_material* mats, int nmat, const m2cx_scheme* sch,
                          const m2cx_levelset* ls, int nls, int device, m2cx_solver** out) {
  if (!grid || !mats || !sch || !out || nls < 0 || nls > M2CX_MAX_LEVELSETS || (nls > 0 && !ls)) return M2CX_E_ARG;
  *out = nullptr;
  unsigned seen = 0;
  for (int q = 0; q < nls; q++) {
    const m2cx_levelset& L = ls[q];
    // a level set tracks a material other than 0, one level set per material
    if (L.materialid <= 0 || L.materialid >= nmat || ((seen >> L.materialid) & 1u)) return M2CX_E_ARG;
    seen |= 1u << L.materialid;
    if (L.reinit_first_layer < M2CX_REINIT_FIXED || L.reinit_first_layer > M2CX_REINIT_ITERATIVE_CONSTRAINED2 ||
        L.solver < M2CX_LSSOLVER_FINITE_VOLUME || L.solver > M2CX_LSSOLVER_FINITE_DIFFERENCE ||
        L.flux < M2CX_LSFLUX_ROE || L.flux > M2CX_LSFLUX_UPWIND)
      return M2CX_E_ARG;
    const int bw = (L.bandwidth > 0 && L.bandwidth < INT32_MAX) ? L.bandwidth : 0;
    if (bw && (bw <= 1 || L.reinit_frequency <= 0 || L.reinit_frequency >= bw)) return M2CX_E_ARG;
  }
  m2cx_scheme s0 = *sch;
  s0.n_levelsets = nls;
  if (nls > 0) scheme_set_ls(s0, ls[0]);
  if (int rc = create_impl(grid, mats, nmat, &s0, device, out, true)) return rc;
  m2cx_solver* h = *out;
  const long n = h->ncells;
  const double diag = std::sqrt((grid->xmax - grid->x0) * (grid->xmax - grid->x0) +
                                (grid->ymax - grid->y0) * (grid->ymax - grid->y0) +
                                (grid->zmax - grid->z0) * (grid->zmax - grid->z0));
  for (int q = 1; q < nls; q++) {
    LSSlot S;
    S.set = ls[q];
    S.lsid = ls[q].materialid;
    for (int f = 0; f < 6; f++) S.lsbc[f] = ls[q].bc[f];
    for (double** p : {&S.phi, &S.phi1, &S.phin, &S.rls, &S.phitmp}) {
      HIPCHK(h, hipMalloc(p, n * 8));
      HIPCHK(h, hipMemsetAsync(*p, 0, n * 8, h->stream));
    }
    S.fv.on = ls[q].solver == M2CX_LSSOLVER_FINITE_VOLUME;
    S.fv.flux = ls[q].flux;
    S.fv.rec = ls[q].rec_type;
    S.fv.limiter = ls[q].limiter;
    S.fv.gmm = ls[q].gmm_coeff;
    S.fv.delta = ls[q].delta;
    const std::vector<int> K = ls_k_coefs(*grid, ls[q].limiter);
    int* lsk = nullptr;
    HIPCHK(h, hipMalloc(&lsk, K.size() * 4));
    HIPCHK(h, hipMemcpy(lsk, K.data(), K.size() * 4, hipMemcpyHostToDevice));
    S.fv.lsk = lsk;
    S.reinit_cfl = ls[q].reinit_cfl;
    S.nb.diag = diag;
    S.nb.bw = (ls[q].bandwidth > 0 && ls[q].bandwidth < INT32_MAX) ? ls[q].bandwidth : 0;
    S.nb.fv = ls[q].solver == M2CX_LSSOLVER_FINITE_VOLUME;
    if (S.nb.bw)
      if (int rc = nb_alloc(S.nb, n, h->stream)) return fail(h, rc, "narrow-band buffers");
    h->lsx.push_back(S);
  }
  if (nls >= 2) HIPCHK(h, hipMalloc(&h->lsr_list, (long)h->mp.list_cap * 4));
  fill_iface_ls(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

static int create_impl(const m2cx_grid* grid, const m2cx_material* mats, int nmat, const m2cx_scheme* sch,
                       int device, m2cx_solver** out, bool several) {
  if (!grid || !mats || !sch || !out || nmat < 1 || nmat > M2C_MAX_MAT) return M2CX_E_ARG;
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return M2CX_E_ARG;
  if ((long)grid->nx * grid->ny * grid->nz >= (1L << 31)) return M2CX_E_ARG;  // 32-bit cell numbering
  if (sch->var_type < M2CX_VAR_PRIMITIVE || sch->var_type > M2CX_VAR_CONSERVATIVE_CHARACTERISTIC) return M2CX_E_ARG;
  if (sch->n_levelsets < 0 || sch->n_levelsets > (several ? M2CX_MAX_LEVELSETS : 1)) return M2CX_E_UNSUPPORTED;
  if (sch->n_levelsets > 0 && (sch->ls_materialid <= 0 || sch->ls_materialid >= nmat)) return M2CX_E_ARG;
  // MultiPhase.CorrectionsUntilReinitialization at least 1 (TimeIntegrator.cpp:46-49)
  if (sch->n_levelsets > 0 && sch->mp_corrections_until_reinit < 1) return M2CX_E_ARG;
  if (sch->reinit_first_layer < M2CX_REINIT_FIXED || sch->reinit_first_layer > M2CX_REINIT_ITERATIVE_CONSTRAINED2)
    return M2CX_E_ARG;
  if (sch->riemann_normal < M2CX_RNORMAL_LEVEL_SET || sch->riemann_normal > M2CX_RNORMAL_AVERAGE) return M2CX_E_ARG;
  if (sch->mp_phase_change < M2CX_PHASECHANGE_RIEMANN_SOLUTION || sch->mp_phase_change > M2CX_PHASECHANGE_EXTRAPOLATION)
    return M2CX_E_ARG;
  if (sch->mp_phase_dir != M2CX_PHASEDIR_ALL && sch->mp_phase_dir != M2CX_PHASEDIR_UPWIND) return M2CX_E_ARG;
  if (sch->ls_solver < M2CX_LSSOLVER_FINITE_VOLUME || sch->ls_solver > M2CX_LSSOLVER_FINITE_DIFFERENCE ||
      sch->ls_flux < M2CX_LSFLUX_ROE || sch->ls_flux > M2CX_LSFLUX_UPWIND)
    return M2CX_E_ARG;
  // the level-set normal needs level sets (faces next to a material that
  // no level set tracks, other than 0, take the mesh normal: the reference
  // asserts there, CalculateGradPhiAtCellInterface)
  if (sch->riemann_normal != M2CX_RNORMAL_MESH && sch->n_levelsets < 1) return M2CX_E_UNSUPPORTED;
  // narrow band (LevelSetOperator.cpp:84-99): bandwidth at least 2 and the
  // reinitialization frequency, in steps, below it
  const int bw = (sch->n_levelsets > 0 && sch->ls_bandwidth > 0 && sch->ls_bandwidth < INT32_MAX)
                     ? sch->ls_bandwidth : 0;
  if (bw && (bw <= 1 || sch->reinit_frequency <= 0 || sch->reinit_frequency >= bw)) return M2CX_E_ARG;
  auto* h = new m2cx_solver;
  *out = h;
  h->grid = *grid;
  h->sch = *sch;
  h->mats.assign(mats, mats + nmat);
  h->device = device;
  HIPCHK(h, hipSetDevice(device));
  HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (int rc = scratch_setup(h)) return rc;
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
  // one geometry table: x, y, z, dx, dy, dz, then Reconstructor CoeffA and
  // CoeffB per axis (Reconstructor.cpp:53-99); CoeffK in an int table
  std::vector<double> geo;
  for (int q = 0; q < 6; q++) geo.insert(geo.end(), ax[q].begin(), ax[q].end());
  int Ns[3] = {NX, NY, NZ};
  std::vector<double> Acoef, Bcoef;
  std::vector<int> Kcoef, LSKcoef;
  for (int a = 0; a < 3; a++) {
    const std::vector<double>& d = ax[3 + a];
    for (int i = 0; i < Ns[a]; i++) {
      double A = (d[i] + d[i + 1]) / (d[i + 1] + d[i + 2]);
      double B = 2.0 * d[i + 1] / (d[i + 1] + d[i + 2]);
      Acoef.push_back(A);
      Bcoef.push_back(B);
      Kcoef.push_back(slope_k(sch->limiter, A, B));
      LSKcoef.push_back(slope_k(sch->ls_limiter, A, B));
    }
  }
  geo.insert(geo.end(), Acoef.begin(), Acoef.end());
  geo.insert(geo.end(), Bcoef.begin(), Bcoef.end());
  HIPCHK(h, hipMalloc(&h->geo, geo.size() * 8));
  HIPCHK(h, hipMemcpy(h->geo, geo.data(), geo.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMalloc(&h->coefK, Kcoef.size() * 4));
  HIPCHK(h, hipMemcpy(h->coefK, Kcoef.data(), Kcoef.size() * 4, hipMemcpyHostToDevice));
  g.geo = h->geo;
  g.kt = h->coefK;
  // scheme + materials
  Scheme& s = h->s;
  s.flux = sch->flux; s.rec_type = sch->rec_type; s.limiter = sch->limiter;
  s.slope_zero = (sch->slope_near_interface == M2CX_SLOPE_ZERO);
  s.integrator = sch->integrator; s.mp_flux = sch->mp_flux; s.mp_recon = sch->mp_recon;
  s.nmat = nmat; s.inactive = nmat;
  s.gmm = sch->gmm_coeff; s.del = sch->entropy_fix; s.cfl = sch->cfl;
  s.time_step = sch->time_step;
  s.rnormal = sch->riemann_normal;
  s.phase_change = sch->mp_phase_change;
  s.phase_dir = sch->mp_phase_dir;
  h->reinit_cfl = sch->reinit_cfl;
  s.var_type = sch->var_type;
  DevMats& M = h->M;
  std::memset(&M, 0, sizeof M);
  M.nmat = nmat;
  for (int q = 0; q < nmat; q++) M.m[q] = make_eos(mats[q]);
  m2cx_material dm{};
  dm.eos = M2CX_EOS_DUMMY;
  M.m[nmat] = make_eos(dm);
  if (int rc = mat_slot_acquire(h)) return rc;
  s.mset = h->mset;
  // fields
  const long n = h->ncells;
  auto field = [&](double** comp) -> int {
    double* base = nullptr;
    if (hipMalloc(&base, 5 * n * 8) != hipSuccess) return M2CX_E_OOM;
    if (hipMemsetAsync(base, 0, 5 * n * 8, h->stream) != hipSuccess) return M2CX_E_HIP;
    for (int p = 0; p < 5; p++) comp[p] = base + p * n;
    return 0;
  };
  for (double** comp : {h->V[0], h->V[1], h->V[2], h->Un, h->U1, h->R}) {
    int rc = field(comp);
    if (rc) return fail(h, rc, "device field allocation");
  }
  HIPCHK(h, hipMalloc(&h->ID, n));
  HIPCHK(h, hipMalloc(&h->IDn, n));
  if (sch->var_type != M2CX_VAR_PRIMITIVE && sch->rec_type == M2CX_REC_LINEAR) {
    HIPCHK(h, hipMalloc(&h->rflag, n));
    HIPCHK(h, hipMemsetAsync(h->rflag, 0, n, h->stream));
    s.rflag = h->rflag;
  }
  HIPCHK(h, hipMemsetAsync(h->ID, 0, n, h->stream));
  HIPCHK(h, hipMemsetAsync(h->IDn, 0, n, h->stream));
  if (sch->flux == M2CX_FLUX_GODUNOV) {
    // 3 axes x 5 components per cell; lanes for the exact solves: one per
    // face up to the interface solver's lane count
    HIPCHK(h, hipMalloc(&h->gflux, 15 * n * 8));
    HIPCHK(h, hipMemsetAsync(h->gflux, 0, 15 * n * 8, h->stream));
    const long faces = (long)(NX + 1) * NY * NZ + (long)NX * (NY + 1) * NZ + (long)NX * NY * (NZ + 1);
    h->glanes = (int)std::min<long>(RIEM_LANES, (faces + 63) / 64 * 64);
    HIPCHK(h, hipMalloc(&h->gslab, (long)h->glanes * 2 * GODUNOV_PATHCAP * 3 * 8));
    h->grc = RiemannCfg{sch->riem_maxits_main, sch->riem_maxits_bracket, sch->riem_maxits_shock,
                        sch->riem_steps_rarefaction, sch->riem_tol_main, sch->riem_tol_shock,
                        sch->riem_tol_rarefaction, sch->riem_min_pressure, sch->riem_failure_threshold,
                        sch->riem_pressure_at_failure};
  }
  HIPCHK(h, hipMalloc(&h->staging, n * 5 * 8));
  HIPCHK(h, hipMalloc(&h->staging_i, n * 4));
  h->dt_blocks = std::min(grid_blocks((long)NX * NY * NZ), DT_BLOCKS);
  HIPCHK(h, hipMalloc(&h->dt_part, h->dt_blocks * 8));
  HIPCHK(h, hipMalloc(&h->cnt, sizeof(DevCounters)));
  DevCounters c0{};
  c0.tmax = (sch->max_time > 0) ? sch->max_time : DBL_MAX;
  HIPCHK(h, hipMemcpy(h->cnt, &c0, sizeof c0, hipMemcpyHostToDevice));
  h->multimat = nmat > 1 || sch->n_levelsets > 0;
  h->nls = sch->n_levelsets;
  if (h->multimat) {
    int rc = mp_create(h->mp, g, *sch, ax, h->stream);
    if (rc) return fail(h, rc, "multi-material buffers: " + std::string(hipGetErrorString(hipGetLastError())));
    if (bw) {
      h->nb.bw = bw;
      const m2cx_grid& gr = *grid;
      h->nb.diag = std::sqrt((gr.xmax - gr.x0) * (gr.xmax - gr.x0) + (gr.ymax - gr.y0) * (gr.ymax - gr.y0) +
                             (gr.zmax - gr.z0) * (gr.zmax - gr.z0));
      h->nb.fv = sch->ls_solver == M2CX_LSSOLVER_FINITE_VOLUME;
      if (int rc2 = nb_alloc(h->nb, h->ncells, h->stream)) return fail(h, rc2, "narrow-band buffers");
    }
    LSFvm& fv = h->mp.fv;
    fv.on = sch->ls_solver == M2CX_LSSOLVER_FINITE_VOLUME;
    fv.flux = sch->ls_flux;
    fv.rec = sch->ls_rec_type;
    fv.limiter = sch->ls_limiter;
    fv.gmm = sch->ls_gmm_coeff;
    fv.delta = sch->ls_delta;
    int* lsk = nullptr;
    HIPCHK(h, hipMalloc(&lsk, LSKcoef.size() * 4));
    HIPCHK(h, hipMemcpy(lsk, LSKcoef.data(), LSKcoef.size() * 4, hipMemcpyHostToDevice));
    fv.lsk = lsk;
    // second stream and tile flags for overlapping the stage kernel with the
    // interface Riemann solve (default priority, like the solver's stream)
    if (!getenv("M2C_NO_OVERLAP")) {
      HIPCHK(h, hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
      HIPCHK(h, hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
      HIPCHK(h, hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
      long mt = 0;
      for (const TileVariant& t : kTiles)
        mt = std::max(mt, (long)((NX + t.bx - 1) / t.bx) * ((NY + t.by - 1) / t.by) * ((NZ + t.bz - 1) / t.bz));
      HIPCHK(h, hipMalloc(&h->tflag, mt));
      h->tflag_n = mt;
      h->overlap = true;
    }
    
  }
  if (h->multimat) fill_iface_ls(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_destroy(m2cx_solver* h) {
  if (!h) return 0;
  hipSetDevice(h->device);
  hipStreamSynchronize(h->stream);
  flush_timers(h);
  // each field is one allocation starting at component 0 (pointer swaps
  // exchange whole fields)
  for (int f = 0; f < 3; f++) hipFree(h->V[f][0]);
  hipFree(h->Un[0]); hipFree(h->U1[0]); hipFree(h->R[0]);
  hipFree(h->ID); hipFree(h->IDn); hipFree(h->rflag);
  hipFree(h->gflux); hipFree(h->gslab);
  hipFree(h->geo); hipFree(h->coefK);
  hipFree(h->staging); hipFree(h->staging_i); hipFree(h->dt_part); hipFree(h->cnt);
  if (h->multimat) mp_destroy(h->mp);
  ri_free(h->rb);
  nb_free(h->nb);
  for (LSSlot& S : h->lsx) {
    for (void* p : {(void*)S.phi, (void*)S.phi1, (void*)S.phin, (void*)S.rls, (void*)S.phitmp, (void*)S.fv.lsk})
      if (p) hipFree(p);
    ri_free(S.rb);
    nb_free(S.nb);
  }
  if (h->lsr_list) hipFree(h->lsr_list);
  if (h->tflag) hipFree(h->tflag);
  if (h->stream2) hipStreamSynchronize(h->stream2);
  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  if (h->ev_join) hipEventDestroy(h->ev_join);
  if (h->stream2) hipStreamDestroy(h->stream2);
  mat_slot_release(h);
  for (void* p : {(void*)h->mlp.theta, (void*)h->mlp.m, (void*)h->mlp.v, (void*)h->mlp.wt1, (void*)h->mlp.wt2,
                  (void*)h->mlp.wt3, (void*)h->mlp.wt4, (void*)h->mlp.w2, (void*)h->mlp.w3, (void*)h->mlp.guess})
    hipFree(p);
  hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

static double** field_ptr(m2cx_solver* h, int field) {
  if (field >= M2CX_FIELD_V && field <= M2CX_FIELD_V2) return h->V[field];
  if (field == M2CX_FIELD_R) return h->R;
  return nullptr;
}

int m2cx_upload_levelset(m2cx_solver* h, int ls, const double* host) {
  if (!h || !host) return M2CX_E_ARG;
  if (ls < 0 || ls >= h->nls) return fail(h, M2CX_E_ARG, "no such level set");
  HIPCHK(h, hipSetDevice(h->device));
  double* dst = ls == 0 ? h->mp.phi : h->lsx[ls - 1].phi;
  HIPCHK(h, hipMemcpyAsync(dst, host, h->ncells * 8, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_download_levelset(m2cx_solver* h, int ls, int what, double* host) {
  if (!h || !host || (what != M2CX_LS_PHI && what != M2CX_LS_RESIDUAL)) return M2CX_E_ARG;
  if (ls < 0 || ls >= h->nls) return fail(h, M2CX_E_ARG, "no such level set");
  HIPCHK(h, hipSetDevice(h->device));
  const bool own = ls == 0;
  const double* src = what == M2CX_LS_PHI ? (own ? h->mp.phi : h->lsx[ls - 1].phi)
                                          : (own ? h->mp.rls : h->lsx[ls - 1].rls);
  HIPCHK(h, hipMemcpyAsync(host, src, h->ncells * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
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
    k_aos_to_soa<<<grid_blocks(n), 256, 0, h->stream>>>(h->staging, f5(h, f), 5, n);
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
    k_soa_to_aos<<<grid_blocks(n), 256, 0, h->stream>>>(cf5(h, f), h->staging, 5, n);
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
  if (c.error & 16) return fail(h, M2CX_E_LEVELSETS, "a node belongs to two material subdomains");
  return 0;
}

// the per-solver integration-path slab of the exact solver, on first use
static int ensure_slab(m2cx_solver* h) {
  if (!h->multimat) return 0;
  if (mp_slab(h->mp)) return fail(h, M2CX_E_OOM, "out of device memory (Riemann integration paths)");
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
  a.Vin = cf5(h, h->V[vin]);
  a.Vout = (vout >= 0) ? f5(h, h->V[vout]) : F5{};
  a.Un = f5(h, h->Un);
  a.U1 = f5(h, h->U1);
  a.R = f5(h, h->R);
  a.ID = h->ID;
  if (h->multimat) {
    for (int d = 0; d < 3; d++) a.fidx[d] = h->mp.fidx[d];
    a.L = h->mp.L;
  }
  a.mode = mode;
  a.cnt = h->cnt;
  a.gflux = h->gflux;
  return a;
}

extern "C++" {
static constexpr int kNumTiles = 4;

template <int BX, int BY, int BZ, bool GEN>
static void launch_stage_tile(m2cx_solver* h, const StageArgs& a, int mode, hipStream_t st) {
  // one work-group per tile, padded to a multiple of the XCD count (see
  // k_stage for the tile order)
  const long ntiles = (long)((h->g.NX + BX - 1) / BX) * ((h->g.NY + BY - 1) / BY) * ((h->g.NZ + BZ - 1) / BZ);
  const long per = (ntiles + NXCD - 1) / NXCD;
  const dim3 grid((unsigned)(per * NXCD));
  constexpr int nt = BX * BY * BZ;
  switch (mode) {
    case MODE_RESIDUAL: k_stage<MODE_RESIDUAL, BX, BY, BZ, GEN><<<grid, nt, 0, st>>>(a); break;
    case MODE_STAGE1: k_stage<MODE_STAGE1, BX, BY, BZ, GEN><<<grid, nt, 0, st>>>(a); break;
    case MODE_RK2_FINAL: k_stage<MODE_RK2_FINAL, BX, BY, BZ, GEN><<<grid, nt, 0, st>>>(a); break;
    case MODE_RK3_STAGE2: k_stage<MODE_RK3_STAGE2, BX, BY, BZ, GEN><<<grid, nt, 0, st>>>(a); break;
    case MODE_RK3_FINAL: k_stage<MODE_RK3_FINAL, BX, BY, BZ, GEN><<<grid, nt, 0, st>>>(a); break;
    case MODE_FE: k_stage<MODE_FE, BX, BY, BZ, GEN><<<grid, nt, 0, st>>>(a); break;
  }
}

// Reconstruction in conservative or characteristic variables runs one tile
// shape (16x4x4) of the general kernel; primitive variables choose the tuned
// tile of the primitive-only kernel.
static void launch_stage_variant(m2cx_solver* h, const StageArgs& a, int mode, int v, hipStream_t st = nullptr) {
  if (!st) st = h->stream;
  if (a.s.var_type != M2CX_VAR_PRIMITIVE && a.s.rec_type == M2CX_REC_LINEAR) {
    launch_stage_tile<16, 4, 4, true>(h, a, mode, st);
    return;
  }
  switch (v) {
    case 1: launch_stage_tile<32, 4, 2, false>(h, a, mode, st); break;
    case 2: launch_stage_tile<8, 8, 4, false>(h, a, mode, st); break;
    case 3: launch_stage_tile<16, 8, 2, false>(h, a, mode, st); break;
    default: launch_stage_tile<16, 4, 4, false>(h, a, mode, st); break;
  }
}

// the tile shape the stage kernel runs with
static void stage_tile_dims(const m2cx_solver* h, int& tx, int& ty, int& tz) {
  if (h->s.var_type != M2CX_VAR_PRIMITIVE && h->s.rec_type == M2CX_REC_LINEAR) {
    tx = 16; ty = 4; tz = 4;
    return;
  }
  const TileVariant& t = kTiles[h->tile < 0 ? 0 : h->tile];
  tx = t.bx; ty = t.by; tz = t.bz;
}

// phase 0: all tiles on the solver's stream; 1: tiles without interface
// faces on stream2 (timed as "stage"); 2: the others on the solver's stream
// (timed as "stage_if")
static void launch_stage(m2cx_solver* h, int vin, int vout, int mode, int phase = 0) {
  StageArgs a = stage_args(h, vin, vout, mode);
  a.phase = phase;
  a.tflag = h->tflag;
  hipStream_t st = (phase == 1) ? h->stream2 : h->stream;
  const char* name = (phase == 2) ? "stage_if" : "stage";
  hipEvent_t e0{};
  tic(h, name, &e0, st);
  launch_stage_variant(h, a, mode, h->tile < 0 ? 0 : h->tile, st);
  toc(h, name, e0, st);
}

// Tile-shape autotuner: times the residual launch of every variant on the
// current state (results are identical across variants: each face flux is
// computed from the same inputs whichever thread does it, and each cell
// accumulates its six faces in the same order) and keeps the fastest.  The
// choice is cached per grid shape and scheme for the process.
static std::map<std::tuple<int, int, int, int, int, int>, int>& tile_cache() {
  static std::map<std::tuple<int, int, int, int, int, int>, int> c;
  return c;
}

static int autotune(m2cx_solver* h, float* ms_out) {
  if (h->rflag && !ms_out) {  // one tile shape for the general reconstruction
    h->tile = 0;
    return 0;
  }
  const auto key = std::make_tuple(h->g.NX, h->g.NY, h->g.NZ, h->s.flux, h->s.limiter, (int)h->multimat);
  auto it = tile_cache().find(key);
  if (it != tile_cache().end() && !ms_out) {
    h->tile = it->second;
    return 0;
  }
  const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
  if (nin < 64L * 64 * 64 && !ms_out) {  // small grids: the default shape
    h->tile = 0;
    return 0;
  }
  int err0 = 0;
  HIPCHK(h, hipMemcpy(&err0, &h->cnt->error, sizeof(int), hipMemcpyDeviceToHost));
  StageArgs a = stage_args(h, 0, -1, MODE_RESIDUAL);
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  for (int v = 0; v < kNumTiles; v++) {
    launch_stage_variant(h, a, MODE_RESIDUAL, v);  // warm
    HIPCHK(h, hipEventRecord(e0, h->stream));
    for (int r = 0; r < 3; r++) launch_stage_variant(h, a, MODE_RESIDUAL, v);
    HIPCHK(h, hipEventRecord(e1, h->stream));
    HIPCHK(h, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    if (ms_out) ms_out[v] = ms / 3;
    if (ms < best_ms) { best_ms = ms; best = v; }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  // the residual launches may have flagged the state: restore the flags
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(&h->cnt->error, &err0, sizeof(int), hipMemcpyHostToDevice));
  h->tile = best;
  tile_cache()[key] = best;
  return 0;
}
}  // extern "C++"

// Interface Riemann problems of this stage (multi-material only).
static MLPDev mlp_dev(const m2cx_solver* h) {
  return MLPDev{h->mlp.wt1, h->mlp.wt2, h->mlp.wt3, h->mlp.wt4, h->mlp.w2, h->mlp.w3, h->mlp.theta};
}

// learned guesses for the faces of this stage (when enabled)
static void guess_stage(m2cx_solver* h, int vin) {
  if (!h->multimat) return;
  const bool learned = h->sch.riem_guess == M2CX_GUESS_MLP && h->mlp.ready && h->mlp.guess;
  h->mp.guess = learned ? 1 : 0;
  h->mp.pguess = h->mlp.guess;
  if (learned) {
    hipEvent_t eg{};
    tic(h, "mlp", &eg);
    const int blocks = (int)std::min<long>((h->mp.L.cap + 15) / 16, 4096L);
    k_mlp_guess<<<blocks, 64, 0, h->stream>>>(h->g, h->s, cf5(h, h->V[vin]), h->ID, h->mp.L, mlp_dev(h),
                                              h->mlp.guess, h->cnt);
    toc(h, "mlp", eg);
  }
}

// Cells falling back to primitive reconstruction for this stage's input
// (conservative / characteristic variable types only).
// What every stage needs before its face fluxes: the cells reconstructed in
// primitive variables instead of the scheme's (other variable types), and
// with Flux = Godunov the exact-solver fluxes of the single-material faces.
static void recon_stage(m2cx_solver* h, int vin) {
  const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
  if (h->rflag)
    k_recon_flags<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, h->s, cf5(h, h->V[vin]), h->ID, h->rflag, h->cnt);
  if (h->gflux) {
    hipEvent_t e0{};
    tic(h, "godunov", &e0);
    GodunovArgs ga{h->g, h->s, h->grc, cf5(h, h->V[vin]), h->ID, h->gflux, h->gslab, h->glanes, GODUNOV_PATHCAP,
                   h->cnt};
    k_godunov<<<h->glanes / 64, 64, 0, h->stream>>>(ga);
    toc(h, "godunov", e0);
  }
}

static void interface_stage(m2cx_solver* h, int vin, int stage1) {
  recon_stage(h, vin);
  if (!h->multimat) return;
  fill_iface_ls(h);
  guess_stage(h, vin);
  hipEvent_t e0{};
  tic(h, "riemann", &e0);
  mp_riemann(h->mp, h->g, h->s, cf5(h, h->V[vin]), h->ID, stage1, h->cnt, h->stream);
  toc(h, "riemann", e0);
}

// the level-set residual in the band, then (which >= 0) the RK combination
// on its useful nodes inside the domain and the boundary fill of the field
// it wrote (LevelSetOperator.cpp:880-955, 152-196, 408-508)
static void levelset_band(m2cx_solver* h, CF5 V, int which) {
  MPBuffers& B = h->mp;
  const int LB = nb_blocks(h);
  LSArgs a = mp_ls_args(B, h->g, V, which, h->cnt);
  if (B.fv.on) {
    const long nf = 2L * ((long)h->g.NY * h->g.NZ + (long)h->g.NX * h->g.NZ + (long)h->g.NX * h->g.NY);
    k_nb_fv_recon<<<LB, 256, 0, h->stream>>>(a, h->nb);
    k_nb_fv_ghost<<<blocks_for(nf), 256, 0, h->stream>>>(a, h->nb);
    k_nb_fv_flux<<<LB, 256, 0, h->stream>>>(a, h->nb);
  } else {
    k_nb_ls_residual<<<LB, 256, 0, h->stream>>>(a, h->nb);
  }
  if (which < 0) return;
  k_nb_ls_combine<<<LB, 256, 0, h->stream>>>(h->g, h->nb, which, B.phi, B.phi1, B.rls, h->cnt);
  double* tgt = (which == 1 || which == 3) ? B.phi1 : B.phi;
  const long nf = 2L * ((long)h->g.NY * h->g.NZ + (long)h->g.NX * h->g.NZ + (long)h->g.NX * h->g.NY);
  k_ls_bc<<<blocks_for(nf), 256, 0, h->stream>>>(h->g, tgt, B.lsbc[0], B.lsbc[1], B.lsbc[2], B.lsbc[3], B.lsbc[4],
                                                B.lsbc[5]);
}

static void levelset_stage(m2cx_solver* h, int vin, int which) {
  if (!h->multimat || h->nls == 0) return;
  hipEvent_t e0{};
  tic(h, "levelset", &e0);
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    if (h->nb.bw)
      levelset_band(h, cf5(h, h->V[vin]), which);
    else
      mp_levelset(h->mp, h->g, cf5(h, h->V[vin]), which, h->cnt, h->stream);
    swap_ls(h, q);
  }
  toc(h, "levelset", e0);
}

int m2cx_compute_residual(m2cx_solver* h, int v_field, int capture_riemann) {
  if (!h || v_field < 0 || v_field > 2) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  if (int rc = ensure_slab(h)) return rc;
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
  k_dt<<<h->dt_blocks, 256, 0, h->stream>>>(h->g, h->s, cf5(h, h->V[0]), h->ID, h->dt_part, h->cnt, 0);
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
  k_apply_bc<<<grid_blocks((long)h->g.NX * h->g.NY * h->g.NZ), 256, 0, h->stream>>>(h->g, f5(h, f));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_clip(m2cx_solver* h, int v_field) {
  double** f = field_ptr(h, v_field);
  if (!f) return M2CX_E_ARG;
  k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
  k_clip<<<grid_blocks((long)h->g.NX * h->g.NY * h->g.NZ), 256, 0, h->stream>>>(h->g, h->s.mset, f5(h, f), h->ID, h->cnt);
  int rc = check_error(h);
  if (rc) return rc;
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  return (int)c.stage_clipped;
}

// The RK stages of TimeIntegratorRK2/RK3::AdvanceOneTimeStep
// (TimeIntegrator.cpp:319-455, 527-715): stage input field, whether the
// Riemann star states are captured, stage kernel mode and output field, and
// the level-set combination of the stage.
struct StageSpec {
  int vin, stage1, mode, vout, resync, ls_which;
};
static int stage_specs(int ti, StageSpec* out) {
  if (ti == M2CX_TI_RK2) {
    out[0] = StageSpec{0, 1, MODE_STAGE1, 1, 1, 1};
    out[1] = StageSpec{1, 0, MODE_RK2_FINAL, 0, 0, 2};
    return 2;
  }
  out[0] = StageSpec{0, 1, MODE_STAGE1, 1, 1, 1};
  out[1] = StageSpec{1, 0, MODE_RK3_STAGE2, 2, 1, 3};
  out[2] = StageSpec{2, 0, MODE_RK3_FINAL, 0, 0, 4};
  return 3;
}

// the part of a stage after its interface Riemann problems
static void stage_after_riemann(m2cx_solver* h, const StageSpec& sp) {
  const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
  if (sp.resync) k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
  launch_stage(h, sp.vin, sp.vout, sp.mode);
  if (sp.resync)
    k_resync<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, h->s.mset, cf5(h, h->V[sp.vout]), f5(h, h->U1), h->ID,
                                                       h->cnt);
  levelset_stage(h, sp.vin, sp.ls_which);
}

static bool reinit_enabled(const m2cx_solver* h) {
  return h->sch.n_levelsets > 0 && (h->sch.reinit_frequency > 0 || h->sch.reinit_frequency_dt > 0);
}

// LevelSetReinitializer::ReinitializeFullDomain (LevelSetReinitializer.cpp:
// 76-176) on level set 0.  Iterations go in chunks; between chunks the host
// reads whether they converged (later kernels of a chunk return at once).
constexpr int REINIT_CHUNK = 16;
static int reinit_full(m2cx_solver* h, int special_maxits) {
  MPBuffers& B = h->mp;
  const int flt = h->sch.reinit_first_layer;
  const bool cr = flt == M2CX_REINIT_CONSTRAINED1 || flt == M2CX_REINIT_CONSTRAINED2;
  const bool hcr = flt == M2CX_REINIT_ITERATIVE_CONSTRAINED1 || flt == M2CX_REINIT_ITERATIVE_CONSTRAINED2;
  const int fix = (flt == M2CX_REINIT_FIXED || cr) ? 1 : 0;
  if (int rc = ri_alloc(h->rb, h->ncells, B.list_cap, cr, h->stream)) return fail(h, rc, "reinitialization buffers");
  RBuf& b = h->rb;
  const DevGrid& g = h->g;
  hipStream_t st = h->stream;
  const long nin = (long)g.NX * g.NY * g.NZ;
  const long nf = 2L * ((long)g.NY * g.NZ + (long)g.NX * g.NZ + (long)g.NX * g.NY);
  const int* bc = B.lsbc;
  double* phi = B.phi;
  hipEvent_t e0{};
  tic(h, "reinit", &e0);
  k_rstate_reset<<<1, 1, 0, st>>>(b.st, 1);
  k_rtag<<<blocks_for(g.ncells), 256, 0, st>>>(g, phi, b);
  ReinitState hs;
  HIPCHK(h, hipMemcpyAsync(&hs, b.st, sizeof hs, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  if (hs.nfl == 0) return 0;  // no interface (TagFirstLayerNodes returns false)
  if (hs.nfl > b.cap) return fail(h, M2CX_E_CAPACITY, "first layer of the level set exceeds the list capacity");
  h->reinit_calls++;
  const int lb = blocks_for(hs.nfl);
  k_rsign<<<blocks_for(nin), 256, 0, st>>>(g, phi, b.sign);
  if (cr) {
    k_raxpy_copy<<<blocks_for(g.ncells), 256, 0, st>>>(b.pa, phi, g.ncells);
    k_rfirst<<<lb, 256, 0, st>>>(g, b, phi, b.pa);
    k_rnormal<<<blocks_for(nin), 256, 0, st>>>(g, b, phi);
    k_rnormal_fl<<<lb, 256, 0, st>>>(b);
    k_rcurv<<<lb, 256, 0, st>>>(g, b, phi, b.pa, flt);
    k_rapply<<<lb, 256, 0, st>>>(b, phi);
    k_ls_bc<<<blocks_for(nf), 256, 0, st>>>(g, phi, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5]);
  }
  k_copy_f64<<<blocks_for(g.ncells), 256, 0, st>>>(b.phibk, phi, g.ncells);
  const double tol = h->sch.reinit_tol;
  for (;;) {  // RETRY_FullDomain
    const int maxits = special_maxits > 0 ? special_maxits : h->sch.reinit_maxits;
    const double cfl = h->reinit_cfl;
    k_rstate_reset<<<1, 1, 0, st>>>(b.st, 0);
    bool done = false;
    hs.iters = 0;
    hs.last_dphi = 0.0;
    for (int launched = 0; launched < maxits;) {
      const int chunk = std::min(REINIT_CHUNK, maxits - launched);
      for (int q = 0; q < chunk; q++) {
        k_rstage1<<<blocks_for(nin), 256, 0, st>>>(g, b, phi, fix, cfl);
        k_rbc<<<blocks_for(nf), 256, 0, st>>>(g, b.pa, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], b.st, tol);
        k_rstage2<<<blocks_for(nin), 256, 0, st>>>(g, b, phi, fix, cfl, tol);
        k_rbc<<<blocks_for(nf), 256, 0, st>>>(g, b.pb, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], b.st, tol);
        k_rstage3<<<blocks_for(nin), 256, 0, st>>>(g, b, phi, fix, hcr ? 1 : 0, cfl, tol);
        k_rbc<<<blocks_for(nf), 256, 0, st>>>(g, phi, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], b.st, tol);
        if (hcr) {
          k_rcorr_f<<<lb, 256, 0, st>>>(g, b, phi, flt, tol);
          k_rcorr_apply<<<lb, 256, 0, st>>>(g, b, phi, cfl, tol, 1);
          k_rbc<<<blocks_for(nf), 256, 0, st>>>(g, phi, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], b.st, tol);
        }
        k_rcheck<<<1, 1, 0, st>>>(b.st, tol);
      }
      launched += chunk;
      HIPCHK(h, hipGetLastError());
      HIPCHK(h, hipMemcpyAsync(&hs, b.st, sizeof hs, hipMemcpyDeviceToHost, st));
      HIPCHK(h, hipStreamSynchronize(st));
      if (hs.done) {
        done = true;
        break;
      }
    }
    h->reinit_iters += hs.iters;
    if (done || hs.last_dphi < 0.02) break;  // converged, or close enough (:157-160)
    if (h->reinit_cfl < 0.02) {
      toc(h, "reinit", e0);
      return fail(h, M2CX_E_REINIT, "level-set reinitialization failed to converge");
    }
    h->reinit_cfl /= 1.5;
    k_copy_f64<<<blocks_for(g.ncells), 256, 0, st>>>(phi, b.phibk, g.ncells);
  }
  toc(h, "reinit", e0);
  return 0;
}

// LevelSetOperator::ConstructNarrowBandInReinitializer (LevelSetOperator.cpp:
// 1328-1334) with ConstructNarrowBand (LevelSetReinitializer.cpp:1479-1592)
static int nb_build(m2cx_solver* h) {
  if (int rc = nb_construct(h->nb, h->g, h->mp.phi, h->stream)) return fail(h, rc, "narrow band: list order");
  HIPCHK(h, hipGetLastError());
  return 0;
}

// LevelSetReinitializer::ReinitializeInBand (LevelSetReinitializer.cpp:
// 181-300) on level set 0: the running bounds, the tags of the band, the new
// band around them, then the iterations of the full-domain method restricted
// to the band (chunks of iterations behind the device's done flag).
static int reinit_band(m2cx_solver* h, int special_maxits) {
  MPBuffers& B = h->mp;
  NBBuf& nb = h->nb;
  if (!nb.built) return fail(h, M2CX_E_ARG, "the narrow band of the level set has not been constructed");
  const int flt = h->sch.reinit_first_layer;
  const bool cr = flt == M2CX_REINIT_CONSTRAINED1 || flt == M2CX_REINIT_CONSTRAINED2;
  const bool hcr = flt == M2CX_REINIT_ITERATIVE_CONSTRAINED1 || flt == M2CX_REINIT_ITERATIVE_CONSTRAINED2;
  const int fix = (flt == M2CX_REINIT_FIXED || cr) ? 1 : 0;
  if (int rc = ri_alloc(h->rb, h->ncells, B.list_cap, cr, h->stream)) return fail(h, rc, "reinitialization buffers");
  RBuf& b = h->rb;
  const DevGrid& g = h->g;
  hipStream_t st = h->stream;
  const long nf = 2L * ((long)g.NY * g.NZ + (long)g.NX * g.NZ + (long)g.NX * g.NY);
  const int* bc = B.lsbc;
  double* phi = B.phi;
  const int LB = nb_blocks(h);
  const NBSkip none{nullptr, 0.0, false};
  hipEvent_t e0{};
  tic(h, "reinit", &e0);
  // UpdatePhiMaxAndPhiMinInBand (:1459-1476)
  k_nb_keys<<<1, 1, 0, st>>>(nb, 0);
  k_nb_minmax<<<LB, 256, 0, st>>>(g, phi, nb, 1);
  k_nb_bounds<<<1, 1, 0, st>>>(nb, 0);
  k_rstate_reset<<<1, 1, 0, st>>>(b.st, 1);
  k_nb_rtag<<<LB, 256, 0, st>>>(g, phi, b, nb);
  ReinitState hs;
  HIPCHK(h, hipMemcpyAsync(&hs, b.st, sizeof hs, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  if (hs.nfl == 0) {  // no interface (TagFirstLayerNodesInBand returns false)
    toc(h, "reinit", e0);
    return 0;
  }
  if (hs.nfl > b.cap) return fail(h, M2CX_E_CAPACITY, "first layer of the level set exceeds the list capacity");
  h->reinit_calls++;
  // UpdateNarrowBand (:1777-1886)
  k_nb_reseed<<<blocks_for(g.ncells), 256, 0, st>>>(g, phi, b.tag, nb);
  nb_propagate(nb, g, st);
  if (nb.fv)
    if (int rc = nb_positions(nb, g, true, st)) return fail(h, rc, "narrow band: list order");
  k_nb_leave_enter<<<blocks_for(g.ncells), 256, 0, st>>>(g, phi, nb);
  k_nb_rsign<<<LB, 256, 0, st>>>(g, phi, b.sign, nb);
  const int lb = blocks_for(hs.nfl);
  if (cr) {
    k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 1, b.pa, 0.0, phi, 1.0, nullptr, 0.0, none);
    k_nb_phig2<<<LB, 256, 0, st>>>(nb, b.pa);
    k_rfirst<<<lb, 256, 0, st>>>(g, b, phi, nb.phig2);
    k_nb_rnormal<<<LB, 256, 0, st>>>(g, b, nb, phi);
    k_rnormal_fl<<<lb, 256, 0, st>>>(b);
    k_rcurv<<<lb, 256, 0, st>>>(g, b, phi, nb.phig2, flt);
    k_rapply<<<lb, 256, 0, st>>>(b, phi);
    k_nb_bc<<<blocks_for(nf), 256, 0, st>>>(g, phi, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], nb.useful, none);
  }
  k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 0, b.phibk, 0.0, phi, 1.0, nullptr, 0.0, none);
  const double tol = h->sch.reinit_tol;
  for (;;) {  // RETRY_NarrowBand
    const int maxits = special_maxits > 0 ? special_maxits : h->sch.reinit_maxits;
    const double cfl = h->reinit_cfl;
    k_rstate_reset<<<1, 1, 0, st>>>(b.st, 0);
    const NBSkip s0{b.st, tol, false}, s1{b.st, tol, true};
    bool done = false;
    hs.iters = 0;
    hs.last_dphi = 0.0;
    for (int launched = 0; launched < maxits;) {
      const int chunk = std::min(REINIT_CHUNK, maxits - launched);
      for (int q = 0; q < chunk; q++) {
        k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 0, b.phi0, 0.0, phi, 1.0, nullptr, 0.0, s0);
        k_nb_rres<<<LB, 256, 0, st>>>(g, b, nb, phi, fix, cfl, 1, s0);
        k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 0, b.pa, 0.0, phi, 1.0, nb.r, 1.0, s1);
        k_nb_bc<<<blocks_for(nf), 256, 0, st>>>(g, b.pa, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], nb.useful, s1);
        k_nb_rres<<<LB, 256, 0, st>>>(g, b, nb, b.pa, fix, cfl, 0, s1);
        k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 0, b.pa, 0.25, phi, 0.75, nb.r, 0.25, s1);
        k_nb_bc<<<blocks_for(nf), 256, 0, st>>>(g, b.pa, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], nb.useful, s1);
        k_nb_rres<<<LB, 256, 0, st>>>(g, b, nb, b.pa, fix, cfl, 0, s1);
        k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 0, phi, 1.0 / 3.0, b.pa, 2.0 / 3.0, nb.r, 2.0 / 3.0, s1);
        k_nb_bc<<<blocks_for(nf), 256, 0, st>>>(g, phi, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], nb.useful, s1);
        if (hcr) {
          k_rcorr_f<<<lb, 256, 0, st>>>(g, b, phi, flt, tol);
          k_rcorr_apply<<<lb, 256, 0, st>>>(g, b, phi, cfl, tol, 0);
          k_nb_bc<<<blocks_for(nf), 256, 0, st>>>(g, phi, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], nb.useful, s1);
        }
        k_nb_relerr<<<LB, 256, 0, st>>>(g, b, nb, phi, s1);
        k_rcheck<<<1, 1, 0, st>>>(b.st, tol);
      }
      launched += chunk;
      HIPCHK(h, hipGetLastError());
      HIPCHK(h, hipMemcpyAsync(&hs, b.st, sizeof hs, hipMemcpyDeviceToHost, st));
      HIPCHK(h, hipStreamSynchronize(st));
      if (hs.done) {
        done = true;
        break;
      }
    }
    h->reinit_iters += hs.iters;
    if (done || hs.last_dphi < 0.02) break;  // converged, or close enough (:274-279)
    if (h->reinit_cfl < 0.02) {
      toc(h, "reinit", e0);
      return fail(h, M2CX_E_REINIT, "level-set reinitialization failed to converge");
    }
    h->reinit_cfl /= 1.5;
    k_nb_axpy<<<LB, 256, 0, st>>>(g, nb, 0, phi, 0.0, b.phibk, 1.0, nullptr, 0.0, none);
  }
  toc(h, "reinit", e0);
  return 0;
}

// the reinitialization over the full domain or in the band
static int reinitialize(m2cx_solver* h, int special_maxits) {
  return h->nb.bw ? reinit_band(h, special_maxits) : reinit_full(h, special_maxits);
}

// isTimeToWrite (Utils.cpp:337-357) for the reinitialization frequency
static bool reinit_due(const m2cx_solver* h, double time, double dt, int time_step) {
  if (h->sch.reinit_frequency_dt > 0)
    return std::floor(time / h->sch.reinit_frequency_dt) != std::floor((time - dt) / h->sch.reinit_frequency_dt);
  return h->sch.reinit_frequency > 0 && time_step % h->sch.reinit_frequency == 0;
}

// TimeIntegratorBase::UpdateSolutionAfterTimeStepping (TimeIntegrator.cpp:
// 720-800), one level set, no embedded surfaces.  FixUnresolvedNodes with the
// failsafe density leaves no cell unresolved, so the level-set correction of
// :771-793 never runs.
// MultiPhaseOperator::ResolveConflictsInLevelSets (MultiPhaseOperator.cpp:
// 1705-1990), with its boundary fill (TimeIntegrator.cpp:737-741).  Part I
// alone (the default): the candidates (nodes inside two or more subdomains,
// the only ones it can change) are collected on the device, sorted into the
// reference's (k, j, i) order here, and resolved in that order by one device
// thread.  At the steps LevelSetCorrectionFrequency divides, part I with its
// optional branches and part II run as sweeps of planes i + j + k = t (one
// launch per plane, the sequential order's results), since the optional
// branches can change any node near an interface.
static int resolve_conflicts(m2cx_solver* h, int* resolved) {
  *resolved = 0;
  if (h->nls == 0) return 0;
  hipStream_t st = h->stream;
  bool opt = false;
  if (h->sch.mp_ls_correction_frequency > 0) {
    long ts = h->host_time_step;
    if (ts < 0) {
      DevCounters c;
      HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, st));
      HIPCHK(h, hipStreamSynchronize(st));
      ts = c.time_step;
    }
    opt = ts % h->sch.mp_ls_correction_frequency == 0;
  }
  if (!opt && h->nls < 2) return 0;
  const LSPtrs P = ls_ptrs(h);
  DevCounters c;
  if (opt) {
    const DevGrid& g = h->g;
    k_lsr_reset<<<1, 1, 0, st>>>(h->cnt);
    if (h->nls >= 2) {
      const long m1 = (long)(g.NY + 2) * (g.NZ + 2);
      for (int t = -3; t <= g.NX + g.NY + g.NZ; t++)
        k_lsr_plane1<<<blocks_for(m1), 256, 0, st>>>(g, P, t, h->cnt);
    }
    const long m2 = (long)g.NY * g.NZ;
    for (int t = 0; t <= g.NX + g.NY + g.NZ - 3; t++) k_lsr_plane2<<<blocks_for(m2), 256, 0, st>>>(g, P, t, h->cnt);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
  } else {
    const int cap = h->mp.list_cap;
    k_lsr_reset<<<1, 1, 0, st>>>(h->cnt);
    k_lsr_mark<<<blocks_for(h->ncells), 256, 0, st>>>(h->g, P, h->lsr_list, cap, h->cnt);
    HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    if (c.n_lsr == 0) return 0;
    if (c.n_lsr > cap) return fail(h, M2CX_E_CAPACITY, "overlapping level-set subdomains exceed the list capacity");
    std::vector<int> list(c.n_lsr);
    HIPCHK(h, hipMemcpy(list.data(), h->lsr_list, list.size() * 4, hipMemcpyDeviceToHost));
    std::sort(list.begin(), list.end());  // linear index order is (k, j, i) order
    HIPCHK(h, hipMemcpy(h->lsr_list, list.data(), list.size() * 4, hipMemcpyHostToDevice));
    k_lsr_resolve<<<1, 64, 0, st>>>(h->g, P, h->lsr_list, c.n_lsr, h->cnt);
    HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
  }
  *resolved = c.n_resolved;
  if (*resolved) {
    const long nf = 2L * ((long)h->g.NY * h->g.NZ + (long)h->g.NX * h->g.NZ + (long)h->g.NX * h->g.NY);
    for (int q = 0; q < h->nls; q++) {
      swap_ls(h, q);
      const int* b = h->mp.lsbc;
      k_ls_bc<<<blocks_for(nf), 256, 0, st>>>(h->g, h->mp.phi, b[0], b[1], b[2], b[3], b[4], b[5]);
      swap_ls(h, q);
    }
  }
  return 0;
}

// materials that no level set tracks
static int untracked_count(const m2cx_solver* h) {
  const LSPtrs P = ls_ptrs(h);
  int n = 0;
  for (int m = 0; m < (int)h->mats.size(); m++) n += !((P.tracked >> m) & 1u);
  return n;
}

// UpdateMaterialIDByLevelSet (MultiPhaseOperator.cpp:153-357) over every
// level set, with Phi_tmp in each level set's phin: the sweep, then the
// neighbour vote -- Jacobi sweeps in one work-group when at most one
// material (material 0) has no level set, where the order cannot change the
// result, else the reference's in-order sweeps (the remaining nodes sorted
// here, one device thread) -- and the ghost IDs
static int update_ids(m2cx_solver* h) {
  MPBuffers& B = h->mp;
  const DevGrid& g = h->g;
  hipStream_t st = h->stream;
  const LSPtrs P = ls_ptrs(h);
  if (untracked_count(h) < 2) {
    mp_end_ids(B, g, h->s, P, h->ID, h->cnt, st);
    return 0;
  }
  const long nin = (long)g.NX * g.NY * g.NZ;
  k_end_reset<<<1, 1, 0, st>>>(h->cnt);
  k_id_sweep<<<blocks_for(nin), 256, 0, st>>>(g, h->s, P, h->ID, B.tag, B.remaining, B.list_cap, h->cnt);
  DevCounters c;
  HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  const int n = std::min(c.n_remaining, B.list_cap);
  if (n > 0) {
    std::vector<int> list(n);
    HIPCHK(h, hipMemcpy(list.data(), B.remaining, n * 4L, hipMemcpyDeviceToHost));
    auto key = [&](int q) {  // (i, j, k), i most significant
      long r = q;
      const long i = r % g.SX;
      r /= g.SX;
      const long j = r % g.SY, k = r / g.SY;
      return (i * g.SY + j) * g.SZ + k;
    };
    std::sort(list.begin(), list.end(), [&](int a, int b) { return key(a) < key(b); });
    HIPCHK(h, hipMemcpy(B.remaining, list.data(), n * 4L, hipMemcpyHostToDevice));
    k_id_vote_seq<<<1, 1, 0, st>>>(g, h->s, h->ID, B.tag, B.remaining, n, P.tracked);
  }
  const long nf = 2L * ((long)g.NY * g.NZ + (long)g.NX * g.NZ + (long)g.NX * g.NY);
  k_ghost_ids<<<blocks_for(nf), 256, 0, st>>>(g, h->ID);
  return 0;
}

// TimeIntegratorBase::UpdateSolutionAfterTimeStepping (TimeIntegrator.cpp:
// 720-800), no embedded surfaces: overlaps of several level sets resolved,
// the material IDs from every level set, each level set reinitialized at
// its own frequency (all of them, at once, when the overlaps resolved reach
// CorrectionsUntilReinitialization), then the state update.  FixUnresolvedNodes
// with the failsafe density leaves no cell unresolved, so the level-set
// correction of :771-793 never runs.
static int end_of_step(m2cx_solver* h) {
  if (!(h->multimat && h->nls > 0)) return 0;
  hipEvent_t e0{};
  bool any_reinit = false;
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    any_reinit |= reinit_enabled(h);
    swap_ls(h, q);
  }
  if (h->nls == 1 && !any_reinit && h->sch.mp_ls_correction_frequency <= 0 && untracked_count(h) < 2) {
    tic(h, "endstep", &e0);
    mp_end_step(h->mp, h->g, h->s, f5(h, h->V[0]), h->ID, h->cnt, h->stream);
    toc(h, "endstep", e0);
    return 0;
  }
  MPBuffers& B = h->mp;
  F5 V = f5(h, h->V[0]);
  int resolved = 0;
  if (int rc = resolve_conflicts(h, &resolved)) return rc;
  if (int rc = update_ids(h)) return rc;
  h->ls_bucket += resolved;
  const bool must = h->ls_bucket >= h->sch.mp_corrections_until_reinit;
  if (must) h->ls_bucket = 0;
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    bool due = must;
    if (!must && reinit_enabled(h)) {
      if (h->sch.reinit_frequency_dt <= 0 && h->host_time_step >= 0) {
        // Frequency mode: the step number alone decides, no device read
        due = reinit_due(h, 0.0, 0.0, h->host_time_step);
      } else {
        DevCounters c;
        HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        due = reinit_due(h, c.t, c.dts, c.time_step);
      }
    }
    int rc = 0;
    if (due && !reinit_enabled(h))
      rc = fail(h, M2CX_E_UNSUPPORTED, "reinitialization requested but not specified (reinit_frequency)");
    else if (due)
      rc = reinitialize(h, 0);
    swap_ls(h, q);
    if (rc) return rc;
  }
  mp_end_states(B, h->g, h->s, V, h->ID, h->cnt, h->stream);
  mp_end_fix(B, h->g, h->s, V, h->ID, h->cnt, h->stream);
  return 0;
}

// One stage with the stage kernel's tiles that touch no interface face
// running on stream2 while the interface Riemann problems are solved: they
// read V_in and write V_out (different fields) and no interface flux.  The
// tiles with interface faces follow the Riemann launch; the level set waits
// for both.  Results are those of the single launch (same code per tile).
static void split_stage(m2cx_solver* h, const StageSpec& sp) {
  const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
  fill_iface_ls(h);
  recon_stage(h, sp.vin);
  if (sp.resync) k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
  guess_stage(h, sp.vin);
  hipEventRecord(h->ev_fork, h->stream);
  // the Riemann launch goes first: its waves (one per SIMD) take their
  // places before the stage kernel fills the rest, so a long Riemann launch
  // is not held back behind stage tiles
  hipEvent_t e0{};
  tic(h, "riemann", &e0);
  mp_riemann(h->mp, h->g, h->s, cf5(h, h->V[sp.vin]), h->ID, sp.stage1, h->cnt, h->stream);
  toc(h, "riemann", e0);
  hipStreamWaitEvent(h->stream2, h->ev_fork, 0);
  launch_stage(h, sp.vin, sp.vout, sp.mode, 1);
  hipEventRecord(h->ev_join, h->stream2);
  launch_stage(h, sp.vin, sp.vout, sp.mode, 2);
  hipStreamWaitEvent(h->stream, h->ev_join, 0);
  if (sp.resync)
    k_resync<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, h->s.mset, cf5(h, h->V[sp.vout]), f5(h, h->U1), h->ID,
                                                       h->cnt);
  if (sp.ls_which >= 0) levelset_stage(h, sp.vin, sp.ls_which);
}

// marks the stage-kernel tiles that hold a cell of an interface face (after
// the step's interface list is built)
static void tile_flags(m2cx_solver* h) {
  int tx, ty, tz;
  stage_tile_dims(h, tx, ty, tz);
  const long nt = (long)((h->g.NX + tx - 1) / tx) * ((h->g.NY + ty - 1) / ty) * ((h->g.NZ + tz - 1) / tz);
  k_clear_u8<<<blocks_for(nt), 256, 0, h->stream>>>(h->tflag, nt);
  k_tile_flags<<<blocks_for(h->mp.L.cap), 256, 0, h->stream>>>(h->g, h->mp.L, h->cnt, h->tflag, tx, ty, tz);
}

// One TimeIntegrator step with dt already in cnt->dt (TimeIntegrator.cpp:167-715)
static int step_kernels(m2cx_solver* h) {
  const int ti = h->sch.integrator;
  if (h->multimat) mp_begin_step(h->mp, h->g, h->s, h->ID, h->cnt, h->stream, false);
  // Phi_tmp = Phi: over the whole brick, with the narrow band only on its
  // useful nodes after the first step (TimeIntegrator.cpp:333-339)
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    const bool unbuilt = h->nb.bw && !h->nb.built;
    if (!unbuilt) {
      if (h->nb.bw && h->host_time_step > 1)
        k_nb_copy_useful<<<nb_blocks(h), 256, 0, h->stream>>>(h->nb, h->mp.phin, h->mp.phi);
      else
        k_copy_f64<<<blocks_for(h->g.ncells), 256, 0, h->stream>>>(h->mp.phin, h->mp.phi, h->g.ncells);
    }
    swap_ls(h, q);
    if (unbuilt) return fail(h, M2CX_E_ARG, "the narrow band of the level set has not been constructed");
  }
  if (h->overlap) tile_flags(h);
  if (ti == M2CX_TI_FORWARD_EULER && h->overlap) {
    split_stage(h, StageSpec{0, 1, MODE_FE, 1, 0, -1});
    for (int p = 0; p < 5; p++) std::swap(h->V[0][p], h->V[1][p]);
    levelset_stage(h, 0, 0);
  } else if (ti == M2CX_TI_FORWARD_EULER) {
    interface_stage(h, 0, 1);
    launch_stage(h, 0, 1, MODE_FE);
    // V <- new state (pointer swap), then the level set with the new
    // velocities (TimeIntegrator.cpp:199-214)
    for (int p = 0; p < 5; p++) std::swap(h->V[0][p], h->V[1][p]);
    levelset_stage(h, 0, 0);
  } else {
    StageSpec sp[3];
    const int ns = stage_specs(ti, sp);
    for (int q = 0; q < ns; q++) {
      if (h->overlap) {
        split_stage(h, sp[q]);
      } else {
        interface_stage(h, sp[q].vin, sp[q].stage1);
        stage_after_riemann(h, sp[q]);
      }
    }
  }
  return end_of_step(h);
}

int m2cx_advance_one_time_step(m2cx_solver* h, double time, double dt, int time_step) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  if (int rc = ensure_slab(h)) return rc;
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  c.dt = dt;
  c.dts = dt;
  c.t = time;
  c.time_step = time_step;
  HIPCHK(h, hipMemcpy(h->cnt, &c, sizeof c, hipMemcpyHostToDevice));
  h->host_time_step = time_step;
  const int rc = step_kernels(h);
  h->host_time_step = -1;
  if (rc) return rc;
  HIPCHK(h, hipGetLastError());
  return check_error(h);
}

int m2cx_run_steps(m2cx_solver* h, int nsteps) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  if (int rc = ensure_slab(h)) return rc;
  if (h->tile < 0)
    if (int rc = autotune(h, nullptr)) return rc;
  unsigned long long err0 = 0;
  int ts = 0;
  {
    DevCounters c;
    HIPCHK(h, hipMemcpyAsync(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    err0 = c.riemann_errors;
    ts = c.time_step;
  }
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
    k_dt<<<h->dt_blocks, 256, 0, h->stream>>>(h->g, h->s, cf5(h, h->V[0]), h->ID, h->dt_part, h->cnt, 1);
    toc(h, "dt", e0);
    h->host_time_step = ++ts;  // k_dt advances the device's step count by one
    const int rc = step_kernels(h);
    h->host_time_step = -1;
    if (rc) return rc;
    HIPCHK(h, hipGetLastError());
  }
  if (int rc = check_error(h)) return rc;
  // Riemann failures of this call (warnings, SpaceOperator.cpp:2031-2034)
  DevCounters c;
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  const unsigned long long nf = c.riemann_errors - err0;
  return nf > (unsigned long long)INT_MAX ? INT_MAX : (int)nf;
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
  out->reinit_calls = h->reinit_calls;
  out->reinit_iterations = h->reinit_iters;
  out->unresolved_cells = (int64_t)c.unresolved_total;
  return 0;
}

int m2cx_levelset_reinitialize(m2cx_solver* h, int special_maxits) {
  if (!h) return M2CX_E_ARG;
  if (!h->multimat || h->nls == 0)
    return fail(h, M2CX_E_UNSUPPORTED, "reinitialization requested but not specified (reinit_frequency)");
  HIPCHK(h, hipSetDevice(h->device));
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    int rc = 0;
    if (!reinit_enabled(h))
      rc = fail(h, M2CX_E_UNSUPPORTED, "reinitialization requested but not specified (reinit_frequency)");
    else
      rc = reinitialize(h, special_maxits);
    swap_ls(h, q);
    if (rc) return rc;
  }
  HIPCHK(h, hipGetLastError());
  return check_error(h);
}

int m2cx_get_riemann_histogram(m2cx_solver* h, uint64_t* iters, uint64_t* fevals) {
  if (!h || !iters || !fevals) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  if (!h->multimat) {
    std::fill(iters, iters + M2CX_RIEMANN_HIST_BINS, (uint64_t)0);
    std::fill(fevals, fevals + M2CX_RIEMANN_HIST_BINS, (uint64_t)0);
    return 0;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(iters, h->mp.L.hist, M2CX_RIEMANN_HIST_BINS * 8, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(fevals, h->mp.L.hist + M2CX_RIEMANN_HIST_BINS, M2CX_RIEMANN_HIST_BINS * 8,
                      hipMemcpyDeviceToHost));
  return 0;
}

int m2cx_reset_counters(m2cx_solver* h) {
  DevCounters c;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->multimat) HIPCHK(h, hipMemset(h->mp.L.hist, 0, 2 * M2CX_RIEMANN_HIST_BINS * 8));
  HIPCHK(h, hipMemcpy(&c, h->cnt, sizeof c, hipMemcpyDeviceToHost));
  c.clipped = c.riemann_calls = c.riemann_trivial = c.riemann_errors = c.riemann_iters = c.riemann_fevals = 0;
  c.riemann_fallbacks = c.guess_fallbacks = c.id_changed = c.steps = 0;
  c.unresolved_total = 0;
  h->reinit_calls = h->reinit_iters = 0;
  HIPCHK(h, hipMemcpy(h->cnt, &c, sizeof c, hipMemcpyHostToDevice));
  return 0;
}

// m2cx_riemann_batch on device arrays, with the fixed scratch above instead
// of any allocation, enqueued on the solver's stream without waiting for it.
int m2cx_riemann_batch_noalloc(m2cx_solver* h, int n, const double* dir, const double* Vm, const int32_t* idm,
                               const double* Vp, const int32_t* idp, double* Vs, int32_t* id, double* Vsm,
                               double* Vsp, int32_t* err, int32_t* iters, int32_t* fevals) {
  if (!h) return M2CX_E_ARG;
  if (n <= 0) return fail(h, M2CX_E_ARG, "m2cx_riemann_batch_noalloc: no problems");
  if (n > M2CX_NOALLOC_MAX_PROBLEMS)
    return fail(h, M2CX_E_ARG, "m2cx_riemann_batch_noalloc: " + std::to_string(n) + " problems, at most " +
                                   std::to_string(M2CX_NOALLOC_MAX_PROBLEMS) + " per call");
  if (!dir || !Vm || !idm || !Vp || !idp || !Vs || !id || !Vsm || !Vsp || !err)
    return fail(h, M2CX_E_ARG, "m2cx_riemann_batch_noalloc: null array");
  if (!h->multimat) return fail(h, M2CX_E_ARG, "Riemann batch needs a multi-material solver");
  HIPCHK(h, hipSetDevice(h->device));
  unsigned char* sc = scratch_ptr(h->device);
  if (!sc) return fail(h, M2CX_E_HIP, "m2cx_riemann_batch_noalloc: scratch not set up on this device");
  double* slab = reinterpret_cast<double*>(sc);
  int* it = iters ? iters : reinterpret_cast<int*>(sc + NOALLOC_IT);
  int* fe = fevals ? fevals : reinterpret_cast<int*>(sc + NOALLOC_FE);
  float* dguess = nullptr;
  if (h->sch.riem_guess == M2CX_GUESS_MLP && h->mlp.ready) {
    float* X = reinterpret_cast<float*>(sc + NOALLOC_X);
    float* y = reinterpret_cast<float*>(sc + NOALLOC_Y);
    double* P = reinterpret_cast<double*>(sc + NOALLOC_P);
    double* pc = reinterpret_cast<double*>(sc + NOALLOC_PC);
    int* ok = reinterpret_cast<int*>(sc + NOALLOC_OK);
    dguess = reinterpret_cast<float*>(sc + NOALLOC_G);
    // every entry of X, y, P, pc, ok and the guesses is written before it is read
    k_mlp_features<<<(n + 255) / 256, 256, 0, h->stream>>>(h->s.mset, n, dir, Vm, idm, Vp, idp, nullptr, X, P, pc,
                                                           nullptr, ok);
    k_mlp_predict<<<(n + 15) / 16, 64, 0, h->stream>>>(n, X, mlp_dev(h), y);
    k_mlp_compose<<<(n + 255) / 256, 256, 0, h->stream>>>(n, y, P, pc, ok, dguess);
  }
  const MPBuffers& B = h->mp;
  k_riemann_batch<<<(NOALLOC_LANES + 127) / 128, 128, 0, h->stream>>>(
      h->s.mset, B.rc, n, dir, Vm, idm, Vp, idp, Vs, id, Vsm, Vsp, err, it, fe, slab, NOALLOC_LANES, NOALLOC_PATHCAP,
      dguess, h->cnt);
  HIPCHK(h, hipGetLastError());
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
  if (h->gflux)
    k_godunov_batch<<<h->glanes / 64, 64, 0, h->stream>>>(h->s.mset, h->grc, n, dir, dVm, dVp, did, dF, h->gslab,
                                                          h->glanes, GODUNOV_PATHCAP, h->cnt);
  else
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
  if (!h->multimat) return fail(h, M2CX_E_ARG, "Riemann batch needs a multi-material solver");
  HIPCHK(h, hipSetDevice(h->device));
  if (int rc = ensure_slab(h)) return rc;
  const long N = n;
  double *dd, *dVm, *dVp, *dVs, *dVsm, *dVsp;
  int *didm, *didp, *did, *derr, *dit, *dfe;
  HIPCHK(h, hipMalloc(&dd, N * 24));
  HIPCHK(h, hipMalloc(&dVm, N * 40));
  HIPCHK(h, hipMalloc(&dVp, N * 40));
  HIPCHK(h, hipMalloc(&dVs, N * 40));
  HIPCHK(h, hipMalloc(&dVsm, N * 40));
  HIPCHK(h, hipMalloc(&dVsp, N * 40));
  HIPCHK(h, hipMalloc(&didm, N * 4));
  HIPCHK(h, hipMalloc(&didp, N * 4));
  HIPCHK(h, hipMalloc(&did, N * 4));
  HIPCHK(h, hipMalloc(&derr, N * 4));
  HIPCHK(h, hipMalloc(&dit, N * 4));
  HIPCHK(h, hipMalloc(&dfe, N * 4));
  HIPCHK(h, hipMemcpy(dd, dir, N * 24, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dVm, Vm, N * 40, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dVp, Vp, N * 40, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(didm, idm, N * 4, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(didp, idp, N * 4, hipMemcpyHostToDevice));
  float* dguess = nullptr;
  if (h->sch.riem_guess == M2CX_GUESS_MLP && h->mlp.ready) {
    // learned starting pressures for these problems
    float *X, *y;
    double *P, *pc;
    int* ok;
    HIPCHK(h, hipMalloc(&dguess, N * 4));
    HIPCHK(h, hipMalloc(&X, N * 32));
    HIPCHK(h, hipMalloc(&y, N * 4));
    HIPCHK(h, hipMalloc(&P, N * 8));
    HIPCHK(h, hipMalloc(&pc, N * 8));
    HIPCHK(h, hipMalloc(&ok, N * 4));
    k_mlp_features<<<(n + 255) / 256, 256, 0, h->stream>>>(h->s.mset, n, dd, dVm, didm, dVp, didp, nullptr, X, P, pc, nullptr,
                                                           ok);
    k_mlp_predict<<<(n + 15) / 16, 64, 0, h->stream>>>(n, X, mlp_dev(h), y);
    k_mlp_compose<<<(n + 255) / 256, 256, 0, h->stream>>>(n, y, P, pc, ok, dguess);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    hipFree(X); hipFree(y); hipFree(P); hipFree(pc); hipFree(ok);
  }
  MPBuffers& B = h->mp;
  k_riemann_batch<<<(B.lanes + 127) / 128, 128, 0, h->stream>>>(h->s.mset, B.rc, n, dd, dVm, didm, dVp, didp, dVs, did, dVsm,
                                                                 dVsp, derr, dit, dfe, B.slab, B.lanes, B.pathcap,
                                                                 dguess, h->cnt);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(Vs, dVs, N * 40, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(Vsm, dVsm, N * 40, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(Vsp, dVsp, N * 40, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(id, did, N * 4, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(err, derr, N * 4, hipMemcpyDeviceToHost));
  if (iters) HIPCHK(h, hipMemcpy(iters, dit, N * 4, hipMemcpyDeviceToHost));
  if (fevals) HIPCHK(h, hipMemcpy(fevals, dfe, N * 4, hipMemcpyDeviceToHost));
  for (void* p : {(void*)dd, (void*)dVm, (void*)dVp, (void*)dVs, (void*)dVsm, (void*)dVsp, (void*)didm, (void*)didp,
                  (void*)did, (void*)derr, (void*)dit, (void*)dfe, (void*)dguess})
    hipFree(p);
  return 0;
}

// ---------------------------------------------------------------- learned guess
static int mlp_alloc(m2cx_solver* h) {
  if (h->mlp.theta) return 0;
  auto& m = h->mlp;
  HIPCHK(h, hipMalloc(&m.theta, MLP_NPAR * 4));
  HIPCHK(h, hipMalloc(&m.m, MLP_NPAR * 4));
  HIPCHK(h, hipMalloc(&m.v, MLP_NPAR * 4));
  HIPCHK(h, hipMalloc(&m.wt1, 64 * 32 * 2));
  HIPCHK(h, hipMalloc(&m.wt2, 64 * 64 * 2));
  HIPCHK(h, hipMalloc(&m.wt3, 64 * 64 * 2));
  HIPCHK(h, hipMalloc(&m.wt4, 16 * 64 * 2));
  HIPCHK(h, hipMalloc(&m.w2, 64 * 64 * 2));
  HIPCHK(h, hipMalloc(&m.w3, 64 * 64 * 2));
  if (h->multimat) HIPCHK(h, hipMalloc(&m.guess, (long)h->mp.L.cap * 4));
  return 0;
}

int m2cx_mlp_set_params(m2cx_solver* h, const float* theta) {
  if (!h || !theta) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  int rc = mlp_alloc(h);
  if (rc) return rc;
  HIPCHK(h, hipMemcpy(h->mlp.theta, theta, MLP_NPAR * 4, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(h->mlp.m, 0, MLP_NPAR * 4));
  HIPCHK(h, hipMemset(h->mlp.v, 0, MLP_NPAR * 4));
  h->mlp.adam_t = 0;
  k_mlp_refresh<<<16, 256, 0, h->stream>>>(h->mlp.theta, mlp_dev(h));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->mlp.ready = true;
  return 0;
}

int m2cx_mlp_get_params(m2cx_solver* h, float* theta) {
  if (!h || !theta || !h->mlp.ready) return M2CX_E_ARG;
  HIPCHK(h, hipMemcpy(theta, h->mlp.theta, MLP_NPAR * 4, hipMemcpyDeviceToHost));
  return 0;
}

int m2cx_mlp_features(m2cx_solver* h, int n, const double* dir, const double* Vm, const int32_t* idm,
                      const double* Vp, const int32_t* idp, const double* pstar, float* X, float* T,
                      int32_t* ok) {
  if (!h || n <= 0 || !X || !ok || (pstar && !T)) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const long N = n;
  double *dd, *dVm, *dVp, *dps = nullptr, *P, *pc;
  int *didm, *didp, *dok;
  float *dX, *dT = nullptr;
  HIPCHK(h, hipMalloc(&dd, N * 24));
  HIPCHK(h, hipMalloc(&dVm, N * 40));
  HIPCHK(h, hipMalloc(&dVp, N * 40));
  HIPCHK(h, hipMalloc(&didm, N * 4));
  HIPCHK(h, hipMalloc(&didp, N * 4));
  HIPCHK(h, hipMalloc(&dok, N * 4));
  HIPCHK(h, hipMalloc(&dX, N * 32));
  HIPCHK(h, hipMalloc(&P, N * 8));
  HIPCHK(h, hipMalloc(&pc, N * 8));
  if (pstar) {
    HIPCHK(h, hipMalloc(&dps, N * 8));
    HIPCHK(h, hipMalloc(&dT, N * 4));
    HIPCHK(h, hipMemcpy(dps, pstar, N * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(h, hipMemcpy(dd, dir, N * 24, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dVm, Vm, N * 40, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dVp, Vp, N * 40, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(didm, idm, N * 4, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(didp, idp, N * 4, hipMemcpyHostToDevice));
  k_mlp_features<<<(n + 255) / 256, 256, 0, h->stream>>>(h->s.mset, n, dd, dVm, didm, dVp, didp, dps, dX, P, pc, dT, dok);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(X, dX, N * 32, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(ok, dok, N * 4, hipMemcpyDeviceToHost));
  if (pstar) HIPCHK(h, hipMemcpy(T, dT, N * 4, hipMemcpyDeviceToHost));
  for (void* p : {(void*)dd, (void*)dVm, (void*)dVp, (void*)dps, (void*)P, (void*)pc, (void*)didm, (void*)didp,
                  (void*)dok, (void*)dX, (void*)dT})
    hipFree(p);
  return 0;
}

int m2cx_mlp_predict(m2cx_solver* h, int n, const float* X, float* y) {
  if (!h || n <= 0 || !X || !y || !h->mlp.ready) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  float *dX, *dy;
  HIPCHK(h, hipMalloc(&dX, (long)n * 32));
  HIPCHK(h, hipMalloc(&dy, (long)n * 4));
  HIPCHK(h, hipMemcpy(dX, X, (long)n * 32, hipMemcpyHostToDevice));
  k_mlp_predict<<<(n + 15) / 16, 64, 0, h->stream>>>(n, dX, mlp_dev(h), dy);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(y, dy, (long)n * 4, hipMemcpyDeviceToHost));
  hipFree(dX);
  hipFree(dy);
  return 0;
}

// One pass over perm[start .. start + count): partial gradients, then their
// sum into `gout` (when given) and an Adam step (when lr > 0).
static int mlp_batch(m2cx_solver* h, const float* dX, const float* dT, const int* dperm, int n, int start, int count,
                     float lr, float* gpart, float* lpart, double* loss_acc, float* gout) {
  const int waves = (count + 31) / 32;
  MLPTrainArgs a{mlp_dev(h), dX, dT, dperm, n, start, count, 1.0f / count, gpart, lpart};
  k_mlp_train_step<<<waves, 64, 0, h->stream>>>(a);
  float c1 = 1.0f, c2 = 1.0f;
  const float b1 = 0.9f, b2 = 0.999f;
  if (lr > 0) {
    h->mlp.adam_t++;
    c1 = 1.0f - std::pow(b1, (float)h->mlp.adam_t);
    c2 = 1.0f - std::pow(b2, (float)h->mlp.adam_t);
  }
  k_mlp_adam<<<(MLP_NPAR + 255) / 256, 256, 0, h->stream>>>(h->mlp.theta, h->mlp.m, h->mlp.v, gpart, waves, lr, b1,
                                                             b2, 1e-8f, c1, c2, gout, lpart, loss_acc);
  if (lr > 0) k_mlp_refresh<<<16, 256, 0, h->stream>>>(h->mlp.theta, mlp_dev(h));
  HIPCHK(h, hipGetLastError());
  return 0;
}

int m2cx_mlp_loss_grad(m2cx_solver* h, int n, const float* X, const float* T, float* grad, float* loss) {
  if (!h || n <= 0 || n > 65536 || !X || !T || !h->mlp.ready) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const long N = n;
  const int waves = (n + 31) / 32;
  float *dX, *dT, *gpart, *lpart, *dg;
  int* dperm;
  double* lacc;
  HIPCHK(h, hipMalloc(&dX, N * 32));
  HIPCHK(h, hipMalloc(&dT, N * 4));
  HIPCHK(h, hipMalloc(&dperm, N * 4));
  HIPCHK(h, hipMalloc(&gpart, (long)waves * MLP_NPAR * 4));
  HIPCHK(h, hipMalloc(&lpart, waves * 4));
  HIPCHK(h, hipMalloc(&dg, MLP_NPAR * 4));
  HIPCHK(h, hipMalloc(&lacc, 8));
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  HIPCHK(h, hipMemcpy(dX, X, N * 32, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dT, T, N * 4, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dperm, perm.data(), N * 4, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(lacc, 0, 8));
  int rc = mlp_batch(h, dX, dT, dperm, n, 0, n, 0.0f, gpart, lpart, lacc, dg);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double l = 0;
  HIPCHK(h, hipMemcpy(&l, lacc, 8, hipMemcpyDeviceToHost));
  if (grad) HIPCHK(h, hipMemcpy(grad, dg, MLP_NPAR * 4, hipMemcpyDeviceToHost));
  if (loss) *loss = (float)(l / n);
  for (void* p : {(void*)dX, (void*)dT, (void*)dperm, (void*)gpart, (void*)lpart, (void*)dg, (void*)lacc}) hipFree(p);
  return 0;
}

int m2cx_mlp_train(m2cx_solver* h, int n, const float* X, const float* T, int epochs, int batch, float lr,
                   uint32_t seed, float* loss_per_epoch) {
  if (!h || n <= 0 || !X || !T || epochs <= 0 || batch <= 0 || batch > 65536 || !(lr > 0) || !h->mlp.ready)
    return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  const long N = n;
  const int waves = (batch + 31) / 32;
  float *dX, *dT, *gpart, *lpart;
  int* dperm;
  double* lacc;
  HIPCHK(h, hipMalloc(&dX, N * 32));
  HIPCHK(h, hipMalloc(&dT, N * 4));
  HIPCHK(h, hipMalloc(&dperm, N * 4));
  HIPCHK(h, hipMalloc(&gpart, (long)waves * MLP_NPAR * 4));
  HIPCHK(h, hipMalloc(&lpart, waves * 4));
  HIPCHK(h, hipMalloc(&lacc, 8));
  HIPCHK(h, hipMemcpy(dX, X, N * 32, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(dT, T, N * 4, hipMemcpyHostToDevice));
  std::mt19937 rng(seed);
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    HIPCHK(h, hipMemcpyAsync(dperm, perm.data(), N * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemsetAsync(lacc, 0, 8, h->stream));
    for (int start = 0; start < n; start += batch) {
      const int count = std::min(batch, n - start);
      int rc = mlp_batch(h, dX, dT, dperm, n, start, count, lr, gpart, lpart, lacc, nullptr);
      if (rc) return rc;
    }
    double l = 0;
    HIPCHK(h, hipMemcpyAsync(&l, lacc, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (loss_per_epoch) loss_per_epoch[ep] = (float)(l / n);
  }
  for (void* p : {(void*)dX, (void*)dT, (void*)dperm, (void*)gpart, (void*)lpart, (void*)lacc}) hipFree(p);
  return 0;
}

int m2cx_levelset_residual(m2cx_solver* h, int v_field) {
  if (!h || !h->mp.phi || v_field < 0 || v_field > 2) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    const bool unbuilt = h->nb.bw && !h->nb.built;
    if (!unbuilt && h->nb.bw)
      levelset_band(h, cf5(h, h->V[v_field]), -1);
    else if (!unbuilt)
      mp_levelset(h->mp, h->g, cf5(h, h->V[v_field]), -1, h->cnt, h->stream);
    swap_ls(h, q);
    if (unbuilt) return fail(h, M2CX_E_ARG, "the narrow band of the level set has not been constructed");
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_download_ls_residual(m2cx_solver* h, double* host) {
  if (!h || !h->mp.rls) return M2CX_E_ARG;
  HIPCHK(h, hipMemcpy(host, h->mp.rls, h->ncells * 8, hipMemcpyDeviceToHost));
  return 0;
}

// ---------------------------------------------------------------- ensembles
// Many independent solvers advanced in lockstep on one stream: per member,
// the dt reduction, the stage kernels, the level set and the end-of-step
// update; per RK stage, ONE launch solves the interface Riemann problems of
// all members (k_riemann_multi).  One time step of the whole ensemble is
// captured once as a HIP graph and replayed.
struct m2cx_ensemble {
  std::vector<m2cx_solver*> m;
  std::vector<char> active;  // members still stepping (not yet at MaxTime)
  int device = 0;
  int ti = 0, nstages = 0;
  hipStream_t stream = nullptr;
  RiemMember* dmem[3] = {};  // member tables, one per RK stage
  int nriem = 0;              // members whose interface problems go to the batched launch
  int* doff = nullptr;
  double* slab = nullptr;
  // lanes of the batched Riemann launch, each with an integration-path slot:
  // the members' faces together are many per lane, the launch is bound by
  // the lanes' sequential work as much as by its slowest face, and more waves
  // than SIMDs keep the SIMDs busy (C4, 64 members: 612.8 ms/step at 65 536
  // lanes, 566.2 at 131 072, 521.0 at 262 144, 555.0 at 524 288)
  int lanes = 4 * RIEM_LANES, pathcap = 512;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // members' stage tiles without interface faces run on stream2 while the
  // batched Riemann launch runs on stream
  bool overlap = false;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // the stage kernel of all split members that share a tile shape, one
  // launch per stage and phase (k_stage_multi): member tables [stage][phase]
  // and tile-count prefix sums
  struct StageClass {
    int v = 0;                 // tile variant (kTiles index)
    std::vector<m2cx_solver*> members;
    StageArgs* tab[3][2] = {};
    int* toff = nullptr;
    int ntiles = 0;
  };
  std::vector<StageClass> classes;
  std::set<const m2cx_solver*> classed;  // members whose stage kernel runs in a class launch
  std::string err;
};

static int efail(m2cx_ensemble* e, int code, const std::string& msg) {
  e->err = msg;
  return code;
}

#define EHIPCHK(e, x)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return efail((e), M2CX_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

// members whose interface Riemann problems the batched launch solves: the
// exact solver is not used where MultiPhase.Flux = LocalLaxFriedrichs
// (mp_riemann returns at once for those)
static bool ens_riemann_member(const m2cx_solver* h) {
  return h->multimat && h->sch.mp_flux != M2CX_MPFLUX_LLF;
}

extern "C++" {
// k_stage_multi for one class, stage and phase (1: tiles without interface
// faces, 2: the others) on stream `st`
template <int BX, int BY, int BZ>
static void launch_class_tile(const m2cx_ensemble::StageClass& c, const StageArgs* tab, int mode, hipStream_t st) {
  const long per = ((long)c.ntiles + NXCD - 1) / NXCD;
  const dim3 grid((unsigned)(per * NXCD));
  const int nm = (int)c.members.size();
  constexpr int nt = BX * BY * BZ;
  switch (mode) {
    case MODE_STAGE1: k_stage_multi<MODE_STAGE1, BX, BY, BZ><<<grid, nt, 0, st>>>(tab, c.toff, nm); break;
    case MODE_RK2_FINAL: k_stage_multi<MODE_RK2_FINAL, BX, BY, BZ><<<grid, nt, 0, st>>>(tab, c.toff, nm); break;
    case MODE_RK3_STAGE2: k_stage_multi<MODE_RK3_STAGE2, BX, BY, BZ><<<grid, nt, 0, st>>>(tab, c.toff, nm); break;
    case MODE_RK3_FINAL: k_stage_multi<MODE_RK3_FINAL, BX, BY, BZ><<<grid, nt, 0, st>>>(tab, c.toff, nm); break;
  }
}

static void launch_class(const m2cx_ensemble::StageClass& c, int q, int phase, int mode, hipStream_t st) {
  const StageArgs* tab = c.tab[q][phase - 1];
  switch (c.v) {
    case 1: launch_class_tile<32, 4, 2>(c, tab, mode, st); break;
    case 2: launch_class_tile<8, 8, 4>(c, tab, mode, st); break;
    case 3: launch_class_tile<16, 8, 2>(c, tab, mode, st); break;
    default: launch_class_tile<16, 4, 4>(c, tab, mode, st); break;
  }
}

}  // extern "C++"

static void free_classes(m2cx_ensemble* e) {
  for (auto& c : e->classes) {
    for (auto& row : c.tab)
      for (auto* p : row) hipFree(p);
    hipFree(c.toff);
  }
  e->classes.clear();
}

// one ensemble time step over the active members, enqueued on e->stream (no
// host synchronisation)
static void ensemble_step(m2cx_ensemble* e) {
  std::vector<m2cx_solver*> act;
  for (size_t i = 0; i < e->m.size(); i++)
    if (e->active[i]) act.push_back(e->m[i]);
  for (m2cx_solver* h : act) {
    k_dt<<<h->dt_blocks, 256, 0, h->stream>>>(h->g, h->s, cf5(h, h->V[0]), h->ID, h->dt_part, h->cnt, 1);
    if (h->multimat) mp_begin_step(h->mp, h->g, h->s, h->ID, h->cnt, h->stream);
  }
  StageSpec sp[3];
  stage_specs(e->ti, sp);
  // split: the member's stage launch is divided by tiles as in split_stage
  auto split = [e](const m2cx_solver* h) { return e->overlap && ens_riemann_member(h); };
  if (e->overlap)
    for (m2cx_solver* h : act)
      if (split(h)) tile_flags(h);
  for (int q = 0; q < e->nstages; q++) {
    for (m2cx_solver* h : act) {
      recon_stage(h, sp[q].vin);
      if (split(h) && sp[q].resync) k_reset_stage<<<1, 1, 0, h->stream>>>(h->cnt);
      guess_stage(h, sp[q].vin);
    }
    if (e->overlap) hipEventRecord(e->ev_fork, e->stream);
    if (e->nriem) {
      k_riem_offsets<<<1, 1, 0, e->stream>>>(e->dmem[q], e->nriem, e->doff);
      k_riemann_multi<<<e->lanes / 64, 64, 0, e->stream>>>(e->dmem[q], e->nriem, e->doff, sp[q].stage1, e->slab,
                                                           e->lanes, e->pathcap);
      k_riem_count<<<e->nriem, 256, 0, e->stream>>>(e->dmem[q]);
    }
    if (e->overlap) {
      // the stage tiles of all members: one launch per tile shape, tiles
      // without interface faces beside the Riemann launch, the others after it
      hipStreamWaitEvent(e->stream2, e->ev_fork, 0);
      for (const auto& c : e->classes) launch_class(c, q, 1, sp[q].mode, e->stream2);
      for (m2cx_solver* h : act)
        if (split(h) && !e->classed.count(h)) launch_stage(h, sp[q].vin, sp[q].vout, sp[q].mode, 1);
      hipEventRecord(e->ev_join, e->stream2);
      for (const auto& c : e->classes) launch_class(c, q, 2, sp[q].mode, e->stream);
      for (m2cx_solver* h : act)
        if (split(h) && !e->classed.count(h)) launch_stage(h, sp[q].vin, sp[q].vout, sp[q].mode, 2);
      hipStreamWaitEvent(e->stream, e->ev_join, 0);
    }
    for (m2cx_solver* h : act) {
      if (!split(h)) {
        // members outside the batched launch solve nothing here (single
        // material, or interface LLF fluxes computed in the stage kernel)
        stage_after_riemann(h, sp[q]);
        continue;
      }
      const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
      if (sp[q].resync)
        k_resync<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, h->s.mset, cf5(h, h->V[sp[q].vout]), f5(h, h->U1),
                                                           h->ID, h->cnt);
      if (sp[q].ls_which >= 0) levelset_stage(h, sp[q].vin, sp[q].ls_which);
    }
  }
  for (m2cx_solver* h : act)
    if (h->multimat && h->sch.n_levelsets > 0)
      mp_end_step(h->mp, h->g, h->s, f5(h, h->V[0]), h->ID, h->cnt, h->stream);
}

// The members borrow the ensemble's streams while a step is captured; this
// gives them their own back on every path out of the capture.
struct EnsembleStreamSwap {
  m2cx_ensemble* e;
  std::vector<hipStream_t> s2;
  explicit EnsembleStreamSwap(m2cx_ensemble* e_) : e(e_), s2(e_->m.size()) {
    for (size_t i = 0; i < e->m.size(); i++) {
      e->m[i]->stream = e->stream;
      s2[i] = e->m[i]->stream2;
      e->m[i]->stream2 = e->stream2;
    }
  }
  ~EnsembleStreamSwap() {
    for (size_t i = 0; i < e->m.size(); i++) {
      e->m[i]->stream = e->m[i]->stream_saved;
      e->m[i]->stream2 = s2[i];
    }
  }
};

// (Re)captures one time step of the active members as a HIP graph, with the
// batched Riemann tables over the active members that use the exact solver.
static int ensemble_build(m2cx_ensemble* e) {
  // a replay still in flight reads the member tables rewritten below
  EHIPCHK(e, hipStreamSynchronize(e->stream));
  if (e->exec) { hipGraphExecDestroy(e->exec); e->exec = nullptr; }
  if (e->graph) { hipGraphDestroy(e->graph); e->graph = nullptr; }
  StageSpec sp[3];
  stage_specs(e->ti, sp);
  std::vector<RiemMember> tab[3];
  for (size_t i = 0; i < e->m.size(); i++) {
    m2cx_solver* h = e->m[i];
    if (!e->active[i] || !ens_riemann_member(h)) continue;
    const bool learned = h->sch.riem_guess == M2CX_GUESS_MLP && h->mlp.ready && h->mlp.guess;
    for (int q = 0; q < e->nstages; q++)
      tab[q].push_back(RiemMember{h->g, h->s, h->mp.rc, cf5(h, h->V[sp[q].vin]), h->ID, h->mp.L,
                                  learned ? h->mlp.guess : nullptr, h->cnt});
  }
  const int nriem = (int)tab[0].size();
  if (nriem > e->nriem) return efail(e, M2CX_E_ARG, "ensemble member table grew");
  e->nriem = nriem;
  if (nriem) {
    for (int q = 0; q < e->nstages; q++)
      EHIPCHK(e, hipMemcpy(e->dmem[q], tab[q].data(), tab[q].size() * sizeof(RiemMember), hipMemcpyHostToDevice));
  }
  // stage classes: active members of the batched Riemann launch whose stage
  // kernel is the primitive-variable one, grouped by tile shape
  free_classes(e);
  e->classed.clear();
  if (e->overlap) {
    std::map<int, m2cx_ensemble::StageClass> bytile;
    for (size_t i = 0; i < e->m.size(); i++) {
      m2cx_solver* h = e->m[i];
      if (!e->active[i] || !ens_riemann_member(h)) continue;
      if (h->s.var_type != M2CX_VAR_PRIMITIVE && h->s.rec_type == M2CX_REC_LINEAR) continue;
      auto& c = bytile[h->tile < 0 ? 0 : h->tile];
      c.v = h->tile < 0 ? 0 : h->tile;
      c.members.push_back(h);
    }
    for (auto& kv : bytile) {
      m2cx_ensemble::StageClass c = kv.second;
      const TileVariant& tv = kTiles[c.v];
      std::vector<int> toff(c.members.size() + 1, 0);
      for (size_t k = 0; k < c.members.size(); k++) {
        const DevGrid& g = c.members[k]->g;
        toff[k + 1] = toff[k] + ((g.NX + tv.bx - 1) / tv.bx) * ((g.NY + tv.by - 1) / tv.by) *
                                    ((g.NZ + tv.bz - 1) / tv.bz);
      }
      c.ntiles = toff.back();
      e->classes.push_back(c);
      m2cx_ensemble::StageClass& cc = e->classes.back();
      EHIPCHK(e, hipMalloc(&cc.toff, toff.size() * sizeof(int)));
      EHIPCHK(e, hipMemcpy(cc.toff, toff.data(), toff.size() * sizeof(int), hipMemcpyHostToDevice));
      for (int q = 0; q < e->nstages; q++)
        for (int phase = 1; phase <= 2; phase++) {
          std::vector<StageArgs> tab;
          for (m2cx_solver* h : cc.members) {
            StageArgs a = stage_args(h, sp[q].vin, sp[q].vout, sp[q].mode);
            a.phase = phase;
            a.tflag = h->tflag;
            tab.push_back(a);
          }
          EHIPCHK(e, hipMalloc(&cc.tab[q][phase - 1], tab.size() * sizeof(StageArgs)));
          EHIPCHK(e, hipMemcpy(cc.tab[q][phase - 1], tab.data(), tab.size() * sizeof(StageArgs),
                               hipMemcpyHostToDevice));
        }
      for (m2cx_solver* h : cc.members) e->classed.insert(h);
    }
  }
  {
    EnsembleStreamSwap swap(e);
    EHIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    ensemble_step(e);
    hipError_t ce = hipStreamEndCapture(e->stream, &e->graph);
    if (ce != hipSuccess) return efail(e, M2CX_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
  }
  EHIPCHK(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  return 0;
}

int m2cx_ensemble_create(m2cx_solver* const* members, int n, m2cx_ensemble** out) {
  if (!members || n < 1 || !out) return M2CX_E_ARG;
  auto* e = new m2cx_ensemble;
  *out = e;
  e->m.assign(members, members + n);
  e->active.assign(n, 1);
  e->device = members[0]->device;
  e->ti = members[0]->sch.integrator;
  for (m2cx_solver* h : e->m) {
    if (!h) return efail(e, M2CX_E_ARG, "null member");
    if (h->device != e->device) return efail(e, M2CX_E_ARG, "ensemble members must share one device");
    if (h->sch.integrator != e->ti) return efail(e, M2CX_E_ARG, "ensemble members must share the integrator");
    if (reinit_enabled(h))
      return efail(e, M2CX_E_UNSUPPORTED, "level-set reinitialization needs host control between steps; "
                                          "step such members alone");
    if (h->nls > 1 || (h->nls > 0 && h->sch.mp_ls_correction_frequency > 0) ||
        (h->nls > 0 && untracked_count(h) >= 2))
      return efail(e, M2CX_E_UNSUPPORTED, "several level sets, level-set corrections and two or more materials "
                                          "without a level set need host control between steps; step such "
                                          "members alone");
    if (h->sch.flux == M2CX_FLUX_GODUNOV && h->multimat && h->sch.mp_flux == M2CX_MPFLUX_NUMERICAL)
      return efail(e, M2CX_E_UNSUPPORTED, "Godunov side fluxes at interfaces (Flux = Godunov with "
                                          "MultiPhase.Flux = Numerical): step such members alone");
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++)
      if (e->m[i] == e->m[j]) return efail(e, M2CX_E_ARG, "a solver appears twice in the ensemble");
  // forward Euler exchanges the state buffers every step, which a replayed
  // graph cannot follow
  if (e->ti == M2CX_TI_FORWARD_EULER) return efail(e, M2CX_E_UNSUPPORTED, "ensembles run RK2 or RK3");
  StageSpec sp[3];
  e->nstages = stage_specs(e->ti, sp);
  EHIPCHK(e, hipSetDevice(e->device));
  EHIPCHK(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  for (m2cx_solver* h : e->m) {
    h->profile = false;
    h->stream_saved = h->stream;
    if (int rc = ensure_slab(h)) return efail(e, rc, h->err);
    if (ens_riemann_member(h)) e->nriem++;
  }
  if (const char* env = getenv("M2C_ENSEMBLE_RIEMANN_LANES")) {  // launch-shape override: a multiple of 64
    const long v = atol(env);
    if (v < 64 || v > 8L * RIEM_LANES || v % 64)
      return efail(e, M2CX_E_ARG, "M2C_ENSEMBLE_RIEMANN_LANES: a multiple of 64 up to " +
                                      std::to_string(8L * RIEM_LANES));
    e->lanes = (int)v;
  }
  if (e->nriem) {
    for (int q = 0; q < e->nstages; q++) EHIPCHK(e, hipMalloc(&e->dmem[q], e->nriem * sizeof(RiemMember)));
    EHIPCHK(e, hipMalloc(&e->doff, (e->nriem + 1) * sizeof(int)));
    EHIPCHK(e, hipMalloc(&e->slab, (long)e->lanes * 2 * e->pathcap * 3 * 8));
  }
  for (m2cx_solver* h : e->m)
    if (h->tile < 0 && autotune(h, nullptr)) return efail(e, M2CX_E_HIP, "autotune: " + h->err);
  // overlap when every member that uses the exact solver has it on (the
  // M2C_NO_OVERLAP environment variable or m2cx_set_overlap turns it off)
  e->overlap = e->nriem > 0;
  for (m2cx_solver* h : e->m)
    if (ens_riemann_member(h) && !h->overlap) e->overlap = false;
  if (e->overlap) {
    EHIPCHK(e, hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking));
    EHIPCHK(e, hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    EHIPCHK(e, hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
  }
  return ensemble_build(e);
}

// Main.cpp's loop condition t < MaxTime - 1e-6 dt, as m2cx_run_steps applies it
static bool member_running(const DevCounters& c, double max_time) {
  return !(max_time > 0) || c.t < c.tmax - 1.0e-6 * c.dts;
}

int m2cx_ensemble_run_steps(m2cx_ensemble* e, int nsteps) {
  if (!e || !e->exec || nsteps < 0) return M2CX_E_ARG;
  EHIPCHK(e, hipSetDevice(e->device));
  // members' own streams may still hold work (uploads)
  for (m2cx_solver* h : e->m) EHIPCHK(e, hipStreamSynchronize(h->stream));
  const size_t nm = e->m.size();
  std::vector<DevCounters> c(nm);
  for (size_t i = 0; i < nm; i++) EHIPCHK(e, hipMemcpy(&c[i], e->m[i]->cnt, sizeof(DevCounters), hipMemcpyDeviceToHost));
  unsigned long long err0 = 0;
  for (size_t i = 0; i < nm; i++) err0 += c[i].riemann_errors;
  bool clamp = false;
  for (m2cx_solver* h : e->m) clamp = clamp || h->sch.max_time > 0;
  for (int q = 0; q < nsteps; q++) {
    if (clamp) {
      // each member stops at its MaxTime exactly as m2cx_run_steps stops: a
      // finished member leaves the captured step, which is captured anew.
      // The counters are read once the previous step has finished (the
      // copies below do not wait for the ensemble's non-blocking stream).
      if (q > 0) EHIPCHK(e, hipStreamSynchronize(e->stream));
      bool changed = false, any = false;
      for (size_t i = 0; i < nm; i++) {
        if (!e->active[i]) continue;
        if (q > 0) EHIPCHK(e, hipMemcpy(&c[i], e->m[i]->cnt, sizeof(DevCounters), hipMemcpyDeviceToHost));
        if (!member_running(c[i], e->m[i]->sch.max_time)) {
          e->active[i] = 0;
          changed = true;
        } else {
          any = true;
        }
      }
      if (!any) break;
      if (changed)
        if (int rc = ensemble_build(e)) return rc;
    }
    EHIPCHK(e, hipGraphLaunch(e->exec, e->stream));
  }
  EHIPCHK(e, hipStreamSynchronize(e->stream));
  unsigned long long err1 = 0;
  for (size_t i = 0; i < nm; i++) {
    int rc = check_error(e->m[i]);
    if (rc) return efail(e, rc, "member " + std::to_string(i) + ": " + e->m[i]->err);
    DevCounters ci;
    EHIPCHK(e, hipMemcpy(&ci, e->m[i]->cnt, sizeof ci, hipMemcpyDeviceToHost));
    err1 += ci.riemann_errors;
  }
  // Riemann failures of this call, as m2cx_run_steps reports them
  const unsigned long long nf = err1 - err0;
  return nf > (unsigned long long)INT_MAX ? INT_MAX : (int)nf;
}

const char* m2cx_ensemble_last_error(const m2cx_ensemble* e) { return e ? e->err.c_str() : "null handle"; }

int m2cx_ensemble_destroy(m2cx_ensemble* e) {
  if (!e) return 0;
  hipSetDevice(e->device);
  if (e->stream) hipStreamSynchronize(e->stream);
  if (e->exec) hipGraphExecDestroy(e->exec);
  if (e->graph) hipGraphDestroy(e->graph);
  for (auto* p : e->dmem) hipFree(p);
  hipFree(e->doff);
  hipFree(e->slab);
  free_classes(e);
  if (e->ev_fork) hipEventDestroy(e->ev_fork);
  if (e->ev_join) hipEventDestroy(e->ev_join);
  if (e->stream2) hipStreamDestroy(e->stream2);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return 0;
}

static int nb_build_all(m2cx_solver* h, bool* any);

int m2cx_initialize(m2cx_solver* h, const double* default_state, int default_id, const m2cx_geometry* geoms,
                    int ngeom) {
  if (!h || !default_state || ngeom < 0 || (ngeom > 0 && !geoms)) return M2CX_E_ARG;
  if (default_id < 0 || default_id > h->s.nmat) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  std::vector<InitGeom> G(ngeom);
  for (int m = 0; m < ngeom; m++) {
    const m2cx_geometry& a = geoms[m];
    if ((a.kind != M2CX_GEOM_PLANE && a.kind != M2CX_GEOM_SPHERE) || a.materialid < 0 || a.materialid > h->s.nmat)
      return fail(h, M2CX_E_ARG, "unsupported geometry");
    InitGeom& b = G[m];
    b.kind = a.kind;
    b.interior = a.side == 0;
    b.id = a.materialid;
    for (int d = 0; d < 3; d++) b.x0[d] = a.point[d];
    if (a.kind == M2CX_GEOM_PLANE) {
      // GeoTools::DistanceFromPointToPlane normalises the direction with
      // Vec3D::operator/=, a multiplication by the reciprocal (Vector3D.h:193-200)
      const double nn = std::sqrt(a.normal[0] * a.normal[0] + a.normal[1] * a.normal[1] + a.normal[2] * a.normal[2]);
      if (!(nn > 0)) return fail(h, M2CX_E_ARG, "plane normal is zero");
      const double inv = 1.0 / nn;
      for (int d = 0; d < 3; d++) b.n[d] = a.normal[d] * inv;
    }
    b.R = a.radius;
    for (int p = 0; p < 5; p++) b.V[p] = a.state[p];
  }
  InitGeom* dG = nullptr;
  double* dV0 = nullptr;
  if (ngeom) {
    HIPCHK(h, hipMalloc(&dG, ngeom * sizeof(InitGeom)));
    HIPCHK(h, hipMemcpy(dG, G.data(), ngeom * sizeof(InitGeom), hipMemcpyHostToDevice));
  }
  HIPCHK(h, hipMalloc(&dV0, 40));
  HIPCHK(h, hipMemcpy(dV0, default_state, 40, hipMemcpyHostToDevice));
  const m2cx_grid& gr = h->grid;
  const double diag = std::sqrt((gr.xmax - gr.x0) * (gr.xmax - gr.x0) + (gr.ymax - gr.y0) * (gr.ymax - gr.y0) +
                                (gr.zmax - gr.z0) * (gr.zmax - gr.z0));
  double* phi = (h->multimat && h->sch.n_levelsets > 0) ? h->mp.phi : nullptr;
  const long nin = (long)h->g.NX * h->g.NY * h->g.NZ;
  k_initialize<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, f5(h, h->V[0]), h->ID, phi, dV0, default_id, dG, ngeom,
                                                        h->sch.ls_materialid, diag);
  // boundary fill: state, ghost IDs, level set (ApplyBoundaryConditions,
  // UpdateMaterialIDAtGhostNodes, LevelSetOperator::ApplyBoundaryConditions)
  k_apply_bc<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, f5(h, h->V[0]));
  const long nf = 2L * ((long)h->g.NY * h->g.NZ + (long)h->g.NX * h->g.NZ + (long)h->g.NX * h->g.NY);
  k_ghost_ids<<<grid_blocks(nf), 256, 0, h->stream>>>(h->g, h->ID);
  if (phi) {
    // every level set from the geometries of its material, with its own
    // boundary fill (InitializePhiByDistance per level set,
    // SpaceInitializer.cpp:1314-1360), then the bands (:1787-1793)
    for (int q = 0; q < h->nls; q++) {
      swap_ls(h, q);
      const MPBuffers& B = h->mp;
      if (q > 0)
        k_initialize<<<grid_blocks(nin), 256, 0, h->stream>>>(h->g, f5(h, h->V[0]), h->ID, B.phi, dV0, default_id,
                                                              dG, ngeom, B.lsid, diag, 0);
      k_ls_bc<<<grid_blocks(nf), 256, 0, h->stream>>>(h->g, B.phi, B.lsbc[0], B.lsbc[1], B.lsbc[2], B.lsbc[3],
                                                      B.lsbc[4], B.lsbc[5]);
      swap_ls(h, q);
    }
    bool any = false;
    if (int rc = nb_build_all(h, &any)) return rc;
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  hipFree(dG);
  hipFree(dV0);
  return 0;
}

// the band of every level set that has one
static int nb_build_all(m2cx_solver* h, bool* any) {
  *any = false;
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    int rc = 0;
    if (h->nb.bw) {
      rc = nb_build(h);
      *any = true;
    }
    swap_ls(h, q);
    if (rc) return rc;
  }
  return 0;
}

int m2cx_levelset_construct_narrow_band(m2cx_solver* h) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  bool any = false;
  if (int rc = nb_build_all(h, &any)) return rc;
  if (!any) return fail(h, M2CX_E_UNSUPPORTED, "no narrow band: ls_bandwidth is not set");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_update_material_id_by_levelset(m2cx_solver* h, const double* phi_tmp) {
  if (!h || !phi_tmp) return M2CX_E_ARG;
  if (!h->multimat || h->nls == 0) return fail(h, M2CX_E_UNSUPPORTED, "no level set");
  HIPCHK(h, hipSetDevice(h->device));
  for (int q = 0; q < h->nls; q++) {
    swap_ls(h, q);
    const hipError_t e = hipMemcpyAsync(h->mp.phin, phi_tmp + (long)q * h->ncells, h->ncells * 8,
                                        hipMemcpyHostToDevice, h->stream);
    swap_ls(h, q);
    if (e != hipSuccess) return fail(h, M2CX_E_HIP, std::string("Phi_tmp upload: ") + hipGetErrorString(e));
  }
  if (int rc = update_ids(h)) return rc;
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return check_error(h);
}

int m2cx_levelset_construct_narrow_band_ls(m2cx_solver* h, int ls) {
  if (!h) return M2CX_E_ARG;
  if (ls < 0 || ls >= h->nls) return fail(h, M2CX_E_ARG, "no such level set");
  HIPCHK(h, hipSetDevice(h->device));
  swap_ls(h, ls);
  int rc = h->nb.bw ? nb_build(h) : fail(h, M2CX_E_UNSUPPORTED, "no narrow band: this level set's bandwidth is not set");
  swap_ls(h, ls);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int m2cx_levelset_reinitialize_ls(m2cx_solver* h, int ls, int special_maxits) {
  if (!h) return M2CX_E_ARG;
  if (ls < 0 || ls >= h->nls) return fail(h, M2CX_E_ARG, "no such level set");
  HIPCHK(h, hipSetDevice(h->device));
  swap_ls(h, ls);
  int rc = reinit_enabled(h)
               ? reinitialize(h, special_maxits)
               : fail(h, M2CX_E_UNSUPPORTED, "reinitialization requested but not specified (reinit_frequency)");
  swap_ls(h, ls);
  if (rc) return rc;
  HIPCHK(h, hipGetLastError());
  return check_error(h);
}

int m2cx_get_narrow_band(m2cx_solver* h, int ls, int32_t* level, uint8_t* useful, uint8_t* active,
                         double* bounds) {
  if (!h) return M2CX_E_ARG;
  if (ls < 0 || ls >= h->nls) return fail(h, M2CX_E_ARG, "no such level set");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const NBBuf& nb = ls == 0 ? h->nb : h->lsx[ls - 1].nb;
  if (!nb.bw || !nb.built) return fail(h, M2CX_E_UNSUPPORTED, "no narrow band constructed");
  const long n = h->ncells;
  if (level) HIPCHK(h, hipMemcpy(level, nb.level, n * 4, hipMemcpyDeviceToHost));
  if (useful) HIPCHK(h, hipMemcpy(useful, nb.useful, n, hipMemcpyDeviceToHost));
  if (active) HIPCHK(h, hipMemcpy(active, nb.active, n, hipMemcpyDeviceToHost));
  if (bounds) {
    NBState st;
    HIPCHK(h, hipMemcpy(&st, nb.st, sizeof st, hipMemcpyDeviceToHost));
    bounds[0] = st.phi_max;
    bounds[1] = st.phi_min;
    bounds[2] = st.phi_out_pos;
    bounds[3] = st.phi_out_neg;
  }
  return 0;
}

int m2cx_autotune(m2cx_solver* h, int* variant, float* ms_per_variant) {
  if (!h) return M2CX_E_ARG;
  HIPCHK(h, hipSetDevice(h->device));
  float ms[kNumTiles];
  int rc = autotune(h, ms);
  if (rc) return rc;
  if (variant) *variant = h->tile;
  if (ms_per_variant)
    for (int v = 0; v < kNumTiles; v++) ms_per_variant[v] = ms[v];
  return 0;
}

int m2cx_set_tile(m2cx_solver* h, int variant) {
  if (!h || variant < 0 || variant >= kNumTiles) return M2CX_E_ARG;
  h->tile = variant;
  return 0;
}

int m2cx_set_overlap(m2cx_solver* h, int on) {
  if (!h) return M2CX_E_ARG;
  h->overlap = on != 0 && h->stream2 != nullptr;
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
