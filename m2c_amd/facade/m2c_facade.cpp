// C++ façades with the reference's method signatures over the C ABI of
// libm2c_amd.so (declared in include/m2c_amd_facade.hpp).  Host code only:
// every computation is a call into the C ABI, which runs the HIP kernels.
#include "m2c_amd_facade.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace m2c_amd {

// ----------------------------------------------------------------- Context

Context::Context(const m2cx_grid& grid, const m2cx_material* mats, int nmat, const m2cx_scheme& scheme, int device)
    : grid_(grid), scheme_(scheme) {
  const int rc = m2cx_create(&grid, mats, nmat, &scheme, device, &h_);
  if (rc < 0) {
    const std::string msg = std::string("m2cx_create: ") + m2cx_last_error(h_);
    if (h_) m2cx_destroy(h_);
    h_ = nullptr;
    throw Error(rc, msg);
  }
  if (scheme.n_levelsets > 0) {
    // the scheme's own LevelSet[0] block
    m2cx_levelset L{};
    L.materialid = scheme.ls_materialid;
    for (int q = 0; q < 6; q++) L.bc[q] = scheme.ls_bc[q];
    L.solver = scheme.ls_solver;
    L.flux = scheme.ls_flux;
    L.delta = scheme.ls_delta;
    L.rec_type = scheme.ls_rec_type;
    L.limiter = scheme.ls_limiter;
    L.gmm_coeff = scheme.ls_gmm_coeff;
    L.reinit_frequency = scheme.reinit_frequency;
    L.reinit_maxits = scheme.reinit_maxits;
    L.reinit_frequency_dt = scheme.reinit_frequency_dt;
    L.reinit_cfl = scheme.reinit_cfl;
    L.reinit_tol = scheme.reinit_tol;
    L.reinit_first_layer = scheme.reinit_first_layer;
    L.bandwidth = scheme.ls_bandwidth;
    levelsets_.push_back(L);
  }
}

Context::Context(const m2cx_grid& grid, const m2cx_material* mats, int nmat, const m2cx_scheme& scheme,
                 const std::vector<m2cx_levelset>& levelsets, int device)
    : grid_(grid), scheme_(scheme), levelsets_(levelsets) {
  const int rc = m2cx_create_levelsets(&grid, mats, nmat, &scheme, levelsets.data(), (int)levelsets.size(), device,
                                       &h_);
  if (rc < 0) {
    const std::string msg = std::string("m2cx_create_levelsets: ") + m2cx_last_error(h_);
    if (h_) m2cx_destroy(h_);
    h_ = nullptr;
    throw Error(rc, msg);
  }
  scheme_.n_levelsets = (int)levelsets.size();
  if (!levelsets.empty()) scheme_.ls_materialid = levelsets[0].materialid;
}

Context::~Context() {
  // variables mirrored on the device keep their latest values on the host
  for (int f = 0; f < kFields; f++) {
    if (!holder[f]) continue;
    try {
      holder[f]->SyncToHost();
    } catch (const Error&) {
    }
    holder[f]->Unbind();
  }
  if (h_) m2cx_destroy(h_);
}

int Context::check(int rc) const {
  if (rc < 0) throw Error(rc, std::string("m2c_amd: ") + m2cx_last_error(h_));
  return rc;
}

// --------------------------------------------------------- SpaceVariable3D

SpaceVariable3D::SpaceVariable3D(const m2cx_grid& grid, int dof)
    : NX_(grid.nx), NY_(grid.ny), NZ_(grid.nz), dof_(dof) {
  if (NX_ < 1 || NY_ < 1 || NZ_ < 1 || dof_ < 1) throw Error(M2CX_E_ARG, "SpaceVariable3D: empty brick");
  const long sx = NX_ + 2, sy = NY_ + 2, sz = NZ_ + 2;
  data_.assign(sx * sy * sz * dof_, 0.0);
  // array[k][j][i*dof + p] for k, j, i from -1 to N (DMDAVecGetArray's global
  // indexing of a ghosted brick whose corner is -1)
  rows_.resize(sy * sz);
  planes_.resize(sz);
  for (long k = 0; k < sz; k++) {
    for (long j = 0; j < sy; j++) rows_[k * sy + j] = &data_[((k * sy + j) * sx + 1) * dof_];
    planes_[k] = &rows_[k * sy + 1];
  }
  array_ = &planes_[1];
}

SpaceVariable3D::~SpaceVariable3D() { Unbind(); }

void SpaceVariable3D::Destroy() { Unbind(); }

void SpaceVariable3D::Unbind() {
  if (ctx_ && field_ >= 0 && ctx_->holder[field_] == this) ctx_->holder[field_] = nullptr;
  ctx_ = nullptr;
  field_ = -1;
  device_newer_ = false;
  host_newer_ = true;
}

void SpaceVariable3D::SyncToHost() {
  if (!device_newer_ || !ctx_) return;
  if (field_ == M2CX_FIELD_ID) {
    std::vector<int32_t> tmp(data_.size());
    ctx_->check(m2cx_download(ctx_->handle(), field_, tmp.data()));
    for (size_t q = 0; q < tmp.size(); q++) data_[q] = (double)tmp[q];
  } else if (field_ >= 6) {
    ctx_->check(m2cx_download_levelset(ctx_->handle(), field_ - 5, M2CX_LS_PHI, data_.data()));
  } else {
    ctx_->check(m2cx_download(ctx_->handle(), field_, data_.data()));
  }
  device_newer_ = false;
}

void SpaceVariable3D::Evict(Context& ctx, int field) {
  SpaceVariable3D* other = ctx.holder[field];
  if (!other) return;
  other->SyncToHost();
  other->Unbind();
}

void SpaceVariable3D::ToDevice(Context& ctx, int field) {
  if (ctx_ == &ctx && field_ == field && !host_newer_) return;  // the device copy is current
  const int want = (field == M2CX_FIELD_V || field == M2CX_FIELD_V1 || field == M2CX_FIELD_V2 ||
                    field == M2CX_FIELD_R) ? 5 : 1;
  if (dof_ != want) throw Error(M2CX_E_ARG, "SpaceVariable3D: wrong number of components for this field");
  SyncToHost();
  if (ctx.holder[field] != this) Evict(ctx, field);
  Unbind();
  if (field == M2CX_FIELD_ID) {
    std::vector<int32_t> tmp(data_.size());
    for (size_t q = 0; q < tmp.size(); q++) tmp[q] = (int32_t)data_[q];  // IDs are integer-valued doubles
    ctx.check(m2cx_upload(ctx.handle(), field, tmp.data()));
  } else if (field >= 6) {
    ctx.check(m2cx_upload_levelset(ctx.handle(), field - 5, data_.data()));
  } else {
    ctx.check(m2cx_upload(ctx.handle(), field, data_.data()));
  }
  ctx.holder[field] = this;
  ctx_ = &ctx;
  field_ = field;
  host_newer_ = false;
  device_newer_ = false;
}

void SpaceVariable3D::ProducedOnDevice(Context& ctx, int field) {
  if (ctx.holder[field] != this) Evict(ctx, field);
  if (!(ctx_ == &ctx && field_ == field)) Unbind();
  ctx.holder[field] = this;
  ctx_ = &ctx;
  field_ = field;
  host_newer_ = false;
  device_newer_ = true;
}

double*** SpaceVariable3D::GetDataPointer() {
  SyncToHost();
  return array_;
}

// Both restores may follow writes through the pointer: the device copy is
// then out of date and the next operator uploads the host copy.
void SpaceVariable3D::RestoreDataPointerAndInsert() { host_newer_ = true; }
void SpaceVariable3D::RestoreDataPointerToLocalVector() { host_newer_ = true; }

void SpaceVariable3D::GetCornerIndices(int* i0, int* j0, int* k0, int* imax, int* jmax, int* kmax) const {
  *i0 = 0; *j0 = 0; *k0 = 0;
  if (imax) *imax = NX_;
  if (jmax) *jmax = NY_;
  if (kmax) *kmax = NZ_;
}

void SpaceVariable3D::GetGhostedCornerIndices(int* i0, int* j0, int* k0, int* imax, int* jmax, int* kmax) const {
  *i0 = -1; *j0 = -1; *k0 = -1;
  if (imax) *imax = NX_ + 1;
  if (jmax) *jmax = NY_ + 1;
  if (kmax) *kmax = NZ_ + 1;
}

void SpaceVariable3D::GetSize(int* nx, int* ny, int* nz) const { *nx = NX_; *ny = NY_; *nz = NZ_; }

void SpaceVariable3D::GetGhostedSize(int* nx, int* ny, int* nz) const { *nx = NX_ + 2; *ny = NY_ + 2; *nz = NZ_ + 2; }

void SpaceVariable3D::SetConstantValue(double a, bool workOnGhost) {
  SyncToHost();
  for (int k = -1; k <= NZ_; k++)
    for (int j = -1; j <= NY_; j++)
      for (int i = -1; i <= NX_; i++) {
        const bool ghost = i < 0 || j < 0 || k < 0 || i >= NX_ || j >= NY_ || k >= NZ_;
        if (ghost && !workOnGhost) continue;
        for (int p = 0; p < dof_; p++) array_[k][j][i * dof_ + p] = a;
      }
  host_newer_ = true;
}

void SpaceVariable3D::AXPlusBY(double a, double b, SpaceVariable3D& y, bool workOnGhost) {
  if (y.NX_ != NX_ || y.NY_ != NY_ || y.NZ_ != NZ_ || y.dof_ != dof_)
    throw Error(M2CX_E_ARG, "AXPlusBY: variables of different shape");
  SyncToHost();
  y.SyncToHost();
  for (int k = -1; k <= NZ_; k++)
    for (int j = -1; j <= NY_; j++)
      for (int i = -1; i <= NX_; i++) {
        const bool ghost = i < 0 || j < 0 || k < 0 || i >= NX_ || j >= NY_ || k >= NZ_;
        if (ghost && !workOnGhost) continue;
        for (int p = 0; p < dof_; p++) {
          double& x = array_[k][j][i * dof_ + p];
          x = a * x + b * y.array_[k][j][i * dof_ + p];
        }
      }
  host_newer_ = true;
}

// ----------------------------------------------------------- SpaceOperator

static void riemann_warning(int nfail) {
  if (nfail > 0)
    std::fprintf(stdout, "Warning: Riemann solver failed to find a bracketing interval or to converge on %d "
                         "edge(s).\n", nfail);
}

void SpaceOperator::ApplyBoundaryConditions(SpaceVariable3D& V) {
  Context& c = *ctx_;
  V.ToDevice(c, M2CX_FIELD_V);
  c.check(m2cx_apply_boundary_conditions(c.handle(), M2CX_FIELD_V));
  V.ProducedOnDevice(c, M2CX_FIELD_V);
}

int SpaceOperator::ClipDensityAndPressure(SpaceVariable3D& V, SpaceVariable3D& ID, bool workOnGhost,
                                          bool checkState) {
  (void)checkState;  // the device clip always checks the clipped states
  if (workOnGhost) throw Error(M2CX_E_UNSUPPORTED, "ClipDensityAndPressure: ghost cells are not clipped here");
  Context& c = *ctx_;
  V.ToDevice(c, M2CX_FIELD_V);
  ID.ToDevice(c, M2CX_FIELD_ID);
  const int n = c.check(m2cx_clip(c.handle(), M2CX_FIELD_V));
  V.ProducedOnDevice(c, M2CX_FIELD_V);
  return n;
}

void SpaceOperator::ComputeTimeStepSize(SpaceVariable3D& V, SpaceVariable3D& ID, double& dt, double& cfl,
                                        SpaceVariable3D* LocalDt) {
  if (LocalDt) throw Error(M2CX_E_UNSUPPORTED, "ComputeTimeStepSize: local time stepping is not on this path");
  Context& c = *ctx_;
  V.ToDevice(c, M2CX_FIELD_V);
  ID.ToDevice(c, M2CX_FIELD_ID);
  c.check(m2cx_compute_time_step(c.handle(), &dt, &cfl));
}

void SpaceOperator::ComputeResidual(SpaceVariable3D& V, SpaceVariable3D& ID, SpaceVariable3D& R, double time,
                                    RiemannSolutions* riemann_solutions, std::vector<int>* ls_mat_id,
                                    std::vector<SpaceVariable3D*>* Phi, std::vector<SpaceVariable3D*>* KappaPhi,
                                    std::vector<std::unique_ptr<EmbeddedBoundaryDataSet> >* EBDS,
                                    SpaceVariable3D* Xi, bool run_heat) {
  (void)time; (void)ls_mat_id; (void)KappaPhi; (void)Xi; (void)run_heat;
  if (EBDS && !EBDS->empty()) throw Error(M2CX_E_UNSUPPORTED, "ComputeResidual: embedded boundaries");
  Context& c = *ctx_;
  V.ToDevice(c, M2CX_FIELD_V);
  ID.ToDevice(c, M2CX_FIELD_ID);
  if (Phi)
    for (int i = 0; i < (int)Phi->size() && i < c.num_levelsets(); i++)
      if ((*Phi)[i]) (*Phi)[i]->ToDevice(c, Context::LevelSetField(i));
  SpaceVariable3D::Evict(c, M2CX_FIELD_R);
  const int nfail = c.check(m2cx_compute_residual(c.handle(), M2CX_FIELD_V, riemann_solutions ? 1 : 0));
  R.ProducedOnDevice(c, M2CX_FIELD_R);
  riemann_warning(nfail);
}

// -------------------------------------------------------- LevelSetOperator

void LevelSetOperator::ComputeResidual(SpaceVariable3D& V, SpaceVariable3D& Phi, SpaceVariable3D& R, double time) {
  (void)time;
  Context& c = *ctx_;
  if (R.NumDOF() != 1) throw Error(M2CX_E_ARG, "LevelSetOperator::ComputeResidual: R has one component");
  V.ToDevice(c, M2CX_FIELD_V);
  Phi.ToDevice(c, Context::LevelSetField(ls_));
  c.check(m2cx_levelset_residual(c.handle(), M2CX_FIELD_V));
  // the level-set residual has no field of its own: it comes to R's host copy
  R.SyncToHost();
  R.Unbind();
  c.check(m2cx_download_levelset(c.handle(), ls_, M2CX_LS_RESIDUAL, R.HostData().data()));
  R.RestoreDataPointerAndInsert();
}

bool LevelSetOperator::Reinitialize(double time, double dt, int time_step, SpaceVariable3D& Phi, int special_maxIts,
                                    bool must_do) {
  const m2cx_levelset& s = ctx_->levelset(ls_);
  if (!must_do) {
    // isTimeToWrite (Utils.cpp:337-357) with the reinitialization frequency
    bool due = false;
    if (s.reinit_frequency_dt > 0)
      due = std::floor(time / s.reinit_frequency_dt) != std::floor((time - dt) / s.reinit_frequency_dt);
    else
      due = s.reinit_frequency > 0 && time_step % s.reinit_frequency == 0;
    if (!due) return false;
  }
  Context& c = *ctx_;
  const int f = Context::LevelSetField(ls_);
  Phi.ToDevice(c, f);
  c.check(m2cx_levelset_reinitialize_ls(c.handle(), ls_, special_maxIts));
  Phi.ProducedOnDevice(c, f);
  return true;
}

void LevelSetOperator::ConstructNarrowBandInReinitializer(SpaceVariable3D& Phi) {
  Context& c = *ctx_;
  const int f = Context::LevelSetField(ls_);
  Phi.ToDevice(c, f);
  c.check(m2cx_levelset_construct_narrow_band_ls(c.handle(), ls_));
  Phi.ProducedOnDevice(c, f);
}

// ---------------------------------------------------------- TimeIntegrator

void TimeIntegratorBase::AdvanceOneTimeStep(SpaceVariable3D& V, SpaceVariable3D& ID,
                                            std::vector<SpaceVariable3D*>& Phi, std::vector<SpaceVariable3D*>& NPhi,
                                            std::vector<SpaceVariable3D*>& KappaPhi, SpaceVariable3D* L,
                                            SpaceVariable3D* Xi, SpaceVariable3D* Vturb, SpaceVariable3D* LocalDt,
                                            double time, double dt, int time_step, int subcycle, double dts) {
  (void)NPhi; (void)KappaPhi; (void)dts;
  Context& c = spo_.context();
  const m2cx_scheme& s = c.scheme();
  if (kind_ >= 0 && kind_ != s.integrator)
    throw Error(M2CX_E_ARG, "AdvanceOneTimeStep: the scheme names another time integrator");
  if (L || Xi || Vturb || LocalDt || subcycle != 0)
    throw Error(M2CX_E_UNSUPPORTED, "AdvanceOneTimeStep: laser, ionization, turbulence, local time stepping and "
                                    "subcycling are not on this path");
  const int nls = c.num_levelsets();
  if ((int)Phi.size() < nls) throw Error(M2CX_E_ARG, "AdvanceOneTimeStep: one Phi per level set");
  for (int i = 0; i < nls; i++)
    if (!Phi[i]) throw Error(M2CX_E_ARG, "AdvanceOneTimeStep: one Phi per level set");
  V.ToDevice(c, M2CX_FIELD_V);
  ID.ToDevice(c, M2CX_FIELD_ID);
  for (int i = 0; i < nls; i++) Phi[i]->ToDevice(c, Context::LevelSetField(i));
  // the step works in the stage buffers and the residual buffer
  SpaceVariable3D::Evict(c, M2CX_FIELD_V1);
  SpaceVariable3D::Evict(c, M2CX_FIELD_V2);
  SpaceVariable3D::Evict(c, M2CX_FIELD_R);
  const int nfail = c.check(m2cx_advance_one_time_step(c.handle(), time, dt, time_step));
  V.ProducedOnDevice(c, M2CX_FIELD_V);
  ID.ProducedOnDevice(c, M2CX_FIELD_ID);
  for (int i = 0; i < nls; i++) Phi[i]->ProducedOnDevice(c, Context::LevelSetField(i));
  riemann_warning(nfail);
}

}  // namespace m2c_amd
