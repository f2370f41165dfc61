/*
 * m2c_amd_facade.hpp -- C++ classes with the reference's SpaceVariable3D /
 * SpaceOperator / LevelSetOperator / TimeIntegrator method signatures, over
 * the C ABI of m2c_amd.h (libm2c_amd.so).  Link with libm2c_amd_facade.so.
 *
 * A reference-style driver keeps its calls as they are:
 *
 *   SpaceVariable3D V(grid, 5), ID(grid, 1), Phi0(grid, 1);     // was (comm, &dms.ghosted1_5dof) ...
 *   double*** v = V.GetDataPointer();  ...  V.RestoreDataPointerAndInsert();
 *   spo.ComputeTimeStepSize(V, ID, dt, cfl);
 *   integrator.AdvanceOneTimeStep(V, ID, Phi, NPhi, KappaPhi, L, Xi, Vturb, LocalDt, t, dt, step, 0, dt);
 *   spo.ComputeResidual(V, ID, R, t);
 *
 * Only the constructors differ: there is no MPI communicator or PETSc DMDA
 * (one GPU holds the whole brick), so a SpaceVariable3D is built from the
 * mesh description and its number of components, and the operators share a
 * Context that owns the device solver (m2cx_create).
 *
 * Data stay where they were last produced.  GetDataPointer() returns the
 * host copy, first fetching the device copy if an operator wrote it since
 * (the lazy device -> host sync of SpaceVariable.h:86-96); both Restore...()
 * calls mark the host copy as changed, so the next operator that reads the
 * variable uploads it.  Indexing is the reference's: v[k][j][i*dof + p] with
 * k, j, i from the ghosted corner (-1) to N (one ghost layer).
 *
 * Errors: where the reference calls print_error + exit_mpi, these methods
 * throw m2c_amd::Error (the caller decides how to end); Riemann failures
 * print the reference's warning (SpaceOperator.cpp:2031-2034) and go on.
 */
#ifndef M2C_AMD_FACADE_HPP
#define M2C_AMD_FACADE_HPP

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "m2c_amd.h"

// Reference types that appear in the signatures.  RiemannSolutions: the star
// states of the interface problems stay on the device (a non-null pointer asks
// for them to be kept, for the end-of-step update); EmbeddedBoundaryDataSet
// is not on this path (a non-null EBDS argument is refused).
class RiemannSolutions;
class EmbeddedBoundaryDataSet;

namespace m2c_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class SpaceVariable3D;

// The device solver shared by the operators of one simulation (the objects
// Main.cpp:85-491 builds, for the explicit Euler + level-set step).
class Context {
 public:
  Context(const m2cx_grid& grid, const m2cx_material* mats, int nmat, const m2cx_scheme& scheme, int device = 0);
  // several LevelSet[i] blocks (Main.cpp:285-330): level set i tracks
  // levelsets[i].materialid; level set 0's block overrides the scheme's
  Context(const m2cx_grid& grid, const m2cx_material* mats, int nmat, const m2cx_scheme& scheme,
          const std::vector<m2cx_levelset>& levelsets, int device = 0);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  m2cx_solver* handle() const { return h_; }
  const m2cx_grid& grid() const { return grid_; }
  const m2cx_scheme& scheme() const { return scheme_; }
  int num_levelsets() const { return (int)levelsets_.size(); }
  const m2cx_levelset& levelset(int i) const { return levelsets_.at(i); }
  // the mirror id of level set i: M2CX_FIELD_PHI for 0, past the C ABI's
  // fields for the others
  static constexpr int kFields = 6 + M2CX_MAX_LEVELSETS - 1;
  static int LevelSetField(int i) { return i == 0 ? M2CX_FIELD_PHI : 5 + i; }
  // device field -> host variable that mirrors it (nullptr: none)
  SpaceVariable3D* holder[kFields] = {};
  int check(int rc) const;  // throws on a negative return code

 private:
  m2cx_solver* h_ = nullptr;
  m2cx_grid grid_;
  m2cx_scheme scheme_;
  std::vector<m2cx_levelset> levelsets_;
};

// SpaceVariable.h:50-208 (single rank, one ghost layer)
class SpaceVariable3D {
 public:
  SpaceVariable3D(const m2cx_grid& grid, int dof);
  ~SpaceVariable3D();
  SpaceVariable3D(const SpaceVariable3D&) = delete;
  SpaceVariable3D& operator=(const SpaceVariable3D&) = delete;

  double*** GetDataPointer();
  void RestoreDataPointerAndInsert();
  void RestoreDataPointerToLocalVector();
  void Destroy();

  int NumDOF() const { return dof_; }
  void GetCornerIndices(int* i0, int* j0, int* k0, int* imax = 0, int* jmax = 0, int* kmax = 0) const;
  void GetGhostedCornerIndices(int* i0, int* j0, int* k0, int* imax = 0, int* jmax = 0, int* kmax = 0) const;
  void GetSize(int* nx, int* ny, int* nz) const;
  void GetGhostedSize(int* nx, int* ny, int* nz) const;
  void SetConstantValue(double a, bool workOnGhost = false);
  // this = a * this + b * y (SpaceVariable.cpp:458-483)
  void AXPlusBY(double a, double b, SpaceVariable3D& y, bool workOnGhost = false);

  // --- device mirroring, used by the operators ---
  // makes device field `field` of `ctx` hold this variable's values
  void ToDevice(Context& ctx, int field);
  // records that the operator wrote this variable's values into `field`
  void ProducedOnDevice(Context& ctx, int field);
  // makes way for an operator that writes `field`: whatever variable
  // mirrors that field keeps its values on the host
  static void Evict(Context& ctx, int field);
  // fetches pending device results into the host copy
  void SyncToHost();
  // host copy as rewritten by an operator that produces it on the host
  std::vector<double>& HostData() { return data_; }
  void Unbind();

 private:
  int NX_, NY_, NZ_, dof_;
  std::vector<double> data_;
  std::vector<double*> rows_;
  std::vector<double**> planes_;
  double*** array_ = nullptr;
  Context* ctx_ = nullptr;  // device field this variable mirrors
  int field_ = -1;
  bool host_newer_ = true;     // host copy changed since the device copy was made
  bool device_newer_ = false;  // device copy written since the host copy was fetched
};

// SpaceOperator.h:100-149: the operations of the explicit step
class SpaceOperator {
 public:
  explicit SpaceOperator(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {}
  void ApplyBoundaryConditions(SpaceVariable3D& V);
  int ClipDensityAndPressure(SpaceVariable3D& V, SpaceVariable3D& ID, bool workOnGhost = false,
                             bool checkState = true);
  void ComputeTimeStepSize(SpaceVariable3D& V, SpaceVariable3D& ID, double& dt, double& cfl,
                           SpaceVariable3D* LocalDt = NULL);
  void ComputeResidual(SpaceVariable3D& V, SpaceVariable3D& ID, SpaceVariable3D& R, double time,
                       RiemannSolutions* riemann_solutions = NULL, std::vector<int>* ls_mat_id = NULL,
                       std::vector<SpaceVariable3D*>* Phi = NULL, std::vector<SpaceVariable3D*>* KappaPhi = NULL,
                       std::vector<std::unique_ptr<EmbeddedBoundaryDataSet> >* EBDS = nullptr,
                       SpaceVariable3D* Xi = NULL, bool run_heat = true);
  Context& context() { return *ctx_; }

 private:
  std::shared_ptr<Context> ctx_;
};

// LevelSetOperator.h:75-110, one per level set (Main.cpp:285-330)
class LevelSetOperator {
 public:
  explicit LevelSetOperator(std::shared_ptr<Context> ctx, int ls = 0) : ctx_(std::move(ctx)), ls_(ls) {}
  void ComputeResidual(SpaceVariable3D& V, SpaceVariable3D& Phi, SpaceVariable3D& R, double time);
  bool Reinitialize(double time, double dt, int time_step, SpaceVariable3D& Phi, int special_maxIts = 0,
                    bool must_do = false);
  // LevelSetOperator.cpp:1328-1334: the narrow band around Phi's zero level
  // set (scheme.ls_bandwidth), Phi cut off outside it
  void ConstructNarrowBandInReinitializer(SpaceVariable3D& Phi);
  int GetMaterialID() const { return ctx_->levelset(ls_).materialid; }

 private:
  std::shared_ptr<Context> ctx_;
  int ls_;
};

// TimeIntegrator.h:60-216.  One step of the scheme's integrator, with the
// end-of-step material update (UpdateSolutionAfterTimeStepping); the
// subclasses check that the scheme names them.
class TimeIntegratorBase {
 public:
  TimeIntegratorBase(SpaceOperator& spo, std::vector<LevelSetOperator*>& lso, int kind = -1)
      : spo_(spo), lso_(lso), kind_(kind) {}
  virtual ~TimeIntegratorBase() {}
  // `time` is the time at the end of the step, as Main.cpp:709-712 passes it
  virtual void AdvanceOneTimeStep(SpaceVariable3D& V, SpaceVariable3D& ID, std::vector<SpaceVariable3D*>& Phi,
                                  std::vector<SpaceVariable3D*>& NPhi, std::vector<SpaceVariable3D*>& KappaPhi,
                                  SpaceVariable3D* L, SpaceVariable3D* Xi, SpaceVariable3D* Vturb,
                                  SpaceVariable3D* LocalDt, double time, double dt, int time_step, int subcycle,
                                  double dts);
  virtual void Destroy() {}

 protected:
  SpaceOperator& spo_;
  std::vector<LevelSetOperator*>& lso_;
  int kind_;  // M2CX_TI_*, or -1: whichever the scheme names
};
class TimeIntegratorFE : public TimeIntegratorBase {
 public:
  TimeIntegratorFE(SpaceOperator& spo, std::vector<LevelSetOperator*>& lso)
      : TimeIntegratorBase(spo, lso, M2CX_TI_FORWARD_EULER) {}
};
class TimeIntegratorRK2 : public TimeIntegratorBase {
 public:
  TimeIntegratorRK2(SpaceOperator& spo, std::vector<LevelSetOperator*>& lso)
      : TimeIntegratorBase(spo, lso, M2CX_TI_RK2) {}
};
class TimeIntegratorRK3 : public TimeIntegratorBase {
 public:
  TimeIntegratorRK3(SpaceOperator& spo, std::vector<LevelSetOperator*>& lso)
      : TimeIntegratorBase(spo, lso, M2CX_TI_RK3) {}
};

}  // namespace m2c_amd

#endif
