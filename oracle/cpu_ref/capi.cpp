// Oracle (test infrastructure only): extern "C" entry points used by tests/
// (through ctypes) and by bench.py's cpu_baseline leg.  Never linked into the
// product library.
#include <cmath>
#include <cstring>
#include <memory>
#include "solver.h"

using namespace orc;

struct orc_solver {
  std::unique_ptr<Solver> s;
};

static void copy_field(const std::vector<double>& src, double* dst) { std::memcpy(dst, src.data(), src.size() * 8); }

extern "C" {

orc_solver* orc_create(const m2cx_grid* g, const m2cx_material* mats, int nmat, const m2cx_scheme* sch) {
  auto* h = new orc_solver;
  h->s.reset(new Solver(*g, mats, nmat, *sch));
  return h;
}

// several level sets (m2cx_create_levelsets)
orc_solver* orc_create_levelsets(const m2cx_grid* g, const m2cx_material* mats, int nmat, const m2cx_scheme* sch,
                                 const m2cx_levelset* ls, int nls) {
  auto* h = new orc_solver;
  m2cx_scheme s = *sch;
  s.n_levelsets = nls;
  if (nls > 0) SchemeSetLS(s, ls[0]);
  h->s.reset(new Solver(*g, mats, nmat, s));
  h->s->SetLevelSets(ls, nls);
  return h;
}

// level set q (what = 0) or its copy at the start of the step, Phi_tmp
// (what = 2)
int orc_set_levelset(orc_solver* h, int q, int what, const double* phi) {
  Solver& s = *h->s;
  if (q < 0 || q >= std::max(s.nls, 1)) return -2;
  s.SwapLS(q);
  std::memcpy((what == 2 ? s.Phi_tmp : s.Phi).data(), phi, s.m.ncells() * 8);
  s.SwapLS(q);
  return 0;
}

// MultiPhaseOperator::UpdateMaterialIDByLevelSet from Phi_tmp to Phi
int orc_update_material_ids(orc_solver* h) {
  Solver& s = *h->s;
  s.UpdateMaterialIDByLevelSets();
  return s.fatal ? -1 : 0;
}

// MultiPhaseOperator::ResolveConflictsInLevelSets; returns the nodes resolved
int orc_resolve_conflicts(orc_solver* h) { return h->s->ResolveConflictsInLevelSets(); }

int orc_get_levelset(orc_solver* h, int q, int what, double* out) {
  Solver& s = *h->s;
  if (q < 0 || q >= std::max(s.nls, 1)) return -2;
  s.SwapLS(q);
  copy_field(what == 1 ? s.Rls : s.Phi, out);
  s.SwapLS(q);
  return 0;
}

void orc_destroy(orc_solver* h) { delete h; }

const char* orc_error(orc_solver* h) { return h->s->msg.c_str(); }

long orc_ncells(orc_solver* h) { return h->s->m.ncells(); }

int orc_set_state(orc_solver* h, const double* V, const int* ID, const double* Phi) {
  Solver& s = *h->s;
  long n = s.m.ncells();
  if (V) std::memcpy(s.V.data(), V, n * 5 * 8);
  if (ID) std::memcpy(s.ID.data(), ID, n * 4);
  if (Phi) std::memcpy(s.Phi.data(), Phi, n * 8);
  return 0;
}

int orc_get_state(orc_solver* h, double* V, int* ID, double* Phi) {
  Solver& s = *h->s;
  long n = s.m.ncells();
  if (V) copy_field(s.V, V);
  if (ID) std::memcpy(ID, s.ID.data(), n * 4);
  if (Phi) copy_field(s.Phi, Phi);
  return 0;
}

int orc_get_field(orc_solver* h, const char* name, double* out) {
  Solver& s = *h->s;
  std::vector<double>* f = nullptr;
  if (!strcmp(name, "V")) f = &s.V;
  else if (!strcmp(name, "V1")) f = &s.V1;
  else if (!strcmp(name, "V2")) f = &s.V2;
  else if (!strcmp(name, "R")) f = &s.R;
  else if (!strcmp(name, "U1")) f = &s.U1;
  else if (!strcmp(name, "Un")) f = &s.Un;
  else if (!strcmp(name, "Vl")) f = &s.Vl;
  else if (!strcmp(name, "Vr")) f = &s.Vr;
  else if (!strcmp(name, "Vb")) f = &s.Vb;
  else if (!strcmp(name, "Vt")) f = &s.Vt;
  else if (!strcmp(name, "Vk")) f = &s.Vk;
  else if (!strcmp(name, "Vf")) f = &s.Vf;
  else if (!strcmp(name, "Phi")) f = &s.Phi;
  else if (!strcmp(name, "Phi1")) f = &s.Phi1;
  else if (!strcmp(name, "Rls")) f = &s.Rls;
  else if (!strcmp(name, "vol")) f = &s.vol;
  if (!f) return -1;
  copy_field(*f, out);
  return 0;
}

// SpaceOperator::ComputeResidual on the current V; R written to `R` (full
// ghosted brick, 5 dof).
int orc_residual(orc_solver* h, double* R, int capture) {
  Solver& s = *h->s;
  s.ComputeResidual(s.V, s.R, capture != 0);
  if (s.fatal) return -1;
  if (R) copy_field(s.R, R);
  return (int)s.cnt.riemann_errors;
}

int orc_apply_bc(orc_solver* h) { h->s->ApplyBoundaryConditions(h->s->V); return 0; }
int orc_update_ghost_ids(orc_solver* h) { h->s->UpdateMaterialIDAtGhostNodes(); return 0; }
int orc_ls_apply_bc(orc_solver* h) {
  Solver& s = *h->s;
  s.ForEachLS([&](int) { s.LSApplyBoundaryConditions(s.Phi); });
  return 0;
}

// residual of every level set; `out` receives level set 0's
int orc_ls_residual(orc_solver* h, double* out) {
  Solver& s = *h->s;
  s.ForEachLS([&](int) { s.LSComputeResidual(s.V, s.Phi, s.Rls); });
  if (out) copy_field(s.Rls, out);
  return 0;
}

int orc_time_step(orc_solver* h, double* dt, double* cfl) {
  h->s->ComputeTimeStepSize(h->s->V, *dt, *cfl);
  return h->s->fatal ? -1 : 0;
}

int orc_advance_one(orc_solver* h, double time, double dt) {
  Solver& s = *h->s;
  s.time_step++;
  s.AdvanceOneTimeStep(time, dt);
  return s.fatal ? -1 : 0;
}

int orc_run_steps(orc_solver* h, int nsteps) { return h->s->RunSteps(nsteps); }

int orc_get_time(orc_solver* h, double* t, double* dt, int* step) {
  *t = h->s->t;
  *dt = h->s->dt_last;
  *step = h->s->time_step;
  return 0;
}

int orc_set_time(orc_solver* h, double t, int step) {
  h->s->t = t;
  h->s->time_step = step;
  return 0;
}

// counters: calls, trivial, errors, iters, fevals, fallbacks, clipped, id_changed, steps,
// reinit calls, reinit iterations, unresolved cells
int orc_counters(orc_solver* h, long* out) {
  Counters& c = h->s->cnt;
  long v[12] = {c.riemann_calls, c.riemann_trivial, c.riemann_errors, c.riemann_iters, c.riemann_fevals,
                c.riemann_fallbacks, c.clipped, c.id_changed, c.steps, c.reinit_calls, c.reinit_iters,
                c.unresolved};
  std::memcpy(out, v, sizeof v);
  return 0;
}

// The Riemann direction of face d of cell (i, j, k) from the current level set
int orc_bimaterial_normal(orc_solver* h, int d, int i, int j, int k, int myid, int neighborid, double* dir) {
  h->s->BimaterialNormal(d, i, j, k, myid, neighborid, dir);
  return 0;
}

// LevelSetOperator::Reinitialize with must_do = true
int orc_reinitialize(orc_solver* h, int special_maxits) {
  Solver& s = *h->s;
  s.ForEachLS([&](int) { s.Reinitialize(0.0, 1.0, 0, special_maxits, true); });
  return s.fatal ? -1 : 0;
}

double orc_reinit_cfl(orc_solver* h) { return h->s->reinit_cfl; }

// LevelSetOperator::ConstructNarrowBandInReinitializer (LevelSetOperator.cpp:
// 1328-1334), after the level set and its boundary values are set
int orc_construct_narrow_band(orc_solver* h) {
  Solver& s = *h->s;
  int any = 0;
  s.ForEachLS([&](int) {
    if (!s.nb_bw) return;
    s.ConstructNarrowBand();
    any = 1;
  });
  if (!any) return -2;
  return s.fatal ? -1 : 0;
}

// The band of level set 0: level per ghosted cell (31415927 outside), the
// useful and active masks, and phi_max, phi_min, phi_out_pos, phi_out_neg.
int orc_get_narrow_band(orc_solver* h, int ls, int* level, unsigned char* useful, unsigned char* active,
                        double* bounds) {
  Solver& s = *h->s;
  if (ls < 0 || ls >= std::max(s.nls, 1)) return -2;
  s.SwapLS(ls);
  struct Back { Solver& s; int q; ~Back() { s.SwapLS(q); } } back{s, ls};
  if (!s.nb_built) return -2;
  const long n = s.m.ncells();
  if (level) std::memcpy(level, s.nbLevel.data(), n * 4);
  if (useful) std::memcpy(useful, s.nbUseful.data(), n);
  if (active) std::memcpy(active, s.nbActive.data(), n);
  if (bounds) {
    bounds[0] = s.phi_max;
    bounds[1] = s.phi_min;
    bounds[2] = s.phi_out_pos;
    bounds[3] = s.phi_out_neg;
  }
  return 0;
}

// Prescribed-velocity level-set test 1, 2 or 3 (0: off)
int orc_set_ls_test(orc_solver* h, int test) {
  h->s->ls_test = test;
  return 0;
}

// Record of interface Riemann problems: enable (clears), then read up to
// `cap` rows of 17 doubles; returns the number of rows recorded.
int orc_riemann_log_enable(orc_solver* h, int on) {
  h->s->riem_log_on = on != 0;
  h->s->riem_log.clear();
  return 0;
}

long orc_riemann_log(orc_solver* h, double* out, long cap) {
  const std::vector<double>& L = h->s->riem_log;
  long n = (long)L.size() / 17;
  for (long r = 0; r < n && r < cap; r++)
    for (int c = 0; c < 17; c++) out[r * 17 + c] = L[r * 17 + c];
  return n;
}

// Point kernels -----------------------------------------------------------

int orc_flux_batch(const m2cx_material* mats, int nmat, const m2cx_scheme* sch, int n, int dir, const double* Vm,
                   const double* Vp, const int* id, double* F) {
  std::vector<Eos> vf;
  for (int i = 0; i < nmat; i++) vf.push_back(Eos::from(mats[i]));
  Flux fl;
  fl.vf = &vf;
  fl.kind = sch->flux;
  fl.del = sch->entropy_fix;
  ExactRiemann godunov(vf, RiemannParams::from(*sch));
  fl.godunov = &godunov;
  for (int q = 0; q < n; q++) fl.Numerical(dir, Vm + 5 * q, Vp + 5 * q, id[q], F + 5 * q);
  return fl.godunov_fatal ? -1 : 0;
}

// The two-material LLF flux of MultiPhase.Flux = LocalLaxFriedrichs
// (FluxFcnLLF.h:119-185): each side with its own material
int orc_llf_interface_batch(const m2cx_material* mats, int nmat, int n, int dir, const double* Vm, const int* idm,
                            const double* Vp, const int* idp, double* F) {
  std::vector<Eos> vf;
  for (int i = 0; i < nmat; i++) vf.push_back(Eos::from(mats[i]));
  Flux fl;
  fl.vf = &vf;
  fl.kind = M2CX_FLUX_LLF;
  fl.del = 0.0;
  for (int q = 0; q < n; q++) fl.Llf(dir, Vm + 5 * q, idm[q], Vp + 5 * q, idp[q], F + 5 * q);
  return 0;
}

int orc_riemann_batch_guess(const m2cx_material* mats, int nmat, const m2cx_scheme* sch, int n, const double* dir,
                            const double* Vm, const int* idm, const double* Vp, const int* idp, double* Vs, int* id,
                            double* Vsm, double* Vsp, int* err, int* iters, int* fevals, int* flags,
                            const double* guess) {
  std::vector<Eos> vf;
  for (int i = 0; i < nmat; i++) vf.push_back(Eos::from(mats[i]));
  ExactRiemann rs(vf, RiemannParams::from(*sch));
  for (int q = 0; q < n; q++) {
    rs.st = RiemannStats();
    // a learned starting pressure (finite and positive) or none
    rs.p0_guess = (guess && std::isfinite(guess[q]) && guess[q] > 0) ? guess[q] : -1.0;
    int mid = -1;
    err[q] = rs.Solve(dir + 3 * q, Vm + 5 * q, idm[q], Vp + 5 * q, idp[q], Vs + 5 * q, mid, Vsm + 5 * q,
                      Vsp + 5 * q);
    id[q] = mid;
    if (rs.st.fatal) {  // the reference exits here; the batch marks the problem instead
      err[q] = -1;
      id[q] = -1;
      for (int p = 0; p < 5; p++) Vs[5 * q + p] = Vsm[5 * q + p] = Vsp[5 * q + p] = std::nan("");
    }
    if (iters) iters[q] = rs.st.it_main;
    if (fevals) fevals[q] = rs.st.fevals;
    if (flags) flags[q] = (rs.st.trivial ? 1 : 0) | (rs.st.fallback ? 2 : 0) | (rs.st.fatal ? 4 : 0);
  }
  return 0;
}

int orc_riemann_batch(const m2cx_material* mats, int nmat, const m2cx_scheme* sch, int n, const double* dir,
                      const double* Vm, const int* idm, const double* Vp, const int* idp, double* Vs, int* id,
                      double* Vsm, double* Vsp, int* err, int* iters, int* fevals, int* flags) {
  return orc_riemann_batch_guess(mats, nmat, sch, n, dir, Vm, idm, Vp, idp, Vs, id, Vsm, Vsp, err, iters, fevals,
                                 flags, nullptr);
}

// One Riemann problem in the PRINT_RIEMANN_SOLUTION build: fills `rows`
// (xi, rho, u, p, id) up to `cap` rows and returns the row count.
int orc_riemann_print(const m2cx_material* mats, int nmat, const m2cx_scheme* sch, const double* Vm, int idm,
                      const double* Vp, int idp, double* rows, int cap, double* star, int nonadaptive) {
  std::vector<Eos> vf;
  for (int i = 0; i < nmat; i++) vf.push_back(Eos::from(mats[i]));
  ExactRiemann rs(vf, RiemannParams::from(*sch));
  rs.print_mode = true;
  rs.nonadaptive = nonadaptive != 0;
  double dir[3] = {1.0, 0.0, 0.0};
  double Vs[5], Vsm[5], Vsp[5];
  int mid = -1;
  rs.Solve(dir, Vm, idm, Vp, idp, Vs, mid, Vsm, Vsp);
  int nr = (int)rs.sol1d.size();
  for (int r = 0; r < nr && r < cap; r++)
    for (int c = 0; c < 5; c++) rows[r * 5 + c] = rs.sol1d[r][c];
  if (star) {
    star[0] = Vsm[4]; star[1] = Vsm[1]; star[2] = Vsm[0]; star[3] = Vsp[0];
  }
  return nr;
}

// Wave curve (ComputeRhoUStar) from a fresh path, for closed-form tests.
int orc_wave(const m2cx_material* mats, int nmat, const m2cx_scheme* sch, int wave, double rho, double u, double p,
             double ps, int idd, double rhos0, double rhos1, double* rhos, double* us) {
  std::vector<Eos> vf;
  for (int i = 0; i < nmat; i++) vf.push_back(Eos::from(mats[i]));
  ExactRiemann rs(vf, RiemannParams::from(*sch));
  ExactRiemann::Path path;
  path.push_back({p, rho, u});
  return rs.RhoUStar(wave, path, rho, u, p, ps, idd, rhos0, rhos1, *rhos, *us) ? 1 : 0;
}

// Equation-of-state point values for closed-form tests: out = {e(rho, p),
// p(rho, e), dp/drho(rho, e), Gamma(rho), c^2(rho, e)} with e = e(rho, p).
int orc_eos_eval(const m2cx_material* mat, double rho, double p, double* out) {
  Eos E = Eos::from(*mat);
  const double e = E.GetInternalEnergyPerUnitMass(rho, p);
  out[0] = e;
  out[1] = E.GetPressure(rho, e);
  out[2] = E.GetDpdrho(rho, e);
  out[3] = E.GetBigGamma(rho, e);
  out[4] = E.ComputeSoundSpeedSquare(rho, e);
  return 0;
}

double orc_cr_root(double x, int n) { return m2c::cr_root(x, n); }

void orc_set_ck_textbook(int on) { orc::g_ck_textbook = on != 0; }

}  // extern "C"
