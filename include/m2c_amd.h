/*
 * m2c_amd.h -- C ABI of the MI355X explicit finite-volume step.
 *
 * This header is the drop-in boundary between M2C-style host code and the
 * gfx950 HIP kernels in libm2c_amd.so.  Every entry point takes plain
 * pointers and sizes; no HIP, torch or C++ types cross it.  Each function
 * names the reference interface it replaces (kevinwgy/m2c, file:line).
 *
 * Conventions
 *  - Return 0 on success, a negative M2CX_E* code on a fatal error (the
 *    reference would have called exit()), a positive count for warnings
 *    (e.g. Riemann failures, SpaceOperator.cpp:2031-2034).  The library never
 *    exits; m2cx_last_error() gives the message.
 *  - Host arrays use the reference's SpaceVariable3D layout: a ghosted brick of
 *    (nz+2) x (ny+2) x (nx+2) cells, one ghost layer, AoS with `dof` doubles per
 *    cell, i fastest: index ((k+1)*(ny+2) + (j+1))*(nx+2) + (i+1), times dof,
 *    plus p (SpaceVariable.h:86-96, SpaceVariable.cpp:55-162).  Material IDs
 *    are int32 in the same brick (the reference stores them as doubles).
 *  - On the device every field is structure-of-arrays FP64 (one array per
 *    component, same brick); the AoS<->SoA transpose happens only in
 *    upload/download.
 *  - One host thread per handle; all work is ordered on the handle's stream.
 */
#ifndef M2C_AMD_H
#define M2C_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Enumerations, numbered as in the reference's IoData / VarFcnBase.   */
/* ------------------------------------------------------------------ */

/* VarFcnBase::Type (VarFcnBase.h:40-43) */
enum { M2CX_EOS_SG = 0, M2CX_EOS_NASG = 1, M2CX_EOS_MG = 2, M2CX_EOS_DUMMY = 8 };

/* MeshData::BcType (IoData.h:283-284) */
enum { M2CX_BC_NONE = 0, M2CX_BC_INLET = 1, M2CX_BC_INLET2 = 2, M2CX_BC_SLIPWALL = 3,
       M2CX_BC_STICKWALL = 4, M2CX_BC_SYMMETRY = 5, M2CX_BC_OVERSET = 6, M2CX_BC_OUTLET = 7 };

/* LevelSetSchemeData::BcType (IoData.h:814) */
enum { M2CX_LSBC_ZERO_NEUMANN = 1, M2CX_LSBC_LINEAR_EXTRAPOLATION = 2, M2CX_LSBC_NON_NEGATIVE = 3 };

/* SchemeData::Flux (IoData.h:710).  GODUNOV: the physical flux of the exact
 * Riemann solution at the face (FluxFcnGodunov.h:37-60), solved on the device
 * for every face with one material on both sides, before each stage. */
enum { M2CX_FLUX_ROE = 0, M2CX_FLUX_LLF = 1, M2CX_FLUX_HLLC = 2, M2CX_FLUX_GODUNOV = 3 };

/* ReconstructionData (IoData.h:662-671) */
enum { M2CX_REC_CONSTANT = 0, M2CX_REC_LINEAR = 1 };
enum { M2CX_LIM_NONE = 0, M2CX_LIM_GENERALIZED_MINMOD = 1, M2CX_LIM_VANALBADA = 2 };
enum { M2CX_SLOPE_ZERO = 0, M2CX_SLOPE_NONZERO = 1 };
enum { M2CX_VAR_PRIMITIVE = 0, M2CX_VAR_CONSERVATIVE = 1, M2CX_VAR_PRIMITIVE_CHARACTERISTIC = 2,
       M2CX_VAR_CONSERVATIVE_CHARACTERISTIC = 3 };

/* ExplicitData::Type (IoData.h:931) */
enum { M2CX_TI_FORWARD_EULER = 0, M2CX_TI_RK2 = 1, M2CX_TI_RK3 = 2 };

/* MultiPhaseData (IoData.h:893-903) */
enum { M2CX_MPFLUX_EXACT = 0, M2CX_MPFLUX_NUMERICAL = 1, M2CX_MPFLUX_LLF = 2 };
enum { M2CX_MPREC_CONSTANT = 0, M2CX_MPREC_LINEAR = 1 };
enum { M2CX_PHASECHANGE_RIEMANN_SOLUTION = 0, M2CX_PHASECHANGE_EXTRAPOLATION = 1 };

/* MultiPhaseData::PhaseChangeDirection (IoData.h:901): the weights of the
 * neighbours a cell that changed material takes its state from
 * (MultiPhaseOperator.cpp:1220-1246, 1396) -- ALL: inverse distance, UPWIND:
 * the neighbour's flow direction towards the cell (the reference default). */
enum { M2CX_PHASEDIR_ALL = 0, M2CX_PHASEDIR_UPWIND = 1 };
enum { M2CX_RNORMAL_LEVEL_SET = 0, M2CX_RNORMAL_MESH = 1, M2CX_RNORMAL_AVERAGE = 2 };

/* LevelSetSchemeData::Solver and ::Flux (IoData.h:806-810) */
enum { M2CX_LSSOLVER_FINITE_VOLUME = 0, M2CX_LSSOLVER_FINITE_DIFFERENCE = 1 };
enum { M2CX_LSFLUX_ROE = 0, M2CX_LSFLUX_LLF = 1, M2CX_LSFLUX_UPWIND = 2 };

/* LevelSetReinitializationData::FirstLayerTreatment (IoData.h:754-756) */
enum { M2CX_REINIT_FIXED = 0, M2CX_REINIT_CONSTRAINED1 = 1, M2CX_REINIT_CONSTRAINED2 = 2,
       M2CX_REINIT_ITERATIVE_CONSTRAINED1 = 3, M2CX_REINIT_ITERATIVE_CONSTRAINED2 = 4 };

/* Initial guess of the interface Riemann solve. */
enum { M2CX_GUESS_ACOUSTIC = 0,   /* FindInitialFeasiblePointsByAcousticTheory */
       M2CX_GUESS_MLP = 1 };      /* learned p* guess, acoustic fallback       */

/* Error codes */
enum { M2CX_OK = 0, M2CX_E_ARG = -1, M2CX_E_HIP = -2, M2CX_E_OOM = -3,
       M2CX_E_HYPERBOLICITY = -4, M2CX_E_RIEMANN_FATAL = -5, M2CX_E_UNRESOLVED = -6,
       M2CX_E_UNSUPPORTED = -7, M2CX_E_CAPACITY = -8, M2CX_E_REINIT = -9,
       M2CX_E_LEVELSETS = -10 };  /* a node inside two level-set subdomains
                                     (MultiPhaseOperator.cpp:200-206 exits)  */

/* ------------------------------------------------------------------ */
/* Problem description                                                 */
/* ------------------------------------------------------------------ */

/* One material: MaterialModelData (IoData.cpp:1040-1080) with the SG, NASG,
 * MG and dummy parameter sets (VarFcnSG.h:129-144, VarFcnNASG.h:100-113,
 * VarFcnMG.h:104-127, VarFcnDummy.h:23-26). */
typedef struct {
  int32_t eos;                 /* M2CX_EOS_*                                  */
  int32_t reserved;
  double gamma, pc;            /* SG and NASG: specific heat ratio, p_c        */
  double b, q;                 /* NASG: covolume, energy constant              */
  double rho0, c0, Gamma0, s, e0; /* MG                                        */
  double rhomin, pmin, rhomax, pmax; /* clipping (DensityCutOff, ...)          */
  double failsafe_density;
  double dummy_rho, dummy_p, dummy_e;  /* VarFcnDummy state                    */
} m2cx_material;

/* Uniform Cartesian mesh (MeshGenerator::ComputeMesh1DUniform,
 * MeshGenerator.cpp:161-173) and its boundary conditions. */
typedef struct {
  int32_t nx, ny, nz, reserved;
  double x0, xmax, y0, ymax, z0, zmax;
  int32_t bc[6];               /* x0, xmax, y0, ymax, z0, zmax: M2CX_BC_*      */
  double inlet[5];             /* BoundaryConditions.Inlet  (rho,u,v,w,p)      */
  double inlet2[5];            /* BoundaryConditions.Inlet2                    */
  int32_t inlet_id, inlet2_id; /* their material IDs                           */
} m2cx_grid;

/* Numerical scheme: Space.NavierStokes, Reconstruction, Time, MultiPhase,
 * ExactRiemannSolution and LevelSet[0] blocks (defaults: IoData.cpp:1265-1274,
 * 1370-1375, 1513-1535, 1622-1639, 1694-1713, 1771-1852). */
typedef struct {
  int32_t flux;                /* M2CX_FLUX_*                                  */
  int32_t rec_type;            /* M2CX_REC_*                                   */
  int32_t limiter;             /* M2CX_LIM_*                                   */
  int32_t var_type;            /* M2CX_VAR_*                                   */
  int32_t slope_near_interface;/* M2CX_SLOPE_*                                 */
  int32_t integrator;          /* M2CX_TI_*                                    */
  double gmm_coeff;            /* GeneralizedMinModCoefficient (1.2)           */
  double entropy_fix;          /* EntropyFixCoefficient (0.2)                  */
  double cfl;                  /* CFL (0.5)                                    */
  double max_time;             /* MaxTime; <= 0 means no clamp                 */
  int32_t mp_flux;             /* M2CX_MPFLUX_*                                */
  int32_t mp_recon;            /* M2CX_MPREC_*                                 */
  int32_t mp_phase_change;     /* M2CX_PHASECHANGE_* (RiemannSolution)       */
  int32_t riemann_normal;      /* RiemannNormal, M2CX_RNORMAL_* (Mesh)          */
  int32_t riem_maxits_main;    /* MaxIts (200)                                 */
  int32_t riem_maxits_bracket; /* MaxItsBracketing (100)                       */
  int32_t riem_maxits_shock;   /* MaxItsShock (200)                            */
  int32_t riem_steps_rarefaction; /* IntegrationSteps (200)                    */
  double riem_tol_main;        /* Tolerance (1e-5)                             */
  double riem_tol_shock;       /* ToleranceShock (1e-12)                       */
  double riem_tol_rarefaction; /* ToleranceRarefaction (1e-8)                  */
  double riem_min_pressure;    /* -1e8                                         */
  double riem_failure_threshold; /* 0.2                                        */
  double riem_pressure_at_failure; /* 1e-8                                     */
  int32_t n_levelsets;         /* 0 or 1 (m2cx_create_levelsets: up to 4)     */
  int32_t ls_materialid;       /* LevelSet[0].MaterialID                       */
  int32_t ls_bc[6];            /* LevelSet[0].BoundaryCondition* (M2CX_LSBC_*) */
  int32_t riem_guess;          /* M2CX_GUESS_*                                 */
  int32_t mp_phase_dir;        /* PhaseChangeDirection, M2CX_PHASEDIR_* (Upwind) */
  /* LevelSet[0].Reinitialization (IoData.cpp:1421-1429); the level set is
   * reinitialized when reinit_frequency > 0 or reinit_frequency_dt > 0      */
  int32_t reinit_frequency;    /* Frequency (-1)                               */
  int32_t reinit_maxits;       /* MaxIts (30)                                  */
  double reinit_frequency_dt;  /* TimeInterval (-1)                            */
  double reinit_cfl;           /* CFL (0.8)                                    */
  double reinit_tol;           /* ConvergenceTolerance (2e-4)                  */
  int32_t reinit_first_layer;  /* FirstLayerTreatment, M2CX_REINIT_* (Fixed)   */
  /* LevelSet[0].Bandwidth (IoData.cpp:1519, INT_MAX): the narrow-band level
   * set of LevelSetOperator.cpp:84-99 when 2 <= ls_bandwidth < INT32_MAX
   * (requires 1 <= reinit_frequency < ls_bandwidth, as the reference does;
   * finite-difference or finite-volume level set);
   * 0 or INT32_MAX: the level set over the full domain.  The band is built
   * by m2cx_initialize, or by m2cx_levelset_construct_narrow_band after the
   * level set is uploaded.                                                   */
  int32_t ls_bandwidth;
  double time_step;            /* Time.TimeStep: fixed dt when > 0 (0)         */
  /* LevelSet[0] Solver, Flux, Reconstruction (IoData.cpp:1513-1535, 1265-1274) */
  int32_t ls_solver;           /* M2CX_LSSOLVER_* (FiniteDifference)           */
  int32_t ls_flux;             /* finite volume: M2CX_LSFLUX_* (Roe)           */
  double ls_delta;             /* Roe entropy-fix coefficient (0.2)            */
  int32_t ls_rec_type;         /* M2CX_REC_* (Linear)                          */
  int32_t ls_limiter;          /* M2CX_LIM_* (GeneralizedMinMod)               */
  double ls_gmm_coeff;         /* GeneralizedMinModCoefficient (1.2)           */
  /* MultiPhase (IoData.cpp:1706-1708): LevelSetCorrectionFrequency (-1:
   * off; at the steps it divides, ResolveConflictsInLevelSets also runs its
   * optional corrections, MultiPhaseOperator.cpp:1766-1968) and
   * CorrectionsUntilReinitialization (1, at least 1): the nodes
   * ResolveConflictsInLevelSets resolves are counted and every level set is
   * reinitialized once the count reaches this number (TimeIntegrator.cpp:
   * 731-768).                                                               */
  int32_t mp_ls_correction_frequency;
  int32_t mp_corrections_until_reinit;
} m2cx_scheme;

/* Several level sets (LevelSet[i] blocks, IoData.cpp:1513-1568; one
 * LevelSetOperator each, Main.cpp:285-330): level set q tracks material
 * `materialid` (1 .. nmat-1, one level set per material); the fields are
 * those of LevelSet[0] in m2cx_scheme.  Passed to m2cx_create_levelsets;
 * m2cx_create takes level set 0 from the scheme. */
#define M2CX_MAX_LEVELSETS 4
typedef struct {
  int32_t materialid;
  int32_t bc[6];                /* M2CX_LSBC_* */
  int32_t solver, flux;         /* M2CX_LSSOLVER_*, M2CX_LSFLUX_* */
  double delta;
  int32_t rec_type, limiter;
  double gmm_coeff;
  int32_t reinit_frequency, reinit_maxits;
  double reinit_frequency_dt, reinit_cfl, reinit_tol;
  int32_t reinit_first_layer, bandwidth;
} m2cx_levelset;

/* Fields held on the device by a solver handle. */
enum {
  M2CX_FIELD_V = 0,   /* primitive state V (5 dof)            */
  M2CX_FIELD_V1 = 1,  /* RK stage state V1                    */
  M2CX_FIELD_V2 = 2,  /* RK stage state V2                    */
  M2CX_FIELD_R = 3,   /* residual R (5 dof)                   */
  M2CX_FIELD_PHI = 4, /* level set phi (1 dof)                */
  M2CX_FIELD_ID = 5   /* material id (int32, 1 dof)           */
};

/* Counters accumulated since the last m2cx_reset_counters(). */
typedef struct {
  int64_t riemann_calls;       /* interface faces solved                        */
  int64_t riemann_trivial;     /* faces that took the trivial branch (:134-139) */
  int64_t riemann_errors;      /* nonzero returns (SpaceOperator.cpp:2031)      */
  int64_t riemann_main_iters;  /* secant main-loop iterations, summed           */
  int64_t riemann_fevals;      /* f(p) evaluations (pairs of ComputeRhoUStar)   */
  int64_t riemann_guess_fallbacks; /* learned guess rejected -> acoustic        */
  int64_t clipped_cells;       /* ClipDensityAndPressure counts                 */
  int64_t id_changed_cells;    /* cells whose material ID changed               */
  int64_t steps;               /* completed time steps                          */
  int64_t reinit_calls;        /* level-set reinitializations                   */
  int64_t reinit_iterations;   /* their pseudo-time iterations, summed          */
  int64_t unresolved_cells;    /* cells UpdateStateVariablesAfterInterfaceMotion
                                  left unresolved                               */
} m2cx_counters;

typedef struct m2cx_solver m2cx_solver;

/* ------------------------------------------------------------------ */
/* Lifetime                                                            */
/* ------------------------------------------------------------------ */

/* Replaces the construction in Main.cpp:85-491 (VarFcn, FluxFcn,
 * ExactRiemannSolverBase, SpaceOperator, LevelSetOperator, MultiPhaseOperator,
 * TimeIntegrator) for the explicit Euler + level-set path.  Allocates every
 * device field; `device` selects the HIP device. */
int m2cx_create(const m2cx_grid* grid, const m2cx_material* mats, int nmat,
                const m2cx_scheme* scheme, int device, m2cx_solver** out);
/* The same with `nls` level sets (0 .. M2CX_MAX_LEVELSETS) described by
 * `ls`; the scheme's n_levelsets and LevelSet[0] fields are replaced by
 * them.  Several level sets are advanced, reinitialized and banded each by
 * its own settings; the material IDs follow all of them
 * (MultiPhaseOperator::UpdateMaterialIDByLevelSet, MultiPhaseOperator.cpp:
 * 153-357) after ResolveConflictsInLevelSets (:1705-1762) has given each
 * node inside two subdomains to one of them.  Level set q is read and
 * written with m2cx_upload_levelset / m2cx_download_levelset
 * (M2CX_FIELD_PHI is level set 0). */
int m2cx_create_levelsets(const m2cx_grid* grid, const m2cx_material* mats, int nmat,
                          const m2cx_scheme* scheme, const m2cx_levelset* ls, int nls, int device,
                          m2cx_solver** out);
/* Replaces the Destroy() calls in Main.cpp:793-855. */
int m2cx_destroy(m2cx_solver* h);
/* Message of the last failure on this handle (empty string if none). */
const char* m2cx_last_error(const m2cx_solver* h);

/* ------------------------------------------------------------------ */
/* State transfer (SpaceVariable3D GetDataPointer / Restore..., SpaceVariable.h:86-96) */
/* ------------------------------------------------------------------ */

/* Copies a host field (reference layout, see above) to the device field.
 * For M2CX_FIELD_ID `host` points to int32.  Ghosts are copied as given. */
int m2cx_upload(m2cx_solver* h, int field, const void* host);
/* Copies a device field to the host (reference layout, ghosts included). */
int m2cx_download(m2cx_solver* h, int field, void* host);
/* Level set q (0 .. nls-1) of a solver made by m2cx_create_levelsets, or its
 * level-set residual R_phi (what = M2CX_LS_RESIDUAL, after
 * m2cx_levelset_residual); ghosts are copied as given. */
enum { M2CX_LS_PHI = 0, M2CX_LS_RESIDUAL = 1 };
int m2cx_upload_levelset(m2cx_solver* h, int ls, const double* host);
int m2cx_download_levelset(m2cx_solver* h, int ls, int what, double* host);
/* Simulation time / step counter held on the device (Main.cpp:653-712). */
int m2cx_set_time(m2cx_solver* h, double t, int time_step);
int m2cx_get_time(m2cx_solver* h, double* t, double* dt, int* time_step);

/* Initial condition from geometric entities, on the device: the default
 * state and ID everywhere, then each geometry in order overrides the
 * interior cells it covers (SpaceInitializer::SetInitialCondition,
 * SpaceInitializer.cpp:237-482); the level set of scheme.ls_materialid starts
 * at the domain diagonal and takes the signed distance of the nearest
 * surface of that material's geometries, negative inside
 * (InitializePhiByDistance, :1390-1514); then the boundary fill.  Plane:
 * cells with (x - point).n > 0; sphere: |x - point| - radius < 0 (side 0,
 * interior) or > 0 (side 1). */
enum { M2CX_GEOM_PLANE = 0, M2CX_GEOM_SPHERE = 3 };
typedef struct {
  int32_t kind;        /* M2CX_GEOM_*                          */
  int32_t side;        /* sphere: 0 interior, 1 exterior        */
  int32_t materialid;
  int32_t reserved;
  double point[3];     /* plane point / sphere centre           */
  double normal[3];    /* plane normal (normalised here)        */
  double radius;       /* sphere radius                         */
  double state[5];     /* rho, u, v, w, p                       */
} m2cx_geometry;
int m2cx_initialize(m2cx_solver* h, const double* default_state, int default_id, const m2cx_geometry* geoms,
                    int ngeom);

/* ------------------------------------------------------------------ */
/* SpaceOperator (SpaceOperator.h:107-149)                             */
/* ------------------------------------------------------------------ */

/* SpaceOperator::ComputeResidual (SpaceOperator.h:144-149,
 * SpaceOperator.cpp:2573-2641): R = -(sum of face fluxes)/vol on interior
 * cells for state field `v_field` (V, V1 or V2), current ID and phi.
 * Interface faces use ExactRiemannSolverBase (multi-material); when
 * `capture_riemann` is nonzero the star states are kept for the end-of-step
 * update, as RiemannSolutions (RiemannSolutions.h:12-30).  Returns the number
 * of Riemann failures (>=0) or a negative error. */
int m2cx_compute_residual(m2cx_solver* h, int v_field, int capture_riemann);

/* SpaceOperator::ComputeTimeStepSize (SpaceOperator.h:135-136,
 * SpaceOperator.cpp:1298-1330): dt = cfl * min over cells and active axes of
 * dx_a/(|u_a|+c). Synchronises to return dt and cfl. */
int m2cx_compute_time_step(m2cx_solver* h, double* dt, double* cfl);

/* ConservativeToPrimitive / PrimitiveToConservative / ClipDensityAndPressure /
 * ApplyBoundaryConditions (SpaceOperator.h:107-123) on a device field pair.
 * m2cx_clip returns the number of clipped cells. */
int m2cx_apply_boundary_conditions(m2cx_solver* h, int v_field);
int m2cx_clip(m2cx_solver* h, int v_field);

/* LevelSetOperator::ComputeResidual + ApplyBoundaryConditions
 * (LevelSetOperator.h:80-107, LevelSetOperator.cpp:768-870, :408-508)
 * for level set 0 with velocities from `v_field`; result in the device R_phi
 * buffer (download with m2cx_download_ls_residual).  With the narrow band
 * (ComputeResidualFDM_NarrowBand, :880-955) R_phi is cleared on the band and
 * computed on its active nodes; elsewhere it keeps its earlier values. */
int m2cx_levelset_residual(m2cx_solver* h, int v_field);
int m2cx_download_ls_residual(m2cx_solver* h, double* host);

/* LevelSetOperator::Reinitialize with must_do = true (LevelSetOperator.cpp:
 * 1338-1363) -> LevelSetReinitializer::ReinitializeFullDomain
 * (LevelSetReinitializer.cpp:76-176): restores level set 0 to a signed
 * distance by pseudo-time RK3 iterations, first-layer nodes treated as
 * scheme.reinit_first_layer says.  special_maxits > 0 replaces
 * scheme.reinit_maxits (the reference's initializers pass 600 or 2000).
 * Requires reinit_frequency > 0 or reinit_frequency_dt > 0, as the reference
 * requires a reinitializer; returns M2CX_E_UNSUPPORTED otherwise, and
 * M2CX_E_REINIT when the iterations fail to converge with CFL below 0.02
 * (the reference exits there).  Time stepping calls it on its own at the
 * configured frequency. */
int m2cx_levelset_reinitialize(m2cx_solver* h, int special_maxits);
/* With scheme.ls_bandwidth set the reinitialization runs in the band
 * (LevelSetReinitializer::ReinitializeInBand, LevelSetReinitializer.cpp:
 * 181-300), which first moves the band to the interface. */

/* LevelSetOperator::ConstructNarrowBandInReinitializer (LevelSetOperator.cpp:
 * 1328-1334 -> LevelSetReinitializer::ConstructNarrowBand, LevelSetReinitializer
 * .cpp:1479-1592): builds the band around the zero level set of every level
 * set with a bandwidth (levels 0 .. bandwidth by face-neighbour distance, edges and corners of
 * the ghost layer excluded) and cuts phi off outside it to 10 x its largest
 * (smallest) value in the band.  Call after uploading the level set with its
 * boundary values, as SpaceInitializer.cpp:1787-1793 does after
 * InitializePhiByDistance; m2cx_initialize calls it itself.  Stepping,
 * m2cx_levelset_residual and m2cx_levelset_reinitialize need it.
 * M2CX_E_UNSUPPORTED without a band. */
int m2cx_levelset_construct_narrow_band(m2cx_solver* h);
/* One level set's operator alone, as the reference calls it per LevelSet[i]
 * (LevelSetOperator::Reinitialize, LevelSetOperator.cpp:1338-1363;
 * ConstructNarrowBandInReinitializer, :1328-1334): level set `ls` reinitialized
 * (must_do), or its band built; m2cx_levelset_reinitialize and
 * m2cx_levelset_construct_narrow_band do this for every level set. */
int m2cx_levelset_reinitialize_ls(m2cx_solver* h, int ls, int special_maxits);
/* MultiPhaseOperator::UpdateMaterialIDByLevelSet (MultiPhaseOperator.h:
 * 82-83, MultiPhaseOperator.cpp:153-357) with the ID ghost fill: the material
 * IDs from every level set's sign changes between phi_tmp (the level sets at
 * the start of the step: nls ghosted fields, level set 0 first; it replaces
 * the solver's copy) and the device level sets.  Error (M2CX_E_LEVELSETS)
 * where a node enters two subdomains, as the reference exits. */
int m2cx_update_material_id_by_levelset(m2cx_solver* h, const double* phi_tmp);
int m2cx_levelset_construct_narrow_band_ls(m2cx_solver* h, int ls);
/* The band of level set `ls` for inspection: per ghosted cell its level (31415927 outside),
 * useful (level <= bandwidth) and active (level < bandwidth) flags, and
 * bounds = {phi_max, phi_min, phi_out_pos, phi_out_neg}; any may be null.
 * No counterpart in the reference (its Level, UsefulG2 and Active fields are
 * members of LevelSetOperator, LevelSetOperator.h:60-62). */
int m2cx_get_narrow_band(m2cx_solver* h, int ls, int32_t* level, uint8_t* useful, uint8_t* active,
                         double* bounds);

/* ------------------------------------------------------------------ */
/* TimeIntegrator (TimeIntegrator.h:77-88, TimeIntegrator.cpp:167-843) */
/* ------------------------------------------------------------------ */

/* TimeIntegratorBase::AdvanceOneTimeStep for the scheme's FE/RK2/RK3 with a
 * host-given dt (the caller did ComputeTimeStepSize), including the end-of-
 * step material update (UpdateSolutionAfterTimeStepping, :720-843). `time`
 * is the time at the END of the step, as Main.cpp:709-712 passes it. */
int m2cx_advance_one_time_step(m2cx_solver* h, double time, double dt, int time_step);

/* The Main.cpp:653-712 loop body `nsteps` times with dt computed on the
 * device (CFL + MaxTime clamp) and no host synchronisation between steps.
 * Stops early at max_time.  Returns Riemann failures seen, or an error. */
int m2cx_run_steps(m2cx_solver* h, int nsteps);

/* Counters (riemann iterations per face etc.). */
int m2cx_get_counters(m2cx_solver* h, m2cx_counters* out);
int m2cx_reset_counters(m2cx_solver* h);

/* Distribution of the interface Riemann work since the last reset:
 * iters[k] = faces whose secant main loop ran k iterations, fevals[k] = faces
 * with k f(p) evaluations (ExactRiemannSolverBase.cpp:219, :741-743,
 * :1122-1152), the last bin counting k >= M2CX_RIEMANN_HIST_BINS - 1.  All
 * zero for a single-material solver.  No reference counterpart: the reference
 * reports only failures (SpaceOperator.cpp:2031-2034). */
enum { M2CX_RIEMANN_HIST_BINS = 512 };
int m2cx_get_riemann_histogram(m2cx_solver* h, uint64_t* iters, uint64_t* fevals);

/* ------------------------------------------------------------------ */
/* Point kernels, batched (device does the work; host arrays in/out)   */
/* ------------------------------------------------------------------ */

/* FluxFcnBase::ComputeNumericalFluxAtCellInterface (FluxFcnBase.h:65-69) with
 * the scheme's flux (FluxFcnHLLC.h:197-245, FluxFcnGenRoe.h:175-212,
 * FluxFcnLLF.h:82-111) for n faces: Vm, Vp are n x 5, id n, F n x 5. */
int m2cx_numerical_flux_batch(m2cx_solver* h, int n, int dir, const double* Vm,
                              const double* Vp, const int32_t* id, double* F);

/* ExactRiemannSolverBase::ComputeRiemannSolution (ExactRiemannSolverBase.h:42-47,
 * ExactRiemannSolverBase.cpp:59-347) for n problems: dir n x 3, Vm/Vp n x 5,
 * idm/idp n; outputs Vs n x 5, id n, Vsm/Vsp n x 5, err n, and per problem
 * the secant main-loop iteration count and f(p) evaluation count.  A problem
 * on which the reference exits (invalid input state, NaN in the rarefaction
 * integration: ExactRiemannSolverBase.cpp:99,110,1326-1335) gets err = -1,
 * id = -1 and NaN states; a problem whose rarefaction integration path
 * outgrows the device path buffer (512 points) gets err = -2 and NaN states;
 * the other problems are unaffected. */
int m2cx_riemann_batch(m2cx_solver* h, int n, const double* dir,
                       const double* Vm, const int32_t* idm,
                       const double* Vp, const int32_t* idp,
                       double* Vs, int32_t* id, double* Vsm, double* Vsp,
                       int32_t* err, int32_t* iters, int32_t* fevals);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The same
 * solves as m2cx_riemann_batch, bit for bit, but every array is a DEVICE
 * pointer (iters, fevals may be NULL), the work is enqueued on the solver's
 * stream without waiting (m2cx_synchronize before reading the outputs), and
 * nothing is allocated: the integration paths and staging live in one fixed
 * device array per GPU (about 29 MB, held until the process ends), set up by
 * m2cx_create.  At most M2CX_NOALLOC_MAX_PROBLEMS problems per call (more:
 * M2CX_E_ARG with a message).  Every call on one GPU shares that array:
 * calls that overlap on one GPU race. */
#define M2CX_NOALLOC_MAX_PROBLEMS 65536
int m2cx_riemann_batch_noalloc(m2cx_solver* h, int n, const double* dir,
                               const double* Vm, const int32_t* idm,
                               const double* Vp, const int32_t* idp,
                               double* Vs, int32_t* id, double* Vsm, double* Vsp,
                               int32_t* err, int32_t* iters, int32_t* fevals);

/* ------------------------------------------------------------------ */
/* Learned p* guess                                                    */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference.  When scheme.riem_guess is
 * M2CX_GUESS_MLP and parameters are loaded, an MLP (8 -> 64 -> 64 -> 64 -> 1,
 * SiLU, BF16 MFMA with FP32 accumulation) predicts p* of every interface
 * problem; the prediction replaces the acoustic-theory starting pressure of
 * FindInitialFeasiblePointsByAcousticTheory (ExactRiemannSolverBase.cpp:
 * 1107-1158), where the acoustic pressure is not positive or the acoustic
 * wave strength lies in [0.3, 10), and the FP64 iteration after it is
 * unchanged.  Accuracy rule (tests/learned_rule.py): where both waves are
 * shocks the result agrees with the converged root within 10 x the Riemann
 * tolerance (p* on the scale max|p + rho u_n^2/2|, u* on max c, densities
 * relative); with a rarefaction, whose integration path depends on the
 * pressures tried before, the result is as accurate as the plain solver's
 * against a solution with ToleranceRarefaction 1e-11 (99th percentile and
 * maximum error within 1.25 x the plain solver's + 10 x tolerance), not
 * equal to it.  A prediction that is not a finite positive pressure falls
 * back to the acoustic guess (counted in riemann_guess_fallbacks).  Features x[8] and target T of a problem:
 * P = max(pL+pcL, pR+pcR), R = max(rhoL, rhoR), C = sqrt(P/R),
 * x = [log(rhoL/R), log(rhoR/R), unL/C, unR/C, log((pL+pcL)/P),
 * log((pR+pcR)/P), gammaL, gammaR], T = log((p* + min(pcL, pcR))/P).
 * Parameters: FP32 vector of M2CX_MLP_NPARAM values, W1[8][64] b1[64]
 * W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4 (W[in][out]). */
#define M2CX_MLP_NPARAM 8961
int m2cx_mlp_set_params(m2cx_solver* h, const float* theta);
int m2cx_mlp_get_params(m2cx_solver* h, float* theta);
/* features X (n x 8), and targets T (n) when pstar (n) is given; ok[i] = 0
 * marks a problem without valid scales (no guess, excluded from training) */
int m2cx_mlp_features(m2cx_solver* h, int n, const double* dir, const double* Vm, const int32_t* idm,
                      const double* Vp, const int32_t* idp, const double* pstar, float* X, float* T,
                      int32_t* ok);
/* network output y (n) for feature rows X (n x 8) */
int m2cx_mlp_predict(m2cx_solver* h, int n, const float* X, float* y);
/* mean-squared-error loss over (X, T) and its gradient (M2CX_MLP_NPARAM
 * values), computed by the training kernels; n <= 65536 */
int m2cx_mlp_loss_grad(m2cx_solver* h, int n, const float* X, const float* T, float* grad, float* loss);
/* Adam (beta 0.9/0.999, eps 1e-8) on the GPU over shuffled mini-batches;
 * loss_per_epoch (epochs values) receives the mean training loss */
int m2cx_mlp_train(m2cx_solver* h, int n, const float* X, const float* T, int epochs, int batch, float lr,
                   uint32_t seed, float* loss_per_epoch);

/* Stage-kernel tile shapes (no counterpart in the reference): variant 0
 * 16x4x4, 1 32x4x2, 2 8x8x4, 3 16x8x2 cells per work-group.  The first
 * m2cx_run_steps (and m2cx_ensemble_create) times the residual launch of each
 * variant on the current state and keeps the fastest, cached per grid shape
 * and scheme for the process (grids below 64^3 use variant 0).  Results do
 * not depend on the variant.  m2cx_autotune re-times (ms_per_variant: 4
 * values, may be NULL); m2cx_set_tile forces a variant. */
#define M2CX_NUM_TILES 4
int m2cx_autotune(m2cx_solver* h, int* variant, float* ms_per_variant);
int m2cx_set_tile(m2cx_solver* h, int variant);

/* No counterpart in the reference (scheduling only).  With material
 * interfaces, each stage runs the stage kernel's tiles that touch no
 * interface face on a second stream while the interface Riemann problems
 * are solved, and the other tiles after them (default: on; results are the
 * same either way).  m2cx_set_overlap(h, 0) runs one launch per stage after
 * the Riemann solve, e.g. to time the stage kernel alone. */
int m2cx_set_overlap(m2cx_solver* h, int on);

/* ------------------------------------------------------------------ */
/* Ensembles                                                           */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (which runs one simulation per MPI job).
 * Advances many independent solvers (uneven sizes, same device, all RK2 or
 * all RK3) in lockstep: per RK stage the interface Riemann problems of all
 * members are solved in ONE launch, everything else runs per member on one
 * stream, and one ensemble time step (each member with its own CFL dt) is
 * captured once as a HIP graph and replayed.  Each member's results are
 * bit-identical to running it alone with m2cx_run_steps.  Members keep
 * working as solvers (upload/download, counters); the ensemble does not own
 * them and must be destroyed before them. */
typedef struct m2cx_ensemble m2cx_ensemble;
int m2cx_ensemble_create(m2cx_solver* const* members, int n, m2cx_ensemble** out);
int m2cx_ensemble_run_steps(m2cx_ensemble* e, int nsteps);
const char* m2cx_ensemble_last_error(const m2cx_ensemble* e);
int m2cx_ensemble_destroy(m2cx_ensemble* e);

/* ------------------------------------------------------------------ */
/* Measurement                                                         */
/* ------------------------------------------------------------------ */

/* When enabled, every launch of the named kernels is bracketed by HIP events on
 * the handle's stream; m2cx_kernel_time returns the summed milliseconds and
 * launch count for kernel `name` ("stage", "riemann", "dt", "levelset", ...). */
int m2cx_profile(m2cx_solver* h, int enable);
int m2cx_kernel_time(m2cx_solver* h, const char* name, double* ms, int64_t* launches);
int m2cx_synchronize(m2cx_solver* h);

#ifdef __cplusplus
}
#endif

#endif /* M2C_AMD_H */
