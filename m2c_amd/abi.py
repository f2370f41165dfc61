"""ctypes mirror of include/m2c_amd.h (the C ABI of libm2c_amd.so).

The structures and enumerations below follow the header field by field.  The
defaults of `default_scheme()` are the reference's IoData defaults
(IoData.cpp:1265-1274, 1370-1375, 1513-1535, 1622-1639, 1694-1713, 1771-1852).
"""
import ctypes as C

EOS_SG, EOS_NASG, EOS_MG, EOS_DUMMY = 0, 1, 2, 8
BC_NONE, BC_INLET, BC_INLET2, BC_SLIPWALL, BC_STICKWALL, BC_SYMMETRY, BC_OVERSET, BC_OUTLET = range(8)
LSBC_ZERO_NEUMANN, LSBC_LINEAR_EXTRAPOLATION, LSBC_NON_NEGATIVE = 1, 2, 3
FLUX_ROE, FLUX_LLF, FLUX_HLLC, FLUX_GODUNOV = 0, 1, 2, 3
REC_CONSTANT, REC_LINEAR = 0, 1
LIM_NONE, LIM_GMM, LIM_VANALBADA = 0, 1, 2
SLOPE_ZERO, SLOPE_NONZERO = 0, 1
VAR_PRIMITIVE, VAR_CONSERVATIVE, VAR_PRIMITIVE_CHARACTERISTIC, VAR_CONSERVATIVE_CHARACTERISTIC = 0, 1, 2, 3
TI_FE, TI_RK2, TI_RK3 = 0, 1, 2
MPFLUX_EXACT, MPFLUX_NUMERICAL, MPFLUX_LLF = 0, 1, 2
MPREC_CONSTANT, MPREC_LINEAR = 0, 1
GUESS_ACOUSTIC, GUESS_MLP = 0, 1
RNORMAL_LEVEL_SET, RNORMAL_MESH, RNORMAL_AVERAGE = 0, 1, 2
PHASECHANGE_RIEMANN_SOLUTION, PHASECHANGE_EXTRAPOLATION = 0, 1
PHASEDIR_ALL, PHASEDIR_UPWIND = 0, 1
LSSOLVER_FINITE_VOLUME, LSSOLVER_FINITE_DIFFERENCE = 0, 1
LSFLUX_ROE, LSFLUX_LLF, LSFLUX_UPWIND = 0, 1, 2
REINIT_FIXED, REINIT_CONSTRAINED1, REINIT_CONSTRAINED2, REINIT_ITERATIVE_CONSTRAINED1, \
    REINIT_ITERATIVE_CONSTRAINED2 = range(5)
FIELD_V, FIELD_V1, FIELD_V2, FIELD_R, FIELD_PHI, FIELD_ID = range(6)
GEOM_PLANE, GEOM_SPHERE = 0, 3
RIEMANN_HIST_BINS = 512
OK, E_ARG, E_HIP, E_OOM, E_HYPERBOLICITY, E_RIEMANN_FATAL, E_UNRESOLVED, E_UNSUPPORTED, E_CAPACITY, \
    E_REINIT, E_LEVELSETS = 0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10

# Python name -> C name where they differ beyond the M2CX_ prefix
C_NAMES = {"EOS_SG": "M2CX_EOS_SG", "LIM_GMM": "M2CX_LIM_GENERALIZED_MINMOD", "TI_FE": "M2CX_TI_FORWARD_EULER"}

DBL_MAX = 1.7976931348623157e308
INT32_MAX = 2147483647


class Material(C.Structure):
    _fields_ = [("eos", C.c_int32), ("reserved", C.c_int32),
                ("gamma", C.c_double), ("pc", C.c_double), ("b", C.c_double), ("q", C.c_double),
                ("rho0", C.c_double), ("c0", C.c_double), ("Gamma0", C.c_double), ("s", C.c_double),
                ("e0", C.c_double),
                ("rhomin", C.c_double), ("pmin", C.c_double), ("rhomax", C.c_double), ("pmax", C.c_double),
                ("failsafe_density", C.c_double),
                ("dummy_rho", C.c_double), ("dummy_p", C.c_double), ("dummy_e", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
                ("x0", C.c_double), ("xmax", C.c_double), ("y0", C.c_double), ("ymax", C.c_double),
                ("z0", C.c_double), ("zmax", C.c_double),
                ("bc", C.c_int32 * 6), ("inlet", C.c_double * 5), ("inlet2", C.c_double * 5),
                ("inlet_id", C.c_int32), ("inlet2_id", C.c_int32)]


class Scheme(C.Structure):
    _fields_ = [("flux", C.c_int32), ("rec_type", C.c_int32), ("limiter", C.c_int32), ("var_type", C.c_int32),
                ("slope_near_interface", C.c_int32), ("integrator", C.c_int32),
                ("gmm_coeff", C.c_double), ("entropy_fix", C.c_double), ("cfl", C.c_double),
                ("max_time", C.c_double),
                ("mp_flux", C.c_int32), ("mp_recon", C.c_int32), ("mp_phase_change", C.c_int32),
                ("riemann_normal", C.c_int32),
                ("riem_maxits_main", C.c_int32), ("riem_maxits_bracket", C.c_int32),
                ("riem_maxits_shock", C.c_int32), ("riem_steps_rarefaction", C.c_int32),
                ("riem_tol_main", C.c_double), ("riem_tol_shock", C.c_double),
                ("riem_tol_rarefaction", C.c_double), ("riem_min_pressure", C.c_double),
                ("riem_failure_threshold", C.c_double), ("riem_pressure_at_failure", C.c_double),
                ("n_levelsets", C.c_int32), ("ls_materialid", C.c_int32), ("ls_bc", C.c_int32 * 6),
                ("riem_guess", C.c_int32), ("mp_phase_dir", C.c_int32),
                ("reinit_frequency", C.c_int32), ("reinit_maxits", C.c_int32),
                ("reinit_frequency_dt", C.c_double), ("reinit_cfl", C.c_double), ("reinit_tol", C.c_double),
                ("reinit_first_layer", C.c_int32), ("ls_bandwidth", C.c_int32), ("time_step", C.c_double),
                ("ls_solver", C.c_int32), ("ls_flux", C.c_int32), ("ls_delta", C.c_double),
                ("ls_rec_type", C.c_int32), ("ls_limiter", C.c_int32), ("ls_gmm_coeff", C.c_double),
                ("mp_ls_correction_frequency", C.c_int32), ("mp_corrections_until_reinit", C.c_int32)]


MAX_LEVELSETS = 4
LS_PHI, LS_RESIDUAL = 0, 1


class Levelset(C.Structure):
    """m2cx_levelset: one LevelSet[i] block"""
    _fields_ = [("materialid", C.c_int32), ("bc", C.c_int32 * 6), ("solver", C.c_int32), ("flux", C.c_int32),
                ("delta", C.c_double), ("rec_type", C.c_int32), ("limiter", C.c_int32), ("gmm_coeff", C.c_double),
                ("reinit_frequency", C.c_int32), ("reinit_maxits", C.c_int32),
                ("reinit_frequency_dt", C.c_double), ("reinit_cfl", C.c_double), ("reinit_tol", C.c_double),
                ("reinit_first_layer", C.c_int32), ("bandwidth", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "riemann_calls", "riemann_trivial", "riemann_errors", "riemann_main_iters", "riemann_fevals",
        "riemann_guess_fallbacks", "clipped_cells", "id_changed_cells", "steps", "reinit_calls",
        "reinit_iterations", "unresolved_cells")]


def material(eos="SG", **kw):
    """MaterialModelData with the reference defaults (IoData.cpp:1040-1053)."""
    m = Material()
    m.eos = {"SG": EOS_SG, "NASG": EOS_NASG, "MG": EOS_MG, "DUMMY": EOS_DUMMY}[eos]
    m.gamma, m.pc = 1.4, 0.0
    m.rhomin, m.pmin, m.rhomax, m.pmax = 0.0, -DBL_MAX, DBL_MAX, DBL_MAX
    m.failsafe_density = 1.0e-10
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def materials(*mats):
    arr = (Material * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = m
    return arr


def grid(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), bc=(BC_OUTLET,) * 6, inlet=None, inlet2=None,
         inlet_id=0, inlet2_id=0):
    g = Grid()
    g.nx, g.ny, g.nz = n
    g.x0, g.y0, g.z0 = lo
    g.xmax, g.ymax, g.zmax = hi
    for i in range(6):
        g.bc[i] = bc[i]
    for i in range(5):
        g.inlet[i] = 0.0 if inlet is None else inlet[i]
        g.inlet2[i] = 0.0 if inlet2 is None else inlet2[i]
    g.inlet_id, g.inlet2_id = inlet_id, inlet2_id
    return g


def levelset_from_scheme(s, **kw):
    """The LevelSet[0] fields of a scheme as one m2cx_levelset, with overrides."""
    L = Levelset()
    L.materialid = s.ls_materialid
    for q in range(6):
        L.bc[q] = s.ls_bc[q]
    L.solver, L.flux, L.delta = s.ls_solver, s.ls_flux, s.ls_delta
    L.rec_type, L.limiter, L.gmm_coeff = s.ls_rec_type, s.ls_limiter, s.ls_gmm_coeff
    L.reinit_frequency, L.reinit_maxits = s.reinit_frequency, s.reinit_maxits
    L.reinit_frequency_dt, L.reinit_cfl, L.reinit_tol = s.reinit_frequency_dt, s.reinit_cfl, s.reinit_tol
    L.reinit_first_layer, L.bandwidth = s.reinit_first_layer, s.ls_bandwidth
    for k, v in kw.items():
        if k == "bc":
            for q in range(6):
                L.bc[q] = v[q]
        else:
            setattr(L, k, v)
    return L


def levelsets(*ls):
    arr = (Levelset * max(1, len(ls)))()
    for i, L in enumerate(ls):
        arr[i] = L
    return arr


def default_scheme(**kw):
    s = Scheme()
    s.flux = FLUX_HLLC
    s.rec_type = REC_LINEAR
    s.limiter = LIM_GMM
    s.var_type = VAR_PRIMITIVE
    s.slope_near_interface = SLOPE_NONZERO
    s.integrator = TI_RK2
    s.gmm_coeff = 1.2
    s.entropy_fix = 0.2
    s.cfl = 0.5
    s.max_time = -1.0
    s.mp_flux = MPFLUX_NUMERICAL
    s.mp_recon = MPREC_CONSTANT
    s.mp_phase_change = PHASECHANGE_RIEMANN_SOLUTION
    s.mp_phase_dir = PHASEDIR_UPWIND
    s.riemann_normal = RNORMAL_MESH
    s.riem_maxits_main = 200
    s.riem_maxits_bracket = 100
    s.riem_maxits_shock = 200
    s.riem_steps_rarefaction = 200
    s.riem_tol_main = 1.0e-5
    s.riem_tol_shock = 1.0e-12
    s.riem_tol_rarefaction = 1.0e-8
    s.riem_min_pressure = -1.0e8
    s.riem_failure_threshold = 0.2
    s.riem_pressure_at_failure = 1.0e-8
    s.n_levelsets = 0
    s.ls_materialid = 1
    for i in range(6):
        s.ls_bc[i] = LSBC_ZERO_NEUMANN
    s.riem_guess = GUESS_ACOUSTIC
    s.reinit_frequency = -1
    s.reinit_maxits = 30
    s.reinit_frequency_dt = -1.0
    s.reinit_cfl = 0.8
    s.reinit_tol = 2.0e-4
    s.reinit_first_layer = REINIT_FIXED
    s.ls_bandwidth = INT32_MAX
    s.mp_ls_correction_frequency = -1
    s.mp_corrections_until_reinit = 1
    s.time_step = 0.0
    s.ls_solver = LSSOLVER_FINITE_DIFFERENCE
    s.ls_flux = LSFLUX_ROE
    s.ls_delta = 0.2
    s.ls_rec_type = REC_LINEAR
    s.ls_limiter = LIM_GMM
    s.ls_gmm_coeff = 1.2
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class Geometry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("side", C.c_int32), ("materialid", C.c_int32), ("reserved", C.c_int32),
                ("point", C.c_double * 3), ("normal", C.c_double * 3), ("radius", C.c_double),
                ("state", C.c_double * 5)]


def geometry(kind, point, normal=(1.0, 0.0, 0.0), radius=0.0, materialid=0, state=(1.0, 0.0, 0.0, 0.0, 1.0),
             side=0):
    """A plane (cells with (x - point).normal > 0) or a sphere (side 0:
    interior) carrying an initial state and material ID."""
    gm = Geometry()
    gm.kind = {"plane": GEOM_PLANE, "sphere": GEOM_SPHERE}[kind]
    gm.side = side
    gm.materialid = materialid
    gm.point[:] = list(point)
    gm.normal[:] = list(normal)
    gm.radius = radius
    gm.state[:] = list(state)
    return gm
