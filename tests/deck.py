"""Turn a parsed reference input deck (tests/golden/riemann_problems.json) into
ABI structures and the initial fields, as Main.cpp:85-382 and
SpaceInitializer.cpp:98-482,1390-1514 set them up (uniform mesh, plane and
sphere geometries, default state = Inlet, level set by signed distance)."""
import numpy as np

from m2c_amd import abi

BC_NAMES = {"Inlet": abi.BC_INLET, "Farfield": abi.BC_INLET, "Inlet2": abi.BC_INLET2, "Farfield2": abi.BC_INLET2,
            "SlipWall": abi.BC_SLIPWALL, "StickWall": abi.BC_STICKWALL, "Symmetry": abi.BC_SYMMETRY,
            "Outlet": abi.BC_OUTLET}
FLUX = {"Roe": abi.FLUX_ROE, "LocalLaxFriedrichs": abi.FLUX_LLF, "HLLC": abi.FLUX_HLLC, "Godunov": abi.FLUX_GODUNOV}
LIM = {"None": abi.LIM_NONE, "GeneralizedMinMod": abi.LIM_GMM, "VanAlbada": abi.LIM_VANALBADA}
VAR = {"Primitive": abi.VAR_PRIMITIVE, "Conservative": abi.VAR_CONSERVATIVE,
       "PrimitiveCharacteristic": abi.VAR_PRIMITIVE_CHARACTERISTIC,
       "ConservativeCharacteristic": abi.VAR_CONSERVATIVE_CHARACTERISTIC}
TI = {"ForwardEuler": abi.TI_FE, "RungeKutta2": abi.TI_RK2, "RungeKutta3": abi.TI_RK3}


# MaterialModelData keys (IoData.cpp:1074-1080).  The deck parser resolves a
# name in the innermost enclosing "under" block that knows it (Parser.y:66-76),
# so these may also be written inside the EOS sub-block.
_CUTOFFS = {"DensityCutOff": "rhomin", "PressureCutOff": "pmin", "DensityUpperLimit": "rhomax",
            "PressureUpperLimit": "pmax", "DensityPrescribedAtFailure": "failsafe_density"}


def _mat(md):
    eos = md.get("EquationOfState", "StiffenedGas")
    if eos == "StiffenedGas":
        sub = md.get("StiffenedGasModel", {})
        m = abi.material("SG", gamma=sub.get("SpecificHeatRatio", 1.4), pc=sub.get("PressureConstant", 0.0))
    elif eos == "NobleAbelStiffenedGas":
        sub = md.get("NobleAbelStiffenedGasModel", {})
        m = abi.material("NASG", gamma=sub.get("SpecificHeatRatio", 1.4), pc=sub.get("PressureConstant", 0.0),
                         b=sub.get("VolumeConstant", 0.0), q=sub.get("EnergyConstant", 0.0))
    elif eos == "MieGruneisen":
        sub = md.get("MieGruneisenModel", {})  # defaults IoData.cpp:708-724
        m = abi.material("MG", rho0=sub.get("ReferenceDensity", 8.96e-3), c0=sub.get("BulkSpeedOfSound", 3.933e6),
                         Gamma0=sub.get("ReferenceGamma", 1.99), s=sub.get("HugoniotSlope", 1.5),
                         e0=sub.get("ReferenceSpecificInternalEnergy", 0.0))
    else:
        raise NotImplementedError(eos)
    for key, field in _CUTOFFS.items():
        if key in sub:
            setattr(m, field, sub[key])
        if key in md:
            setattr(m, field, md[key])
    return m


def supported_eos(deck):
    mats = deck["Equations"]
    return all(mats[k].get("EquationOfState", "StiffenedGas") in ("StiffenedGas", "MieGruneisen",
                                                                   "NobleAbelStiffenedGas") for k in mats)


def materials(deck):
    mats = deck["Equations"]
    keys = sorted(mats, key=lambda k: int(k.split("[")[1][:-1]))
    return abi.materials(*[_mat(mats[k]) for k in keys]), len(keys)


def state(sd):
    return [sd.get("Density", 1.0), sd.get("VelocityX", 0.0), sd.get("VelocityY", 0.0),
            sd.get("VelocityZ", 0.0), sd.get("Pressure", 1.0)], int(sd.get("MaterialID", 0))


def setup(deck):
    """Returns (grid, mats, nmat, scheme, V, ID, Phi) with V/ID/Phi in the
    reference's ghosted AoS layout."""
    me = deck["Mesh"]
    n = (int(me["NumberOfCellsX"]), int(me["NumberOfCellsY"]), int(me["NumberOfCellsZ"]))
    bcs = [BC_NAMES[me.get(k, "Outlet")] for k in ("BoundaryConditionX0", "BoundaryConditionXmax",
                                                   "BoundaryConditionY0", "BoundaryConditionYmax",
                                                   "BoundaryConditionZ0", "BoundaryConditionZmax")]
    # "Farfield" and "Farfield2" name the same states (IoData.cpp:1974-1979)
    bcd = deck.get("BoundaryConditions", {})
    inlet, inlet_id = state(bcd.get("Inlet", bcd.get("Farfield", {})))
    inlet2, inlet2_id = state(bcd.get("Inlet2", bcd.get("Farfield2", {})))
    g = abi.grid(n, (me["X0"], me["Y0"], me["Z0"]), (me["Xmax"], me["Ymax"], me["Zmax"]), bcs, inlet, inlet2,
                 inlet_id, inlet2_id)
    mats, nmat = materials(deck)
    ns = deck["Space"]["NavierStokes"]
    rec = ns.get("Reconstruction", {})
    mp = deck.get("MultiPhase", {})
    ti = deck["Time"]
    s = abi.default_scheme()
    s.flux = FLUX[ns.get("Flux", "HLLC")]
    s.rec_type = abi.REC_LINEAR if rec.get("Type", "Linear") == "Linear" else abi.REC_CONSTANT
    s.limiter = LIM[rec.get("Limiter", "GeneralizedMinMod")]
    s.gmm_coeff = rec.get("GeneralizedMinModCoefficient", 1.2)
    s.var_type = VAR[rec.get("VariableType", "Primitive")]
    s.slope_near_interface = abi.SLOPE_ZERO if rec.get("SlopeNearInterface", "NonZero") == "Zero" else abi.SLOPE_NONZERO
    s.entropy_fix = ns.get("EntropyFixCoefficient", 0.2)
    s.integrator = TI[ti.get("Explicit", {}).get("Type", "RungeKutta2")]
    s.cfl = ti.get("CFL", 0.5)
    s.max_time = ti.get("MaxTime", -1.0)
    s.mp_flux = {"Exact": abi.MPFLUX_EXACT, "Numerical": abi.MPFLUX_NUMERICAL,
                 "LocalLaxFriedrichs": abi.MPFLUX_LLF}[mp.get("Flux", "Numerical")]
    s.mp_recon = abi.MPREC_LINEAR if mp.get("ReconstructionAtInterface", "Constant") == "Linear" else abi.MPREC_CONSTANT
    s.mp_phase_change = {"RiemannSolution": abi.PHASECHANGE_RIEMANN_SOLUTION,
                         "Extrapolation": abi.PHASECHANGE_EXTRAPOLATION}[mp.get("PhaseChange", "RiemannSolution")]
    s.mp_phase_dir = {"All": abi.PHASEDIR_ALL, "Upwind": abi.PHASEDIR_UPWIND}[mp.get("PhaseChangeDirection", "Upwind")]
    s.riemann_normal = {"LevelSet": abi.RNORMAL_LEVEL_SET, "Mesh": abi.RNORMAL_MESH,
                        "Average": abi.RNORMAL_AVERAGE}[mp.get("RiemannNormal", "Mesh")]
    ls = [k for k in deck["Space"] if k.startswith("LevelSet")]
    s.n_levelsets = len(ls)
    if ls:
        lsd = deck["Space"][ls[0]]
        s.ls_materialid = int(lsd["MaterialID"])
        s.ls_solver = {"FiniteVolume": abi.LSSOLVER_FINITE_VOLUME,
                       "FiniteDifference": abi.LSSOLVER_FINITE_DIFFERENCE}[lsd.get("Solver", "FiniteDifference")]
        s.ls_flux = {"Roe": abi.LSFLUX_ROE, "LocalLaxFriedrichs": abi.LSFLUX_LLF,
                     "Upwind": abi.LSFLUX_UPWIND}[lsd.get("Flux", "Roe")]
        s.ls_delta = lsd.get("EntropyFixCoefficient", 0.2)
        lrec = lsd.get("Reconstruction", {})
        s.ls_rec_type = abi.REC_LINEAR if lrec.get("Type", "Linear") == "Linear" else abi.REC_CONSTANT
        s.ls_limiter = LIM[lrec.get("Limiter", "GeneralizedMinMod")]
        s.ls_gmm_coeff = lrec.get("GeneralizedMinModCoefficient", 1.2)
        for q, key in enumerate(("BoundaryConditionX0", "BoundaryConditionXmax", "BoundaryConditionY0",
                                 "BoundaryConditionYmax", "BoundaryConditionZ0", "BoundaryConditionZmax")):
            s.ls_bc[q] = {"ZeroNeumann": abi.LSBC_ZERO_NEUMANN, "LinearExtrapolation": abi.LSBC_LINEAR_EXTRAPOLATION,
                          "NonNegative": abi.LSBC_NON_NEGATIVE}[lsd.get(key, "ZeroNeumann")]
        rei = lsd.get("Reinitialization", {})
        s.reinit_frequency = int(rei.get("Frequency", -1))
        s.reinit_frequency_dt = rei.get("TimeInterval", -1.0)
        s.reinit_maxits = int(rei.get("MaxIts", 30))
        s.reinit_cfl = rei.get("CFL", 0.8)
        s.reinit_tol = rei.get("ConvergenceTolerance", 2.0e-4)
        s.reinit_first_layer = {"Fixed": abi.REINIT_FIXED, "ConstrainedMethod1": abi.REINIT_CONSTRAINED1,
                                "ConstrainedMethod2": abi.REINIT_CONSTRAINED2,
                                "IterativelyConstrainedMethod1": abi.REINIT_ITERATIVE_CONSTRAINED1,
                                "IterativelyConstrainedMethod2": abi.REINIT_ITERATIVE_CONSTRAINED2}[
                                    rei.get("FirstLayerTreatment", "Fixed")]
    s.time_step = ti.get("TimeStep", 0.0)
    # ExactRiemannSolverData (IoData.cpp:1640-1675): keys the deck sets
    ers = deck.get("ExactRiemannSolution", {})
    for key, field in (("MaxIts", "riem_maxits_main"), ("MaxItsBracketing", "riem_maxits_bracket"),
                       ("MaxItsShock", "riem_maxits_shock"), ("IntegrationSteps", "riem_steps_rarefaction")):
        if key in ers:
            setattr(s, field, int(ers[key]))
    for key, field in (("Tolerance", "riem_tol_main"), ("ToleranceShock", "riem_tol_shock"),
                       ("ToleranceRarefaction", "riem_tol_rarefaction"), ("MinPressure", "riem_min_pressure"),
                       ("FailureThreshold", "riem_failure_threshold"),
                       ("PrescribedPressureUponFailure", "riem_pressure_at_failure")):
        if key in ers:
            setattr(s, field, float(ers[key]))
    # initial condition
    px, py, pz = n[0] + 2, n[1] + 2, n[2] + 2
    V = np.zeros((pz, py, px, 5))
    ID = np.zeros((pz, py, px), dtype=np.int32)
    V[...] = inlet
    ID[...] = inlet_id
    x = _axis(n[0], me["X0"], me["Xmax"])
    y = _axis(n[1], me["Y0"], me["Ymax"])
    z = _axis(n[2], me["Z0"], me["Zmax"])
    diag = np.sqrt((me["Xmax"] - me["X0"]) ** 2 + (me["Ymax"] - me["Y0"]) ** 2 + (me["Zmax"] - me["Z0"]) ** 2)
    Phi = np.full((pz, py, px), diag)
    geo = deck.get("InitialCondition", {}).get("GeometricEntities", {})
    X, Y, Z = np.meshgrid(x[1:-1], y[1:-1], z[1:-1], indexing="ij")
    X, Y, Z = X.transpose(2, 1, 0), Y.transpose(2, 1, 0), Z.transpose(2, 1, 0)
    # geometries in the reference's default order: planes, then spheres, each
    # by index (SpaceInitializer.cpp:148-198)
    kinds = {"Plane": 0, "Sphere": 3}
    for key in sorted(geo, key=lambda k: (kinds.get(k.split("[")[0], 99), int(k.split("[")[1][:-1]))):
        kind = key.split("[")[0]
        if kind not in kinds:
            raise NotImplementedError(key)
        pl = geo[key]
        if kind == "Plane":
            x0 = np.array([pl.get("Point_x", 0.0), pl.get("Point_y", 0.0), pl.get("Point_z", 0.0)])
            d = np.array([pl.get("Normal_x", 0.0), pl.get("Normal_y", 0.0), pl.get("Normal_z", 0.0)])
            d = d * (1.0 / np.sqrt(d @ d))
            dist = (X - x0[0]) * d[0] + (Y - x0[1]) * d[1] + (Z - x0[2]) * d[2]
            sel = dist > 0
        else:
            # DistanceFromPointToSphere (GeoTools/DistancePointToSphere.h):
            # |x - x0| - R, negative inside; Side = Interior by default
            if pl.get("Side", "Interior") != "Interior":
                raise NotImplementedError("Sphere Side = " + str(pl.get("Side")))
            x0 = (pl.get("Center_x", 0.0), pl.get("Center_y", 0.0), pl.get("Center_z", 0.0))
            dx, dy, dz = X - x0[0], Y - x0[1], Z - x0[2]
            dist = np.sqrt(dx * dx + dy * dy + dz * dz) - pl["Radius"]
            sel = dist < 0
            # the level set takes +dist for an interior sphere (:1490-1514)
            dist = -dist
        st, mid = state(pl["InitialState"])
        V[1:-1, 1:-1, 1:-1][sel] = st
        ID[1:-1, 1:-1, 1:-1][sel] = mid
        if ls and mid == s.ls_materialid:
            inner = Phi[1:-1, 1:-1, 1:-1]
            upd = np.abs(dist) < np.abs(inner)
            inner[upd] = -dist[upd]
    return g, mats, nmat, s, V, ID, Phi


def _axis(N, a0, a1):
    h = (a1 - a0) / N
    c = np.empty(N + 2)
    for i in range(N):
        c[i + 1] = a0 + 0.5 * h + i * h
    c[0] = c[1] - h
    c[N + 1] = c[N] + h
    return c


def probe_rows(deck, g, V, ID, Phi):
    """The deck's first line plot (Output.LinePlot[0]) sampled from a state in
    the reference's format: rows of coordinate, rho, u, v, w, p, T, ID, phi."""
    from m2c_amd import output
    lp = deck["Output"]["LinePlot[0]"]
    p0 = [lp.get("X0", 0.0), lp.get("Y0", 0.0), lp.get("Z0", 0.0)]
    p1 = [lp.get("Xmax", 0.0), lp.get("Ymax", 0.0), lp.get("Zmax", 0.0)]
    _, rows = output.line_plot_rows(g, V, np.asarray(ID, dtype=np.float64), [Phi], p0, p1,
                                    int(lp["NumberOfPoints"]))
    return np.array(rows)


_AXIS_KEYS = (("NumberOfCells%s", "XYZ"), ("%s0", "XYZ"), ("%smax", "XYZ"), ("BoundaryCondition%s0", "XYZ"),
              ("BoundaryCondition%smax", "XYZ"), ("Velocity%s", "XYZ"), ("Point_%s", "xyz"), ("Normal_%s", "xyz"),
              ("Center_%s", "xyz"))


def permute_deck(deck, perm):
    """The same deck with the coordinate axes relabelled: old axis a becomes
    axis perm[a] (mesh extents, cell counts, boundary conditions, velocity
    components, geometry points and normals, line-plot end points)."""
    import copy

    def walk(d):
        out = {}
        for key, val in d.items():
            new = key
            for pat, names in _AXIS_KEYS:
                for a in range(3):
                    if key == pat % names[a]:
                        new = pat % names[perm[a]]
            out[new] = walk(val) if isinstance(val, dict) else val
        return out

    return walk(copy.deepcopy(deck))


def permute_fields(V, ID, Phi, perm):
    """Fields of an axis-permuted run brought back to the original axes:
    array axes (k, j, i) and velocity components (1, 2, 3).  perm is the map
    given to permute_deck."""
    # array axis of coordinate a is 2 - a; the permuted run stores old axis a
    # along array axis 2 - perm[a]
    order = [2 - perm[2 - ax] for ax in range(3)]
    Vb = np.transpose(V, order + [3]).copy()
    Vb[..., 1:4] = Vb[..., [1 + perm[a] for a in range(3)]]
    return Vb, np.transpose(ID, order).copy(), None if Phi is None else np.transpose(Phi, order).copy()
