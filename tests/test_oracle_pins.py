"""Pins of the CPU oracle (the restatement of the reference's explicit step)
against the reference's own fixtures, closed forms and exact arithmetic.
CPU only."""
import math
import random
from decimal import Decimal, getcontext

import numpy as np
import pytest

import deck as dk
import oracle_lib as ol
from m2c_amd import abi

SG14 = abi.materials(abi.material("SG", gamma=1.4), abi.material("SG", gamma=1.4))


# --------------------------------------------------------------------------
# correctly rounded roots used by the rarefaction step-size control
# --------------------------------------------------------------------------
def _cr_root_exact(x, n):
    getcontext().prec = 60
    r = Decimal(x) ** (Decimal(1) / Decimal(n))
    lo = float(r)
    # pick the double nearest to r among lo and its neighbours
    cands = [np.nextafter(lo, -np.inf), lo, np.nextafter(lo, np.inf)]
    return min(cands, key=lambda c: abs(Decimal(float(c)) - r))


@pytest.mark.parametrize("n", [4, 5])
def test_cr_root_is_correctly_rounded(n):
    rnd = random.Random(1234 + n)
    xs = [rnd.uniform(-40, 40) for _ in range(3000)]
    xs = [10.0 ** e for e in xs] + [1.0, 2.0, 16.0, 32.0, 1e-300, 1e300, 0.5, 3.0]
    bad = [(x, ol.cr_root(x, n), _cr_root_exact(x, n)) for x in xs if ol.cr_root(x, n) != _cr_root_exact(x, n)]
    assert not bad, bad[:5]


# --------------------------------------------------------------------------
# wave curves: shock branch vs the closed-form stiffened-gas Hugoniot
# --------------------------------------------------------------------------
@pytest.mark.parametrize("gam,pc,rho,p,ps", [(1.4, 0.0, 0.125, 0.1, 0.30313), (1.4, 0.0, 1.0, 1.0, 50.0),
                                             (6.12, 3.43e8, 1000.0, 1e5, 2.5e9), (4.4, 6e8, 1000.0, 1e9, 3e9)])
def test_shock_branch_matches_sg_hugoniot(gam, pc, rho, p, ps):
    mats = abi.materials(abi.material("SG", gamma=gam, pc=pc))
    s = abi.default_scheme()
    ok, rhos, us = ol.wave(mats, 1, s, 3, rho, 0.0, p, ps, 0, rho, rho * 1.1)
    assert ok
    mu = (gam - 1) / (gam + 1)
    r = (ps + pc) / (p + pc)
    rho_ex = rho * (r + mu) / (mu * r + 1)
    u_ex = math.sqrt((ps - p) * (1 / rho - 1 / rho_ex))
    assert abs(rhos - rho_ex) / rho_ex < 1e-11
    assert abs(us - u_ex) / u_ex < 1e-9


def test_shock_branch_matches_nasg_hugoniot():
    gam, pc, b, q = 1.19, 6.2178e8, 6.7212e-4, -1.177788e6
    mats = abi.materials(abi.material("NASG", gamma=gam, pc=pc, b=b, q=q))
    s = abi.default_scheme()
    rho, p, ps = 1000.0, 1e5, 3e9
    ok, rhos, us = ol.wave(mats, 1, s, 1, rho, 0.0, p, ps, 0, rho, rho * 1.1)
    assert ok
    # e(v,p) = (p + gam pc)(v - b)/(gam-1) + q makes the Hugoniot linear in v*
    v = 1.0 / rho
    g1 = gam - 1
    vs = ((ps + gam * pc) * b / g1 - (p + gam * pc) * (v - b) / g1 * -1 + 0.5 * (p + ps) * v) / \
         ((ps + gam * pc) / g1 + 0.5 * (p + ps))
    assert abs(1.0 / rhos - vs) / vs < 1e-11


# --------------------------------------------------------------------------
# rarefaction branch: with the published Cash-Karp abscissae the integrator
# reproduces the isentrope; the reference's stage-3 abscissa (p0 + dp/2,
# ExactRiemannSolverBase.cpp:1814) leaves an O(1e-3) defect, which the oracle
# keeps because the product must match the reference.
# --------------------------------------------------------------------------
def _sg_isentrope(gam, pc, rho, p, ps, wave=1):
    c = math.sqrt(gam * (p + pc) / rho)
    r = (ps + pc) / (p + pc)
    rhos = rho * r ** (1 / gam)
    du = 2 * c / (gam - 1) * (r ** ((gam - 1) / (2 * gam)) - 1)
    return rhos, (du if wave == 1 else -du) * -1.0


@pytest.mark.parametrize("gam,pc,rho,p,ps", [(1.4, 0.0, 1.0, 1.0, 0.30313), (1.4, 0.0, 1.0, 1.0, 0.01),
                                             (7.15, 2.8951e8, 1000.0, 1e7, 1.02583e5),
                                             (1.667, 0.0, 1.0, 500.0, 237.6)])
def test_rarefaction_branch_textbook_switch(gam, pc, rho, p, ps):
    mats = abi.materials(abi.material("SG", gamma=gam, pc=pc))
    s = abi.default_scheme()
    rho_ex, u_ex = _sg_isentrope(gam, pc, rho, p, ps)
    try:
        ol.set_ck_textbook(True)
        ok, rhos, us = ol.wave(mats, 1, s, 1, rho, 0.0, p, ps, 0, rho, rho * 0.9)
    finally:
        ol.set_ck_textbook(False)
    assert ok
    assert abs(rhos - rho_ex) / rho_ex < 2e-5
    assert abs(us - u_ex) / abs(u_ex) < 2e-5
    ok, rhos_ref, us_ref = ol.wave(mats, 1, s, 1, rho, 0.0, p, ps, 0, rho, rho * 0.9)
    assert ok
    # the reference's abscissa: a first-order defect, bounded
    assert 1e-6 < abs(rhos_ref - rho_ex) / rho_ex < 5e-2


# --------------------------------------------------------------------------
# exact two-material Riemann fixtures (Tests/RiemannProblems/*/exact)
# --------------------------------------------------------------------------
def _star_from_fixture(rows):
    """p*, u* and the two star densities: the rows next to the contact."""
    rows = np.asarray(rows)
    # FinalizeSolution pushes (u*, rho_R*, u*, p*) right after
    # (u* - du, rho_L*, u*, p*) (ExactRiemannSolverBase.cpp:370-373)
    k = next(k for k in range(1, len(rows)) if abs(rows[k, 0] - rows[k, 2]) <= 1e-6 * abs(rows[k, 2])
             and abs(rows[k - 1, 3] - rows[k, 3]) <= 1e-6 * abs(rows[k, 3]))
    left, right = rows[k - 1], rows[k]
    return left[3], left[2], left[1], right[1]


def _riemann_from_deck(fx, name, textbook):
    d = fx[name]["deck"]
    g, mats, nmat, s, *_ = dk.setup(d)
    bc = d["BoundaryConditions"]
    L, idl = dk.state(bc["Inlet"])
    R, idr = dk.state(bc["Inlet2"])
    try:
        ol.set_ck_textbook(textbook)
        out = ol.riemann_batch(mats, nmat, s, [[1.0, 0, 0]], [L], [idl], [R], [idr])
    finally:
        ol.set_ck_textbook(False)
    return out


@pytest.mark.parametrize("name", ["Test4", "Test8", "Test9", "Test10"])
def test_exact_fixture_shock_shock(fixtures, name):
    out = _riemann_from_deck(fixtures, name, False)
    ps, us, rl, rr = _star_from_fixture(fixtures[name]["exact_rows"])
    Vsm, Vsp = out["Vsm"][0], out["Vsp"][0]
    assert out["err"][0] == 0
    for mine, ref in [(Vsm[4], ps), (Vsm[1], us), (Vsm[0], rl), (Vsp[0], rr)]:
        assert abs(mine - ref) <= 3e-6 * abs(ref) + 1e-300, (name, mine, ref)


def _scales(fx, name):
    """Convergence scales of the secant loop: err_p is relative to
    max|p + rho u^2/2| and err_u to max(c) (ExactRiemannSolverBase.cpp:295-296)."""
    d = fx[name]["deck"]
    g, mats, nmat, s, *_ = dk.setup(d)
    bc = d["BoundaryConditions"]
    (rl, ul, _, _, pl), _ = dk.state(bc["Inlet"])
    (rr, ur, _, _, pr), _ = dk.state(bc["Inlet2"])
    out = ol.riemann_batch(mats, nmat, s, [[1.0, 0, 0]], [[rl, ul, 0, 0, pl]], [0], [[rr, ur, 0, 0, pr]], [1])
    return max(abs(pl + 0.5 * rl * ul * ul), abs(pr + 0.5 * rr * ur * ur))


@pytest.mark.parametrize("name", ["Test1", "Test2", "Test3", "Test5", "Test6", "Test11", "Test12"])
def test_exact_fixture_with_rarefaction(fixtures, name):
    ps, us, rl, rr = _star_from_fixture(fixtures[name]["exact_rows"])
    pscale = _scales(fixtures, name)
    out = _riemann_from_deck(fixtures, name, True)
    Vsm, Vsp = out["Vsm"][0], out["Vsp"][0]
    assert abs(Vsm[4] - ps) <= 3e-5 * pscale, (name, Vsm[4], ps)
    # next to vacuum (Test11/12) the loop stops on err_p while u*(p) is steep,
    # so u* is fixed only to ~|du/dp| * tol * pscale
    utol = 3e-5 if ps > 1e-3 * pscale else 1e-2
    assert abs(Vsm[1] - us) <= utol * abs(us), (name, Vsm[1], us)
    if ps > 1e-3 * pscale:  # star densities are ill-conditioned next to vacuum
        assert abs(Vsm[0] - rl) <= 3e-5 * abs(rl), (name, Vsm[0], rl)
        assert abs(Vsp[0] - rr) <= 3e-5 * abs(rr), (name, Vsp[0], rr)
    ref_mode = _riemann_from_deck(fixtures, name, False)
    assert abs(ref_mode["Vsm"][0][4] - ps) <= 1e-2 * pscale


# --------------------------------------------------------------------------
# full pipeline against the reference's line plots: RK2/RK3, level set FD3,
# exact interface Riemann, RiemannSolution phase change, Inlet/Inlet2/
# Symmetry boundaries, and the four reconstruction variable types:
# Primitive (Test1 van Albada, Test6 GMM 2.0), ConservativeCharacteristic
# (Test2, Test3 Roe GMM; Test8 HLLC van Albada with MG/MG; Test9, Test10 RK3
# MG/MG), Conservative (Test5 HLLC GMM 1.0).  The tolerance is a few times
# the deviation seen, relative to max|column|; the files print 9 digits.
# --------------------------------------------------------------------------
LINE_TOL = {"Test1": 5e-8, "Test2": 1.5e-7, "Test3": 2e-8, "Test4": 4e-8, "Test5": 1e-6, "Test6": 3e-9, "Test8": 6e-7,
            "Test9": 4e-8, "Test10": 8e-8}


def _replay(fx, name, nmax=None):
    d = fx[name]["deck"]
    g, mats, nmat, s, V, ID, Phi = dk.setup(d)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    steps = fx[name]["log_steps"]
    dts = []
    for _ in range(len(steps) + 3 if nmax is None else nmax):
        o.run_steps(1)
        t, dt, st = o.time()
        dts.append(dt)
        if t >= s.max_time:
            break
    return d, g, o, np.array(dts), st


@pytest.mark.parametrize("name", sorted(LINE_TOL))
def test_line_fixture_full_pipeline(fixtures, name):
    d, g, o, dts, st = _replay(fixtures, name)
    steps = fixtures[name]["log_steps"]
    assert st == len(steps)
    ref_dt = np.array([x[2] for x in steps])
    # the last step is clamped to MaxTime - t, a difference of printed times
    assert np.max(np.abs(dts[:-1] - ref_dt[:-1]) / ref_dt[:-1]) < 6e-7  # %e prints 7 digits
    Vf, IDf, Phif = o.get_state()
    rows = np.array(fixtures[name]["line_rows"])
    mine = dk.probe_rows(d, g, Vf, IDf, Phif)
    for col in (1, 2, 5, 7, 8):  # rho, u, p, material ID, phi
        ref = rows[:, col]
        assert np.max(np.abs(mine[:, col] - ref)) <= LINE_TOL[name] * np.max(np.abs(ref)), (name, col)


# PrimitiveCharacteristic (Test11 LLF with exact interface fluxes, Test12 LLF
# SG/MG) next to vacuum: the solid pulls away and its pressure sits on the
# cut-off, so clipping decisions amplify rounding differences (built with FMA
# contraction, the oracle's time steps already leave these logs by step 10).
# The time steps are pinned while the runs are still deterministic; the
# transform pair itself is pinned by the invariants below.
@pytest.mark.parametrize("name,nsteps", [("Test11", 11), ("Test12", 4)])
def test_primitive_characteristic_fixture_early_steps(fixtures, name, nsteps):
    d, g, o, dts, st = _replay(fixtures, name, nsteps)
    ref_dt = np.array([x[2] for x in fixtures[name]["log_steps"][:nsteps]])
    assert np.max(np.abs(dts - ref_dt) / ref_dt) < 6e-7


def _smooth_state(n, rho0, p0, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    x = np.linspace(0, 1, nx + 2)[None, None, :]
    y = np.linspace(0, 1, ny + 2)[None, :, None]
    z = np.linspace(0, 1, nz + 2)[:, None, None]
    V = np.zeros((nz + 2, ny + 2, nx + 2, 5))
    ph = rng.uniform(0, 1, 4)
    V[..., 0] = rho0 * (1.0 + 0.2 * np.sin(5 * x + 2 * y + ph[0]) * np.cos(3 * z))
    V[..., 1] = 0.3 * np.cos(4 * y + z + ph[1])
    V[..., 2] = -0.2 * np.sin(3 * x + 2 * z + ph[2])
    V[..., 3] = 0.25 * np.cos(6 * x * y + ph[3])
    V[..., 4] = p0 * (1.0 + 0.3 * np.cos(4 * z + x))
    return V


@pytest.mark.parametrize("eos", ["SG", "NASG", "MG"])
def test_characteristic_transforms_cancel_without_limiter(eos):
    """Without a limiter the slope is linear in the differences, so limiting
    in characteristic variables (L Q dV, then Q^-1 R sigma; L dU, then R
    sigma) must give the primitive (conservative) residual up to rounding:
    a wrong entry of L, R, Q or Q^-1 breaks the cancellation."""
    if eos == "SG":
        mat, rho0, p0 = abi.material("SG", gamma=1.4, pc=0.5), 1.0, 1.0
    elif eos == "NASG":  # order-one parameters keep L and R well conditioned
        mat, rho0, p0 = abi.material("NASG", gamma=1.19, pc=2.0, b=0.1, q=-0.5), 1.0, 1.0
    else:
        mat, rho0, p0 = abi.material("MG", rho0=1.0, c0=2.0, s=1.5, Gamma0=2.0), 1.05, 1.0
    n = (10, 8, 6)
    g = abi.grid(n, (0, 0, 0), (1.0, 0.8, 0.6))
    mats = abi.materials(mat)
    V = _smooth_state(n, rho0, p0, 5)
    ID = np.zeros(V.shape[:3], dtype=np.int32)
    res = {}
    for vt in range(4):
        s = abi.default_scheme(flux=abi.FLUX_HLLC, limiter=abi.LIM_NONE, var_type=vt)
        o = ol.Oracle(g, mats, 1, s)
        o.set_state(V, ID)
        o.apply_bc()
        res[vt] = o.residual()[1:-1, 1:-1, 1:-1]
    scale = np.max(np.abs(res[0]), axis=(0, 1, 2))
    for a, b in ((abi.VAR_PRIMITIVE_CHARACTERISTIC, abi.VAR_PRIMITIVE),
                 (abi.VAR_CONSERVATIVE_CHARACTERISTIC, abi.VAR_CONSERVATIVE)):
        err = np.max(np.abs(res[a] - res[b]), axis=(0, 1, 2)) / scale
        assert np.all(err < 1e-9), (eos, a, err)
    # and the variable types differ from each other (the test is not vacuous)
    assert np.max(np.abs(res[abi.VAR_CONSERVATIVE] - res[abi.VAR_PRIMITIVE])) > 1e-6 * np.max(scale)
