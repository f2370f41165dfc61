"""The oracle's level-set reinitialization (LevelSetReinitializer.cpp) pinned
by the reference's own results, and by properties for what they leave open.

Pins:
- the full-domain level-set tests (Tests/LevelSetTests/*/FullDomain): the
  final phi of MergingCircles (Fixed first layer, 640 steps, 10
  reinitializations), SlottedDisk (Fixed, 2094 steps, 20) and
  VortexDeformation (IterativelyConstrainedMethod2, 2667 steps, 266; run with
  M2C_LONG_FIXTURES=1, ~5 min), within 1e-12 of max|phi| (the reference's
  FD3 coefficients come from a QR solve and differ from ours in the last bit);
- ConstrainedMethod1/2 and IterativelyConstrainedMethod1, which no fixture
  covers: a stretched distance function returns to a distance function with
  its zero level set in place (parity of their exact arithmetic is otherwise
  unpinned);
- the failsafe (CFL / 1.5 and restart until the last relative change is
  below 0.02), and that unresolved cells force no reinitialization with the
  default failsafe density.  The failsafe's error branch (CFL below 0.02
  and still no convergence) is not exercised: every pseudo-time update
  changes a node by at most about CFL times its own value, so with a CFL
  below 0.02 the last change is below 0.02 already.
"""
import hashlib
import os

import numpy as np
import pytest

import levelset_cases as lc
import oracle_lib as ol
from m2c_amd import abi

_FIX = {}


def _fix(key):
    """The converted reference results (loaded on first use: the file does not
    travel with the GPU tests)."""
    if not _FIX:
        with np.load(os.path.join(os.path.dirname(__file__), "golden", "levelset_tests.npz")) as f:
            _FIX.update({k: f[k] for k in f.files})
    return _FIX[key]


@pytest.mark.parametrize("name", sorted(lc.CASES))
def test_initial_conditions_match_reference_snapshot(name):
    phi0 = lc.initial_phi(name)
    assert hashlib.sha256(phi0[0].tobytes()).digest() == _fix(name + "_initial_sha256").tobytes()


def _replay(name):
    g, mats, nmat, s, V, ID, Phi = lc.setup(name)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_ls_test(lc.CASES[name]["test"])
    o.set_state(V, ID, Phi)
    o.ls_apply_bc()
    o.run_steps(1 << 30)
    return o


@pytest.mark.parametrize("name", [
    "MergingCircles", "SlottedDisk",
    pytest.param("VortexDeformation", marks=pytest.mark.skipif(
        os.environ.get("M2C_LONG_FIXTURES") != "1", reason="about 5 minutes; set M2C_LONG_FIXTURES=1"))])
def test_levelset_test_deck_final_phi(name):
    o = _replay(name)
    assert o.time()[2] == int(_fix(name + "_steps"))
    c = o.counters()
    assert c["reinit_calls"] == int(_fix(name + "_steps")) // lc.CASES[name]["freq"]
    _, _, P = o.get_state()
    ref = _fix(name + "_final")
    err = np.abs(P[1:-1, 1:-1, 1:-1][0] - ref).max()
    assert err <= 1e-12 * np.abs(ref).max(), err


@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_CONSTRAINED1, abi.REINIT_CONSTRAINED2,
                                 abi.REINIT_ITERATIVE_CONSTRAINED1, abi.REINIT_ITERATIVE_CONSTRAINED2])
def test_reinit_restores_distance_and_keeps_interface(flt):
    g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere(flt=flt)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, phi)
    o.ls_apply_bc()
    o.reinitialize()
    _, _, P = o.get_state()
    P, D, P0 = P[1:-1, 1:-1, 1:-1], d[1:-1, 1:-1, 1:-1], phi[1:-1, 1:-1, 1:-1]
    h = 1.4 / 28
    band, near = np.abs(D) < 4 * h, np.abs(D) < h
    err_band, err_near = np.abs(P - D)[band].max(), np.abs(P - D)[near].max()
    print("flt", flt, "error / h within 4h:", err_band / h, "within h:", err_near / h)
    # a distance function again near the interface (first order in h) ...
    assert err_band < 0.8 * h
    assert np.abs(P0 - D)[band].max() > 2 * h  # ... which it was not before
    # the first layer: kept as it was by Fixed, put in place by the others
    if flt == abi.REINIT_FIXED:
        assert err_near > 0.5 * h
    else:
        assert err_near < 0.15 * h
    # ... with the same inside and outside
    assert np.array_equal(P < 0, P0 < 0)
    c = o.counters()
    assert c["reinit_calls"] == 1 and c["reinit_iterations"] > 5


def test_first_layer_treatments_differ_where_they_should():
    """Fixed leaves the first layer alone; CR-1/CR-2 rescale it by |grad phi|
    before the iterations; HCR-1/HCR-2 move it during them."""
    out = {}
    for flt in range(5):
        g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere(flt=flt)
        o = ol.Oracle(g, mats, nmat, s)
        o.set_state(V, ID, phi)
        o.ls_apply_bc()
        o.reinitialize()
        out[flt] = o.get_state()[2]
    inner = (slice(1, -1),) * 3
    first = np.zeros(phi.shape, bool)
    p = phi
    for ax in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(1, None)
        sl_b[ax] = slice(None, -1)
        cross = p[tuple(sl_a)] * p[tuple(sl_b)] <= 0
        first[tuple(sl_a)] |= cross
        first[tuple(sl_b)] |= cross
    first[~np.pad(np.ones((20, 24, 28), bool), 1)] = False
    # Fixed: untouched but for the rounding of the RK3 combinations
    # (0.25 x + 0.75 x need not be x), which the reference has too
    assert np.allclose(out[0][first], phi[first], rtol=1e-14, atol=0)
    for flt in (1, 2, 3, 4):
        assert np.abs(out[flt][first] - phi[first]).max() > 1e-3, flt
    assert not np.array_equal(out[1][inner], out[2][inner])
    assert not np.array_equal(out[3][inner], out[4][inner])


def _cfl_after(retries):
    cfl = 0.8
    for _ in range(retries):
        cfl /= 1.5
    return cfl


def test_failsafe_lowers_cfl_and_accepts_small_changes():
    """Too few iterations: each retry divides the CFL by 1.5 and starts again
    from the saved phi, until the last relative change is below 0.02."""
    g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere(maxits=2)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, phi)
    o.ls_apply_bc()
    o.reinitialize()
    assert o.reinit_cfl() == _cfl_after(8)
    assert o.counters()["reinit_iterations"] == 2 * 9


def test_unresolved_cells_are_fixed_without_reinitialization():
    """FixUnresolvedNodes with the failsafe density (the reference's default)
    resolves every cell and reports none still unresolved
    (MultiPhaseOperator.cpp:1486-1525), so no forced reinitialization
    follows, with or without a reinitializer."""
    for reinit in (False, True):
        g, mats, nmat, s, V, ID, phi = lc.unresolved_case(reinit)
        o = ol.Oracle(g, mats, nmat, s)
        o.set_state(V, ID, phi)
        o.run_steps(8)
        c = o.counters()
        assert c["unresolved_cells"] > 0 and c["reinit_calls"] == 0
        _, IDo, _ = o.get_state()
        assert (IDo[1:-1, 1:-1, 1:-1] == 1).sum() > 0


def _first_layer(phi):
    """Interior nodes with a sign change to an axis neighbour."""
    first = np.zeros(phi.shape, bool)
    for ax in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(1, None)
        sl_b[ax] = slice(None, -1)
        cross = phi[tuple(sl_a)] * phi[tuple(sl_b)] <= 0
        first[tuple(sl_a)] |= cross
        first[tuple(sl_b)] |= cross
    inner = np.zeros(phi.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    return first & inner


@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_CONSTRAINED1, abi.REINIT_CONSTRAINED2])
def test_reinit_of_doubled_plane_distance(flt):
    """phi = 2 d for a tilted plane.  The first-layer rescaling phi/|grad phi|
    (LevelSetReinitializer.cpp:669-716) is exact on linear data, and both
    constrained averages (CR-1, CR-2: :784-833) then reproduce phi/2 exactly,
    so CR-1 and CR-2 put the first layer on d to rounding; the iterations then
    bring the band around it close to d.  Fixed keeps 2 d there, and the band
    follows it.  (HCR-1/HCR-2 move the first layer by a forcing term during
    the iterations, which has no such closed form.)"""
    g, mats, nmat, s, V, ID, phi, d = lc.doubled_plane(flt=flt)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, phi)
    o.ls_apply_bc()
    o.reinitialize()
    _, _, P = o.get_state()
    # nodes two cells or more from the domain boundary: next to it the
    # curvature test of the correction set and the first-layer differences
    # use one-sided stencils and extrapolated ghosts
    away = np.zeros(phi.shape, bool)
    away[3:-3, 3:-3, 3:-3] = True
    first = _first_layer(phi) & away
    h = 0.05
    band = (np.abs(d) < 4 * h) & away
    if flt == abi.REINIT_FIXED:
        assert np.allclose(P[first], phi[first], rtol=1e-14, atol=0)
        assert np.abs(P[band] - d[band]).max() > 0.1 * h
    else:
        assert np.abs(P[first] - d[first]).max() <= 1e-13 * h, flt
        # beyond the first layer: what the boundary stencils leave (0.04 h
        # seen), far below the 2 d it started from
        assert np.abs(P[band] - d[band]).max() <= 0.1 * h, flt
    assert np.array_equal(P[1:-1, 1:-1, 1:-1] < 0, phi[1:-1, 1:-1, 1:-1] < 0)
