"""The oracle's narrow-band level set (LevelSetReinitializer.cpp:181-300,
440-647, 1045-1133, 1459-2031; LevelSetOperator.cpp:84-99, 152-196, 880-955)
pinned by the reference's own narrow-band results, and by properties of the
band for what they leave open.

Pins:
- Tests/LevelSetTests/*/NarrowBand: the field written at t = 0, after the band
  is built and phi is cut off outside it, bit for bit (SHA-256), and the final
  phi within 1e-12 of max|phi| (FD3 coefficients differ from the reference's
  QR solve in the last bit) with the reference's number of reinitializations:
  MergingCircles (Fixed first layer, 640 steps), SlottedDisk
  (IterativelyConstrainedMethod1, 2094 steps, ~80 s) and VortexDeformation
  (IterativelyConstrainedMethod2, 2667 steps, ~4 min, M2C_LONG_FIXTURES=1);
- in 3D, where the decks (one cell thick) do not reach: the band levels are
  one plus the face-neighbour (L1) distance to the nearest node of levels 0
  and 1, computed here as a brute-force distance transform; for a plane
  normal to an axis that distance has a closed form; edge and corner ghosts
  never join the band; phi outside the band is 10 x the extreme in-band
  value; useful = level <= bandwidth, active = level < bandwidth;
- the in-band reinitialization of twice a plane's distance returns the
  distance on the band's active nodes and keeps the zero level set.
"""
import hashlib
import os

import numpy as np
import pytest

import levelset_cases as lc
import oracle_lib as ol
from m2c_amd import abi, configs

LARGE = 31415927
_FIX = {}


def _fix(key):
    if not _FIX:
        with np.load(os.path.join(os.path.dirname(__file__), "golden", "levelset_tests.npz")) as f:
            _FIX.update({k: f[k] for k in f.files})
    return _FIX[key]


def _deck(name):
    g, mats, nmat, s, V, ID, Phi = lc.setup(name, narrow_band=True)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_ls_test(lc.CASES_NB[name]["test"])
    o.set_state(V, ID, Phi)
    o.ls_apply_bc()
    o.construct_narrow_band()
    return o


@pytest.mark.parametrize("name", sorted(lc.CASES_NB))
def test_narrow_band_deck_initial_field(name):
    o = _deck(name)
    P = o.get_state()[2][1:-1, 1:-1, 1:-1][0]
    assert hashlib.sha256(np.ascontiguousarray(P).tobytes()).digest() == _fix(name + "NB_initial_sha256").tobytes()


@pytest.mark.parametrize("name", [
    "MergingCircles", "SlottedDisk",
    pytest.param("VortexDeformation", marks=pytest.mark.skipif(
        os.environ.get("M2C_LONG_FIXTURES") != "1", reason="about 4 minutes; set M2C_LONG_FIXTURES=1"))])
def test_narrow_band_deck_final_phi(name):
    o = _deck(name)
    o.run_steps(1 << 30)
    assert o.time()[2] == int(_fix(name + "NB_steps"))
    assert o.counters()["reinit_calls"] == int(_fix(name + "NB_reinit_calls"))
    P = o.get_state()[2][1:-1, 1:-1, 1:-1][0]
    ref = _fix(name + "NB_final")
    err = np.abs(P - ref).max()
    assert err <= 1e-12 * np.abs(ref).max(), err


def _populated(shape):
    """ghosted brick without its edge and corner ghosts"""
    pz, py, px = shape
    out = np.zeros(shape, int)
    for ax, n in enumerate((pz, py, px)):
        idx = np.arange(n)
        o = ((idx == 0) | (idx == n - 1)).astype(int)
        out += o.reshape([-1 if a == ax else 1 for a in range(3)])
    return out < 2


def _seeds(phi, pop):
    """levels 0 and 1: phi = 0, or a zero or negative product with a
    populated face neighbour"""
    s = pop & (phi == 0)
    for ax in range(3):
        for sh in (1, -1):
            nb = np.roll(phi, sh, axis=ax)
            nbp = np.roll(pop, sh, axis=ax)
            edge = np.zeros(phi.shape, bool)
            sl = [slice(None)] * 3
            sl[ax] = 0 if sh == 1 else -1
            edge[tuple(sl)] = True
            s |= pop & nbp & ~edge & (phi * nb <= 0)
    return s


def _l1_distance(seeds, pop):
    """brute force: min over seeds of |di| + |dj| + |dk| (paths between
    populated nodes never need an edge or corner ghost)"""
    sz = np.argwhere(seeds)
    pts = np.argwhere(pop)
    d = np.full(len(pts), LARGE)
    for c in range(0, len(sz), 256):
        blk = sz[c:c + 256]
        dd = np.abs(pts[:, None, :] - blk[None, :, :]).sum(-1).min(1)
        d = np.minimum(d, dd)
    out = np.full(pop.shape, LARGE)
    out[tuple(pts.T)] = d
    return out


def _sphere_case(n=(18, 14, 12), bw=4, lsbc=abi.LSBC_LINEAR_EXTRAPOLATION):
    g = abi.grid(n, (0, 0, 0), (1.8, 1.4, 1.2))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, reinit_frequency=1, reinit_maxits=300, reinit_tol=1e-6, ls_bandwidth=bw)
    for q in range(6):
        s.ls_bc[q] = lsbc
    x = np.array(configs._axis(n[0], 0, 1.8))[None, None, :]
    y = np.array(configs._axis(n[1], 0, 1.4))[None, :, None]
    z = np.array(configs._axis(n[2], 0, 1.2))[:, None, None]
    # off-centre sphere touching the x = 0 boundary's ghosts
    phi = np.sqrt((x - 0.25) ** 2 + (y - 0.7) ** 2 + (z - 0.6) ** 2) - 0.45
    V = np.zeros(phi.shape + (5,))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    ID = (phi < 0).astype(np.int32)
    return g, mats, 2, s, V, ID, phi


@pytest.mark.parametrize("bw", [2, 4, 6])
def test_band_levels_are_the_distance_to_the_interface_nodes(bw):
    g, mats, nmat, s, V, ID, phi = _sphere_case(bw=bw)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, phi)
    o.ls_apply_bc()
    phi_bc = o.get_state()[2].copy()
    o.construct_narrow_band()
    level, useful, active, (pmax, pmin, pos, neg) = o.narrow_band()
    pop = _populated(phi.shape)
    seeds = _seeds(phi_bc, pop)
    dist = _l1_distance(seeds, pop)
    expect = np.where(dist + 1 <= bw, dist + 1, LARGE)
    expect[(phi_bc == 0) & pop] = 0
    # levels 0 and 1 are both seeds; compare "seed or distance + 1"
    got = np.where(level == 0, 1, level)
    exp = np.where(expect == 0, 1, expect)
    assert np.array_equal(got, exp)
    assert not level[~pop].any() or (level[~pop] == LARGE).all()
    assert np.array_equal(useful, level <= bw) and np.array_equal(active, level < bw)
    # cut-off: 10 x the extremes over the band
    assert pmax == phi_bc[useful].max() and pmin == phi_bc[useful].min()
    assert pos == 10 * pmax and neg == 10 * pmin
    P = o.get_state()[2]
    assert np.array_equal(P[useful], phi_bc[useful])
    assert np.array_equal(P[~useful], np.where(phi_bc[~useful] >= 0, pos, neg))


def test_band_of_a_plane_in_closed_form():
    """phi = x - x0 between two cell columns: levels 1 on both columns, then
    one more per column away from them, the same on every row and in the face
    ghosts of y and z."""
    n = (20, 6, 5)
    bw = 5
    g = abi.grid(n, (0, 0, 0), (2.0, 0.6, 0.5))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, reinit_frequency=1, ls_bandwidth=bw)
    x = np.array(configs._axis(n[0], 0, 2.0))
    phi = np.broadcast_to((x - 0.83)[None, None, :], (n[2] + 2, n[1] + 2, n[0] + 2)).copy()
    V = np.zeros(phi.shape + (5,))
    V[..., 0] = V[..., 4] = 1.0
    o = ol.Oracle(g, mats, 2, s)
    o.set_state(V, (phi < 0).astype(np.int32), phi)
    o.ls_apply_bc()
    o.construct_narrow_band()
    level = o.narrow_band()[0]
    assert len(x) == n[0] + 2  # ghosts included
    i0 = int(np.searchsorted(x, 0.83)) - 1  # last column with x < 0.83 (ghosted index)
    ii = np.arange(n[0] + 2)
    d = np.minimum(np.abs(ii - i0), np.abs(ii - (i0 + 1)))
    col = np.where(d + 1 <= bw, d + 1, LARGE)
    pop = _populated(phi.shape)
    exp = np.where(pop, col[None, None, :], LARGE)
    assert np.array_equal(level, exp)


@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_CONSTRAINED1, abi.REINIT_CONSTRAINED2,
                                 abi.REINIT_ITERATIVE_CONSTRAINED1, abi.REINIT_ITERATIVE_CONSTRAINED2])
def test_band_reinitialization_agrees_with_the_full_domain(flt):
    """Twice a tilted plane's distance, reinitialized in a band of 6 and over
    the full domain (a separate code path, pinned by its own fixtures): on the
    band's active nodes three or more cells from the boundary they agree to
    the convergence tolerance, in half the iterations.  Near the boundary the
    two paths differ by design: the full-domain tags count the unpopulated
    edge ghosts as zero (LevelSetReinitializer.cpp:322-370), the band's tags
    skip them (:466-512), which changes CR-1 / CR-2 at boundary nodes."""
    out, iters = {}, {}
    for bw in (6, 0):
        g, mats, nmat, s, V, ID, phi, d = lc.doubled_plane(flt=flt)
        s.reinit_tol = 1.0e-4
        for q in range(6):
            s.ls_bc[q] = abi.LSBC_ZERO_NEUMANN
        if bw:
            s.ls_bandwidth = bw
            s.reinit_frequency = 2
        o = ol.Oracle(g, mats, nmat, s)
        o.set_state(V, ID, phi)
        o.ls_apply_bc()
        if bw:
            o.construct_narrow_band()
            level, useful, active, _ = o.narrow_band()
        o.reinitialize()
        out[bw] = o.get_state()[2]
        iters[bw] = o.counters()["reinit_iterations"]
    h = 1.2 / 24
    core = np.zeros(phi.shape, bool)
    core[3:-3, 3:-3, 3:-3] = True
    m = active & core
    assert m.sum() > 1000
    assert np.abs(out[6] - out[0])[m].max() < 1.0e-4 * h
    assert np.array_equal(out[6][m] < 0, phi[m] < 0)
    assert iters[6] < iters[0]
    # the full-domain result itself: |grad phi| = 1 away from the first layer
    far = m & (level >= 3) & (level <= 4)
    gx, gy, gz = (np.gradient(out[0], axis=a) / h for a in (2, 1, 0))
    assert np.median(np.abs(np.sqrt(gx ** 2 + gy ** 2 + gz ** 2)[far] - 1)) < 0.05  # 1.0 before


def test_band_edge_ghost_filled_from_outside_the_band():
    """A reference quirk kept on purpose: a ghost can join the band through
    the ghost layer while the interior node two cells in is outside it;
    LinearExtrapolation then fills that ghost from the cut-off value
    (LevelSetReinitializer.cpp:1281-1300 reads phi[j-2] without a band test),
    and the band-edge node next to it keeps drifting, so the iterations hit
    MaxIts and the failsafe lowers the CFL."""
    g, mats, nmat, s, V, ID, phi, d = lc.doubled_plane(flt=abi.REINIT_FIXED)
    s.ls_bandwidth = 6
    s.reinit_frequency = 2
    s.reinit_tol = 1.0e-4
    s.reinit_maxits = 60
    o = ol.Oracle(g, mats, 2, s)
    o.set_state(V, ID, phi)
    o.ls_apply_bc()
    o.construct_narrow_band()
    level = o.narrow_band()[0]
    # the top y ghost at (i=0, k=3) is in the band, the interior nodes two
    # and three rows below it are not
    assert level[4, -1, 1] <= 6 and level[4, -3, 1] > 6
    o.reinitialize()
    P = o.get_state()[2]
    assert abs(P[4, -1, 1]) > 5 * abs(d[4, -1, 1])  # extrapolated from a cut-off value
    assert o.reinit_cfl() < 0.8 and o.counters()["reinit_iterations"] > 60
