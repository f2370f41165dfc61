"""GPU parity of the narrow-band level set (LevelSetOperator.cpp:84-99,
152-196, 880-955; LevelSetReinitializer.cpp:181-300, 440-647, 1045-1133,
1459-2031) with the oracle, bit for bit: the band's construction (levels,
masks, bounds and the cut-off phi), the in-band reinitialization with each
first-layer treatment and boundary condition, its failsafe, the band built
by m2cx_initialize, and shock-bubble time steps with the band moved by the
reinitialization every few steps (FE, RK2, RK3)."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi, configs
import levelset_cases as lc

pytestmark = pytest.mark.gpu


def _used(shape):
    """interior cells and face ghosts (edge and corner ghosts are never read)"""
    out = np.zeros(shape, dtype=int)
    for a in range(3):
        o = np.zeros(shape[a], dtype=int)
        o[0] = o[-1] = 1
        sh = [1, 1, 1]
        sh[a] = shape[a]
        out = out + o.reshape(sh)
    return out <= 1


def _pair(g, mats, nmat, s, V, ID, phi, ls_bc=True):
    """oracle and GPU solver on the same state, both with the band built"""
    from m2c_amd import runtime
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, phi)
    if ls_bc:
        o.ls_apply_bc()
    V0, ID0, P0 = o.get_state()
    o.construct_narrow_band()
    gpu = runtime.Solver(g, mats, nmat, s)
    gpu.set_state(V0, ID0, P0)
    gpu.construct_narrow_band()
    return gpu, o


def _same_band(gpu, o, what):
    lg, ug, ag, bg = gpu.narrow_band()
    lo, uo, ao, bo = o.narrow_band()
    assert np.array_equal(lg, lo), (what, "level", np.argwhere(lg != lo)[:5].tolist())
    assert np.array_equal(ug, uo) and np.array_equal(ag, ao), (what, "masks")
    assert bg == bo, (what, "bounds", bg, bo)


def _same(gpu, o, what):
    Vg, IDg, Pg = gpu.get_state()
    Vo, IDo, Po = o.get_state()
    u = _used(Pg.shape)
    bad = np.argwhere(u & (Pg != Po))
    assert np.array_equal(Pg[u], Po[u]), (what, "phi", np.abs(Pg[u] - Po[u]).max(), bad[:5].tolist())
    assert np.array_equal(IDg[u], IDo[u]), (what, "ID")
    assert np.array_equal(Vg[u], Vo[u]), (what, "V", np.abs(Vg[u] - Vo[u]).max())


def _sphere(n=(22, 18, 16), bw=4, bc=abi.LSBC_LINEAR_EXTRAPOLATION, flt=abi.REINIT_FIXED, maxits=200, freq=1):
    """off-centre sphere that crosses the x = 0 boundary, distance stretched
    by a smooth factor (same zero level set, |grad phi| far from 1)"""
    g = abi.grid(n, (0, 0, 0), (2.2, 1.8, 1.6))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, reinit_frequency=freq, reinit_maxits=maxits, reinit_tol=1e-6,
                           reinit_first_layer=flt, ls_bandwidth=bw)
    for q in range(6):
        s.ls_bc[q] = bc
    x = np.array(configs._axis(n[0], 0, 2.2))[None, None, :]
    y = np.array(configs._axis(n[1], 0, 1.8))[None, :, None]
    z = np.array(configs._axis(n[2], 0, 1.6))[:, None, None]
    d = np.sqrt((x - 0.3) ** 2 + (y - 0.9) ** 2 + (z - 0.8) ** 2) - 0.55
    phi = d * (1.6 + 0.5 * np.sin(3 * x + 2 * y) * np.cos(2 * z))
    V = np.zeros(phi.shape + (5,))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    ID = (phi < 0).astype(np.int32)
    return g, mats, 2, s, V, ID, phi


@pytest.mark.parametrize("bw", [2, 3, 6])
def test_construct_narrow_band_bitwise(bw):
    print("construct bw", bw, flush=True)
    gpu, o = _pair(*_sphere(bw=bw))
    _same_band(gpu, o, "construct")
    _same(gpu, o, "construct")
    lg, ug, ag, _ = gpu.narrow_band()
    assert 0 < ag.sum() < ug.sum() < ug.size


def test_construct_without_interface():
    """no sign change: an empty band and phi cut off to +/- the diagonal"""
    g, mats, nmat, s, V, ID, phi = _sphere()
    phi = np.abs(phi) + 0.1
    gpu, o = _pair(g, mats, nmat, s, V, ID, phi)
    _same_band(gpu, o, "empty")
    _same(gpu, o, "empty")
    assert not gpu.narrow_band()[1].any()


@pytest.mark.parametrize("bc", [abi.LSBC_LINEAR_EXTRAPOLATION, abi.LSBC_ZERO_NEUMANN, abi.LSBC_NON_NEGATIVE])
@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_CONSTRAINED1, abi.REINIT_CONSTRAINED2,
                                 abi.REINIT_ITERATIVE_CONSTRAINED1, abi.REINIT_ITERATIVE_CONSTRAINED2])
def test_reinitialize_in_band_bitwise(flt, bc):
    print("reinitialize in band flt", flt, "bc", bc, flush=True)
    gpu, o = _pair(*_sphere(bw=5, bc=bc, flt=flt, freq=2, maxits=60))
    gpu.reinitialize()
    o.reinitialize()
    _same(gpu, o, "reinit")
    _same_band(gpu, o, "reinit")
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] == 1
    assert cg["reinit_iterations"] == co["reinit_iterations"] > 3
    # a second call moves the band from where the first left it
    gpu.reinitialize(special_maxits=4)
    o.reinitialize(special_maxits=4)
    _same(gpu, o, "second call")
    _same_band(gpu, o, "second call")


def test_reinitialize_in_band_failsafe_bitwise():
    """The band edge that cannot converge (a ghost filled from outside the
    band): MaxIts reached, the CFL lowered by 1.5 per retry, on both."""
    g, mats, nmat, s, V, ID, phi, d = lc.doubled_plane(flt=abi.REINIT_FIXED)
    s.ls_bandwidth = 6
    s.reinit_frequency = 2
    s.reinit_tol = 1.0e-4
    s.reinit_maxits = 30
    gpu, o = _pair(g, mats, nmat, s, V, ID, phi)
    gpu.reinitialize()
    o.reinitialize()
    _same(gpu, o, "failsafe")
    assert gpu.counters()["reinit_iterations"] == o.counters()["reinit_iterations"] > 30


def test_initialize_builds_the_band():
    from m2c_amd import runtime
    n = 24
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(n)
    s.ls_bandwidth = 4
    s.reinit_frequency = 2
    st, sid, geoms = configs.shock_bubble_geometries(1.0, 0.15, 1.5, 0.2, configs.WATER_SG)
    gpu = runtime.Solver(g, mats, nmat, s)
    gpu.initialize(st, sid, geoms)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    o.ls_apply_bc()
    o.construct_narrow_band()
    _same_band(gpu, o, "initialize")
    _same(gpu, o, "initialize")


@pytest.mark.parametrize("ti,bw,freq", [(abi.TI_RK3, 4, 2), (abi.TI_RK2, 5, 3), (abi.TI_FE, 3, 1)])
def test_band_steps_bitwise(ti, bw, freq):
    """Shock-bubble steps with the band: FD3 advection on the band, the band
    moved by each reinitialization, Phi_tmp copied on the band after the
    first step."""
    print("band steps ti", ti, "bw", bw, "freq", freq, flush=True)
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(20, shock_x=0.34)  # shock at the bubble
    s.integrator = ti
    s.ls_bandwidth = bw
    s.reinit_frequency = freq
    s.reinit_maxits = 25
    s.reinit_first_layer = abi.REINIT_ITERATIVE_CONSTRAINED2
    if ti == abi.TI_FE:
        s.cfl = 0.2
    gpu, o = _pair(g, mats, nmat, s, V, ID, Phi, ls_bc=False)
    for step in range(7):
        gpu.run_steps(1)
        o.run_steps(1)
        assert gpu.time()[:2] == o.time()[:2], step
        _same(gpu, o, "step %d" % step)
        _same_band(gpu, o, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] == 7 // freq
    assert cg["reinit_iterations"] == co["reinit_iterations"]
    assert cg["id_changed_cells"] == co["id_changed"] > 0


def test_band_levelset_residual_bitwise():
    """ComputeResidualFDM_NarrowBand alone: cleared on the band, computed on
    its active nodes"""
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(18)
    s.ls_bandwidth = 4
    s.reinit_frequency = 2
    gpu, o = _pair(g, mats, nmat, s, V, ID, Phi, ls_bc=False)
    rg = gpu.levelset_residual()
    ro = o.ls_residual()
    useful = o.narrow_band()[1]
    assert np.array_equal(rg[useful], ro[useful])
    assert np.abs(ro[useful]).max() > 0


def test_band_settings_refused():
    from m2c_amd import runtime
    g, mats, nmat, s, V, ID, phi = _sphere(bw=4, freq=4)  # frequency not below the bandwidth
    with pytest.raises(runtime.M2CError):
        runtime.Solver(g, mats, nmat, s)
    g, mats, nmat, s, V, ID, phi = _sphere(bw=1, freq=1)  # bandwidth below 2
    with pytest.raises(runtime.M2CError):
        runtime.Solver(g, mats, nmat, s)
    g, mats, nmat, s, V, ID, phi = _sphere(bw=4, freq=2)
    gpu = runtime.Solver(g, mats, nmat, s)
    gpu.set_state(V, ID, phi)
    with pytest.raises(runtime.M2CError, match="narrow band"):
        gpu.run_steps(1)  # the band was never built
