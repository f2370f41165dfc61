"""GPU parity of the level-set reinitialization (LevelSetReinitializer.cpp)
with the oracle, bit for bit: the direct call with each first-layer
treatment and boundary condition, its failsafe, reinitialization inside time
stepping at a step or time frequency, cells left to FixUnresolvedNodes, and
the fixed time step."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi, configs
import levelset_cases as lc

pytestmark = pytest.mark.gpu


def _used(shape):
    """Interior cells and face ghosts (edge and corner ghosts are not kept)."""
    out = np.zeros(shape, dtype=int)
    for a in range(3):
        o = np.zeros(shape[a], dtype=int)
        o[0] = o[-1] = 1
        sh = [1, 1, 1]
        sh[a] = shape[a]
        out = out + o.reshape(sh)
    return out <= 1


def _pair(g, mats, nmat, s, V, ID, phi, ls_bc=True):
    from m2c_amd import runtime
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, phi)
    if ls_bc:
        o.ls_apply_bc()
    V0, ID0, P0 = o.get_state()
    gpu = runtime.Solver(g, mats, nmat, s)
    gpu.set_state(V0, ID0, P0)
    return gpu, o


def _same(gpu, o, what):
    Vg, IDg, Pg = gpu.get_state()
    Vo, IDo, Po = o.get_state()
    u = _used(Pg.shape)
    assert np.array_equal(Pg[u], Po[u]), (what, "phi", np.abs(Pg[u] - Po[u]).max())
    assert np.array_equal(IDg[u], IDo[u]), (what, "ID")
    assert np.array_equal(Vg[u], Vo[u]), (what, "V", np.abs(Vg[u] - Vo[u]).max())


@pytest.mark.parametrize("bc", [abi.LSBC_LINEAR_EXTRAPOLATION, abi.LSBC_ZERO_NEUMANN, abi.LSBC_NON_NEGATIVE])
@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_CONSTRAINED1, abi.REINIT_CONSTRAINED2,
                                 abi.REINIT_ITERATIVE_CONSTRAINED1, abi.REINIT_ITERATIVE_CONSTRAINED2])
def test_reinitialize_bitwise(flt, bc):
    print("reinitialize flt", flt, "bc", bc, flush=True)
    g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere(flt=flt, bc=bc)
    gpu, o = _pair(g, mats, nmat, s, V, ID, phi)
    gpu.reinitialize()
    o.reinitialize()
    _same(gpu, o, "reinit")
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] == 1
    assert cg["reinit_iterations"] == co["reinit_iterations"] > 5


@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_ITERATIVE_CONSTRAINED2])
def test_reinitialize_failsafe_bitwise(flt):
    """Two iterations only: the CFL is lowered by 1.5 per retry on both."""
    print("failsafe flt", flt, flush=True)
    g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere(flt=flt, maxits=2)
    gpu, o = _pair(g, mats, nmat, s, V, ID, phi)
    gpu.reinitialize()
    o.reinitialize()
    _same(gpu, o, "failsafe")
    assert gpu.counters()["reinit_iterations"] == o.counters()["reinit_iterations"] >= 2 * 9
    # the lowered CFL stays: a second call starts from it on both
    gpu.reinitialize(special_maxits=3)
    o.reinitialize(special_maxits=3)
    _same(gpu, o, "second call")


def test_reinitialize_without_interface_does_nothing():
    g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere()
    phi = np.abs(phi) + 0.1
    gpu, o = _pair(g, mats, nmat, s, V, ID, phi)
    gpu.reinitialize()
    o.reinitialize()
    _same(gpu, o, "no interface")
    assert gpu.counters()["reinit_calls"] == 0 == o.counters()["reinit_calls"]


def test_reinitialize_needs_settings():
    from m2c_amd import runtime
    g, mats, nmat, s, V, ID, phi, d = lc.stretched_sphere()
    s.reinit_frequency = -1
    gpu = runtime.Solver(g, mats, nmat, s)
    gpu.set_state(V, ID, phi)
    with pytest.raises(runtime.M2CError, match="error -7"):
        gpu.reinitialize()


@pytest.mark.parametrize("flt", [abi.REINIT_FIXED, abi.REINIT_CONSTRAINED2, abi.REINIT_ITERATIVE_CONSTRAINED1])
def test_steps_with_reinitialization_bitwise(flt):
    """Shock-bubble steps with the level set reinitialized every 2 steps."""
    print("steps with reinit flt", flt, flush=True)
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(20)
    s.reinit_frequency = 2
    s.reinit_maxits = 25
    s.reinit_first_layer = flt
    gpu, o = _pair(g, mats, nmat, s, V, ID, Phi, ls_bc=False)
    for step in range(6):
        gpu.run_steps(1)
        o.run_steps(1)
        assert gpu.time()[:2] == o.time()[:2], step
        _same(gpu, o, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] == 3
    assert cg["reinit_iterations"] == co["reinit_iterations"]


def test_reinitialization_time_interval_and_fixed_time_step():
    """TimeInterval scheduling (floor(t/T) changes) with Time.TimeStep."""
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(16)
    o0 = ol.Oracle(g, mats, nmat, s)
    o0.set_state(V, ID, Phi)
    dt_cfl = o0.time_step()[0]
    s.time_step = 0.6 * dt_cfl
    s.reinit_frequency_dt = 1.5 * s.time_step
    s.reinit_maxits = 20
    gpu, o = _pair(g, mats, nmat, s, V, ID, Phi, ls_bc=False)
    gpu.run_steps(5)
    o.run_steps(5)
    assert gpu.time()[:2] == o.time()[:2]
    assert o.time()[1] == s.time_step
    _same(gpu, o, "time interval")
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] == 3  # floor(t / 1.5 dt) changes at steps 2, 3, 5


@pytest.mark.parametrize("reinit", [False, True])
def test_unresolved_cells_bitwise(reinit):
    print("unresolved cells, reinitializer", reinit, flush=True)
    g, mats, nmat, s, V, ID, phi = lc.unresolved_case(reinit)
    gpu, o = _pair(g, mats, nmat, s, V, ID, phi, ls_bc=False)
    for step in range(8):
        gpu.run_steps(1)
        o.run_steps(1)
        _same(gpu, o, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["unresolved_cells"] == co["unresolved_cells"] > 0
    assert cg["reinit_calls"] == co["reinit_calls"] == 0
