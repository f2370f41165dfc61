"""GPU parity of the interface Riemann direction (MultiPhase.RiemannNormal
= LevelSet / Average, SpaceOperator.cpp:2186-2415) with the oracle, bit for
bit: the exact solver along general unit directions, and shock-bubble
residuals and steps with the level-set normals."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi, configs
from test_gpu_multiphase import _random_riemann

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["sg", "nasg", "mg"])
def test_riemann_batch_general_directions_bitwise(kind):
    from m2c_amd import runtime
    mats, _, Vm, idm, Vp, idp = _random_riemann(1500, 23, kind)
    rng = np.random.default_rng(5)
    dirs = rng.normal(size=(len(Vm), 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    s = abi.default_scheme()
    gpu = runtime.Solver(abi.grid((4, 4, 4)), mats, 2, s)
    g = gpu.riemann_batch(dirs, Vm, idm, Vp, idp)
    o = ol.riemann_batch(mats, 2, s, dirs, Vm, idm, Vp, idp)
    for key in ("Vs", "Vsm", "Vsp", "id", "err", "iters", "fevals"):
        a, b = g[key], o[key]
        bad = np.nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))[0]
        assert bad.size == 0, (kind, key, bad[:5])


def _interior(a):
    return a[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("rn", [abi.RNORMAL_LEVEL_SET, abi.RNORMAL_AVERAGE])
def test_shock_bubble_with_levelset_normals_bitwise(rn):
    from m2c_amd import runtime
    print("riemann normal", rn, flush=True)
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(20)
    # a moving bubble with a tangential slip: the interface problems depend
    # on the direction from the first step
    V[..., 1] += 30.0
    V[..., 2] -= 20.0
    V[ID == 1, 3] += 15.0
    s.riemann_normal = rn
    gpu = runtime.Solver(g, mats, nmat, s)
    gpu.set_state(V, ID, Phi)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    assert np.array_equal(_interior(gpu.compute_residual()), _interior(o.residual()))
    # the normals change the solution: compare with the mesh normal
    s2 = abi.Scheme.from_buffer_copy(s)
    s2.riemann_normal = abi.RNORMAL_MESH
    om = ol.Oracle(g, mats, nmat, s2)
    om.set_state(V, ID, Phi)
    assert not np.array_equal(_interior(om.residual()), _interior(o.residual()))
    c0g, c0o = gpu.counters()["riemann_calls"], o.counters()["riemann_calls"]
    for step in range(4):
        gpu.run_steps(1)
        o.run_steps(1)
        assert gpu.time()[:2] == o.time()[:2], step
        Vg, IDg, Pg = gpu.get_state()
        Vo, IDo, Po = o.get_state()
        assert np.array_equal(_interior(Vg), _interior(Vo)), step
        assert np.array_equal(_interior(IDg), _interior(IDo)), step
        assert np.array_equal(_interior(Pg), _interior(Po)), step
    assert gpu.counters()["riemann_calls"] - c0g == o.counters()["riemann_calls"] - c0o > 0


def test_levelset_normals_need_one_level_set_and_two_materials():
    from m2c_amd import runtime
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(8)
    s.riemann_normal = abi.RNORMAL_LEVEL_SET
    s.n_levelsets = 0
    with pytest.raises(runtime.M2CError):
        runtime.Solver(g, mats, nmat, s)
