"""GPU parity along every axis and in two dimensions.

* The reference's line-fixture decks relabelled to run along y and along z,
  stepped on the GPU and on the oracle: bitwise equal at every step (the
  oracle's y and z paths are tied to x by tests/test_oracle_axes.py).
* The reference's 2D explosion deck (Tests/ExplosionShyue98, 500 x 500):
  the first steps bitwise against the oracle, then the whole 326-step run on
  the GPU against the reference's line plot and log."""
import numpy as np
import pytest

import deck as dk
import oracle_lib as ol
import test_oracle_shyue98 as shy

pytestmark = pytest.mark.gpu
INNER = (slice(1, -1),) * 3


def _rt():
    from m2c_amd import runtime
    return runtime


@pytest.mark.parametrize("name", ["Test1", "Test4", "Test5", "Test8"])
@pytest.mark.parametrize("perm", [(1, 2, 0), (2, 0, 1)])
def test_line_deck_along_other_axes_on_gpu(fixtures, name, perm):
    d = dk.permute_deck(fixtures[name]["deck"], perm)
    g, mats, nmat, s, V, ID, Phi = dk.setup(d)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    gpu = _rt().Solver(g, mats, nmat, s)
    gpu.set_state(*o.get_state())
    nsteps = len(fixtures[name]["log_steps"])
    for step in range(nsteps):
        gpu.run_steps(1)
        o.run_steps(1)
        if step % 10 == 9 or step == nsteps - 1:
            for a, b, what in zip(gpu.get_state(), o.get_state(), ("V", "ID", "phi")):
                assert np.array_equal(a[INNER], b[INNER]), (name, perm, step, what)
    assert gpu.time() == o.time()


def test_shyue98_on_gpu():
    fx = shy.fixture()
    g, mats, nmat, s, o = shy.initial(fx)
    gpu = _rt().Solver(g, mats, nmat, s)
    gpu.set_state(*o.get_state())
    for step in range(3):
        gpu.run_steps(1)
        o.run_steps(1)
        for a, b, what in zip(gpu.get_state(), o.get_state(), ("V", "ID", "phi")):
            assert np.array_equal(a[INNER], b[INNER]), (step, what)
        assert gpu.time() == o.time()
    assert gpu.counters()["riemann_calls"] == o.counters()["riemann_calls"] > 0
    # the rest of the run on the GPU alone, against the reference's results
    steps = fx["log_steps"]
    dts = []
    while True:
        gpu.run_steps(1)
        t, dt, st = gpu.time()
        dts.append(dt)
        if t >= s.max_time or st > len(steps) + 2:
            break
    assert st == len(steps)
    ref = np.array([x[2] for x in steps[3:]])
    assert np.max(np.abs(np.array(dts[:-1]) - ref[:-1]) / ref[:-1]) < 6e-7
    shy.check_line(fx, g, *gpu.get_state())
