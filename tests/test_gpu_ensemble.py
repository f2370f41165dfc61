"""Ensembles: members advanced in lockstep (one Riemann launch per RK stage
for all of them, one captured graph per step) give every member the state it
reaches when run alone, bit for bit."""
import numpy as np
import pytest

from m2c_amd import abi, configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ti", [abi.TI_RK2, abi.TI_RK3])
def test_ensemble_matches_members_run_alone(ti):
    from m2c_amd import runtime
    cases = configs.c4_instances(count=4, seed=7, edges=(12, 16, 20))
    alone, members = [], []
    for _, (g, mats, nmat, s, V, ID, Phi) in cases:
        s.integrator = ti
        a = runtime.Solver(g, mats, nmat, s)
        a.set_state(V, ID, Phi)
        a.run_steps(5)
        alone.append((a.get_state(), a.counters(), a.time()))
        m = runtime.Solver(g, mats, nmat, s)
        m.set_state(V, ID, Phi)
        members.append(m)
    ens = runtime.Ensemble(members)
    ens.run_steps(2)
    ens.run_steps(3)
    for i, m in enumerate(members):
        (Va, IDa, Pa), ca, ta = alone[i]
        Vb, IDb, Pb = m.get_state()
        cb = m.counters()
        assert np.array_equal(Vb[1:-1, 1:-1, 1:-1], Va[1:-1, 1:-1, 1:-1]), i
        assert np.array_equal(IDb, IDa), i
        assert np.array_equal(Pb[1:-1, 1:-1, 1:-1], Pa[1:-1, 1:-1, 1:-1]), i
        assert m.time() == ta, i
        for k in ("riemann_calls", "riemann_main_iters", "riemann_fevals", "riemann_errors", "steps"):
            assert cb[k] == ca[k], (i, k, cb[k], ca[k])
    ens.close()


def test_ensemble_rejects_forward_euler():
    from m2c_amd import runtime
    (_, (g, mats, nmat, s, V, ID, Phi)), = configs.c4_instances(count=1, seed=1, edges=(12,))
    s.integrator = abi.TI_FE
    m = runtime.Solver(g, mats, nmat, s)
    with pytest.raises(runtime.M2CError, match="error -7"):
        runtime.Ensemble([m])
