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


def _alone_and_members(cases, steps, prep=None):
    from m2c_amd import runtime
    alone, members = [], []
    for i, (_, (g, mats, nmat, s, V, ID, Phi)) in enumerate(cases):
        if prep:
            prep(i, s)
        a = runtime.Solver(g, mats, nmat, s)
        a.set_state(V, ID, Phi)
        rc = a.run_steps(steps)
        alone.append((a.get_state(), a.counters(), a.time(), rc))
        m = runtime.Solver(g, mats, nmat, s)
        m.set_state(V, ID, Phi)
        members.append(m)
    return alone, members


def _same_as_alone(alone, members):
    for i, m in enumerate(members):
        (Va, IDa, Pa), ca, ta, _ = alone[i]
        Vb, IDb, Pb = m.get_state()
        cb = m.counters()
        assert np.array_equal(Vb[1:-1, 1:-1, 1:-1], Va[1:-1, 1:-1, 1:-1]), i
        assert np.array_equal(IDb, IDa), i
        assert np.array_equal(Pb[1:-1, 1:-1, 1:-1], Pa[1:-1, 1:-1, 1:-1]), i
        assert m.time() == ta, (i, m.time(), ta)
        for k in ("riemann_calls", "riemann_main_iters", "riemann_fevals", "riemann_errors", "steps",
                  "clipped_cells", "id_changed_cells"):
            assert cb[k] == ca[k], (i, k, cb[k], ca[k])


def test_ensemble_member_with_llf_interface_flux():
    """A member whose interface fluxes are MultiPhase.Flux = LocalLaxFriedrichs
    solves no exact Riemann problems, in the ensemble as alone."""
    from m2c_amd import runtime
    cases = configs.c4_instances(count=3, seed=11, edges=(12, 16))

    def prep(i, s):
        if i == 1:
            s.mp_flux = abi.MPFLUX_LLF
    alone, members = _alone_and_members(cases, 4, prep)
    assert alone[1][1]["riemann_calls"] == 0 and alone[0][1]["riemann_calls"] > 0
    ens = runtime.Ensemble(members)
    ens.run_steps(4)
    _same_as_alone(alone, members)
    ens.close()


def test_ensemble_stops_each_member_at_max_time():
    """Members with MaxTime stop where m2cx_run_steps stops them (Main.cpp's
    t < MaxTime - 1e-6 dt), at different steps, while the others go on."""
    from m2c_amd import runtime
    cases = configs.c4_instances(count=3, seed=12, edges=(12, 16, 20))
    # one member runs unclamped; the others reach their MaxTime after about 2
    # and about 4 of their own steps
    probe = []
    for _, (g, mats, nmat, s, V, ID, Phi) in cases:
        a = runtime.Solver(g, mats, nmat, s)
        a.set_state(V, ID, Phi)
        a.run_steps(1)
        probe.append(a.time()[1])
        a.close()
    tmax = [-1.0, 2.5 * probe[1], 4.5 * probe[2]]

    def prep(i, s):
        s.max_time = tmax[i]
    alone, members = _alone_and_members(cases, 8, prep)
    steps = [c[1]["steps"] for c in alone]
    assert steps[0] == 8 and 2 <= steps[1] < 8 and 2 <= steps[2] < 8 and steps[1] != steps[2], steps
    ens = runtime.Ensemble(members)
    ens.run_steps(5)
    ens.run_steps(3)
    _same_as_alone(alone, members)
    ens.close()


def test_ensemble_with_learned_guess_matches_members_alone():
    import os
    from m2c_amd import runtime
    theta = np.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "m2c_amd", "data",
                                 "guess_params.npy"))
    cases = configs.c4_instances(count=3, seed=13, edges=(12, 16, 20))
    alone, members = [], []
    for _, (g, mats, nmat, s, V, ID, Phi) in cases:
        s.riem_guess = abi.GUESS_MLP
        for lst in (alone, members):
            sol = runtime.Solver(g, mats, nmat, s)
            sol.mlp_set_params(theta)
            sol.set_state(V, ID, Phi)
            if lst is alone:
                rc = sol.run_steps(4)
                lst.append((sol.get_state(), sol.counters(), sol.time(), rc))
            else:
                lst.append(sol)
    ens = runtime.Ensemble(members)
    rc = ens.run_steps(4)
    assert rc == sum(a[3] for a in alone)
    _same_as_alone(alone, members)
    ens.close()


def test_ensemble_rejects_several_levelsets_and_corrections():
    from m2c_amd import runtime
    import levelset_cases as lc
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(16, 8, 8), mat2=abi.material("SG", gamma=1.667))
    m = runtime.Solver(g, mats, nmat, s, levelsets=ls)
    with pytest.raises(runtime.M2CError, match="error -7"):
        runtime.Ensemble([m])
    (_, (g, mats, nmat, s, V, ID, Phi)), = configs.c4_instances(count=1, seed=1, edges=(12,))
    s.mp_ls_correction_frequency = 2
    m = runtime.Solver(g, mats, nmat, s)
    with pytest.raises(runtime.M2CError, match="error -7"):
        runtime.Ensemble([m])


def test_ensemble_at_the_c4_member_sizes():
    """members drawn from C4's own edge sizes (64, 96, 128 cells), with the
    learned guess as in the benchmark, 3 steps: each member bit for bit as
    run alone"""
    import os
    from m2c_amd import runtime
    print("ensemble at C4 member sizes", flush=True)
    theta = np.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "m2c_amd", "data",
                                 "guess_params.npy"))
    cases = configs.c4_instances(count=3, seed=5, edges=(64, 96, 128))
    alone, members = [], []
    for _, (g, mats, nmat, s, V, ID, Phi) in cases:
        s.riem_guess = abi.GUESS_MLP
        a = runtime.Solver(g, mats, nmat, s)
        a.mlp_set_params(theta)
        a.set_state(V, ID, Phi)
        rc = a.run_steps(3)
        alone.append((a.get_state(), a.counters(), a.time(), rc))
        a.close()
        m = runtime.Solver(g, mats, nmat, s)
        m.mlp_set_params(theta)
        m.set_state(V, ID, Phi)
        members.append(m)
    assert len({m.shape for m in members}) > 1  # uneven sizes
    ens = runtime.Ensemble(members)
    rc = ens.run_steps(3)
    assert rc == sum(a[3] for a in alone)
    _same_as_alone(alone, members)
    ens.close()
