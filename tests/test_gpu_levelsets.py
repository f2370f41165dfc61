"""GPU parity with several level sets (m2cx_create_levelsets) against the
oracle, bit for bit: two bubbles of two materials, each tracked by its own
level set, with full-domain or narrow-band reinitialization, Mesh or
LevelSet normals; bubbles that overlap at the start (ResolveConflictsInLevelSets)
and touch (interface faces between two tracked materials, whose Riemann
normal uses both level sets); the initial condition built on the device;
the optional corrections of ResolveConflictsInLevelSets
(LevelSetCorrectionFrequency)."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi, configs
import levelset_cases as lc

pytestmark = pytest.mark.gpu

HELIUM = abi.material("SG", gamma=1.667, pc=0.0)


def _used(shape):
    out = np.zeros(shape, dtype=int)
    for a in range(3):
        o = np.zeros(shape[a], dtype=int)
        o[0] = o[-1] = 1
        sh = [1, 1, 1]
        sh[a] = shape[a]
        out = out + o.reshape(sh)
    return out <= 1


def _pair(g, mats, nmat, s, V, ID, phis, ls, band=False):
    from m2c_amd import runtime
    o = ol.Oracle(g, mats, nmat, s, levelsets=ls)
    o.set_state(V, ID, phis[0])
    for q in range(1, len(phis)):
        o.set_levelset(q, phis[q])
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    V0, ID0, _ = o.get_state()
    P0 = [o.levelset(q) for q in range(len(phis))]
    if band:
        o.construct_narrow_band()
    gpu = runtime.Solver(g, mats, nmat, s, levelsets=ls)
    gpu.set_state(V0, ID0, P0[0])
    for q in range(1, len(phis)):
        gpu.upload_levelset(q, P0[q])
    if band:
        gpu.construct_narrow_band()
    return gpu, o


def _same(gpu, o, nls, what):
    Vg, IDg, _ = gpu.get_state()
    Vo, IDo, _ = o.get_state()
    u = _used(IDg.shape)
    for q in range(nls):
        Pg, Po = gpu.download_levelset(q), o.levelset(q)
        assert np.array_equal(Pg[u], Po[u]), (what, "phi", q, np.abs(Pg[u] - Po[u]).max())
    assert np.array_equal(IDg[u], IDo[u]), (what, "ID")
    assert np.array_equal(Vg[u], Vo[u]), (what, "V", np.abs(Vg[u] - Vo[u]).max())


@pytest.mark.parametrize("rnormal,freq", [(abi.RNORMAL_MESH, 3), (abi.RNORMAL_LEVEL_SET, 2),
                                          (abi.RNORMAL_AVERAGE, -1)])
def test_two_bubbles_two_levelsets_bitwise(rnormal, freq):
    print("two level sets, normal", rnormal, "reinit every", freq, flush=True)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM, rnormal=rnormal,
                                                          reinit_frequency=freq, shocks=(0.29, 1.71),
                                                          first_layer=abi.REINIT_ITERATIVE_CONSTRAINED2)
    gpu, o = _pair(g, mats, nmat, s, V, ID, phis, ls)
    for step in range(10):
        gpu.run_steps(1)
        o.run_steps(1)
        assert gpu.time()[:2] == o.time()[:2], step
        _same(gpu, o, 2, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["id_changed_cells"] == co["id_changed"] > 0
    assert cg["reinit_calls"] == co["reinit_calls"]
    assert cg["reinit_iterations"] == co["reinit_iterations"]


def test_two_levelsets_narrow_bands_bitwise():
    print("two level sets in narrow bands", flush=True)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM,
                                                          rnormal=abi.RNORMAL_LEVEL_SET, reinit_frequency=2,
                                                          bandwidth=4, shocks=(0.29, 1.71))
    ls[1].bandwidth = 5  # each level set with its own band
    gpu, o = _pair(g, mats, nmat, s, V, ID, phis, ls, band=True)
    for step in range(9):  # the shocks reach the bubbles at the eighth step
        gpu.run_steps(1)
        o.run_steps(1)
        _same(gpu, o, 2, "step %d" % step)
    for q in range(2):
        lg, ug, ag, bg = gpu.narrow_band(q)
        lo, uo, ao, bo = o.narrow_band(q)
        assert np.array_equal(lg, lo) and bg == bo, q
    assert gpu.counters()["reinit_calls"] == o.counters()["reinit_calls"] == 8
    assert gpu.counters()["id_changed_cells"] == o.counters()["id_changed"] > 0


def test_overlapping_and_touching_bubbles_bitwise():
    """The bubbles overlap at the start: the first step's
    ResolveConflictsInLevelSets gives each overlap node to one level set and
    forces a reinitialization of both; afterwards faces between materials 1
    and 2 carry exact Riemann problems along the difference of the two
    level sets' normals."""
    print("overlapping bubbles", flush=True)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM,
                                                          centres=((0.8, 0.5, 0.5), (1.15, 0.5, 0.5)),
                                                          radii=(0.2, 0.2), rnormal=abi.RNORMAL_LEVEL_SET,
                                                          reinit_frequency=4)
    o_probe = ol.Oracle(g, mats, nmat, s, levelsets=ls)
    o_probe.set_levelset(0, phis[0])
    o_probe.set_levelset(1, phis[1])
    assert o_probe.resolve_conflicts() > 0  # the case does overlap
    gpu, o = _pair(g, mats, nmat, s, V, ID, phis, ls)
    for step in range(6):
        gpu.run_steps(1)
        o.run_steps(1)
        _same(gpu, o, 2, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] >= 2  # both forced after the first step
    assert cg["riemann_calls"] == co["riemann_calls"] > 0


def test_initialize_two_levelsets():
    from m2c_amd import runtime
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM)
    rho_w, p_w = 1000.0, 1.0e5
    rho2, u2, p2 = configs.sg_post_shock(rho_w, p_w, configs.WATER_SG["gamma"], configs.WATER_SG["pc"], 1.5)
    geoms = [abi.geometry("plane", (0.2, 0, 0), normal=(-1.0, 0, 0), materialid=0, state=(rho2, u2, 0, 0, p2)),
             abi.geometry("plane", (1.8, 0, 0), normal=(1.0, 0, 0), materialid=0, state=(rho2, -u2, 0, 0, p2)),
             abi.geometry("sphere", (1.5, 0.5, 0.5), radius=0.2, materialid=2, state=(1.2, 0, 0, 0, p_w)),
             abi.geometry("sphere", (0.5, 0.5, 0.5), radius=0.2, materialid=1, state=(1.2, 0, 0, 0, p_w))]
    gpu = runtime.Solver(g, mats, nmat, s, levelsets=ls)
    gpu.initialize((rho_w, 0.0, 0.0, 0.0, p_w), 0, geoms)
    o = ol.Oracle(g, mats, nmat, s, levelsets=ls)
    o.set_state(V, ID, phis[0])
    o.set_levelset(1, phis[1])
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    _same(gpu, o, 2, "initialize")


def test_levelset_settings_refused():
    from m2c_amd import runtime
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(16, 8, 8))
    ls[1].materialid = 1  # two level sets for one material
    with pytest.raises(runtime.M2CError):
        runtime.Solver(g, mats, nmat, s, levelsets=ls)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(16, 8, 8))
    ls[1].materialid = 0  # material 0 cannot have a level set
    with pytest.raises(runtime.M2CError):
        runtime.Solver(g, mats, nmat, s, levelsets=ls)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(16, 8, 8))
    s.mp_corrections_until_reinit = 0  # CorrectionsUntilReinitialization >= 1
    with pytest.raises(runtime.M2CError):
        runtime.Solver(g, mats, nmat, s, levelsets=ls)


def _planted(phis):
    """an isolated background node inside each bubble (part II resolves it)"""
    out = []
    for q, p in enumerate(phis):
        p = p.copy()
        p[np.unravel_index(np.argmin(p), p.shape)] = 0.01 * (q + 1)
        out.append(p)
    return out


@pytest.mark.parametrize("freq,until", [(1, 1), (2, 3)])
def test_levelset_corrections_bitwise(freq, until):
    """LevelSetCorrectionFrequency: part I with its optional branches and
    part II as plane sweeps on the device, against the oracle's sequential
    sweeps; the planted nodes are resolved at the first correction step, which
    forces a reinitialization of both level sets when CorrectionsUntil
    Reinitialization is 1."""
    print("corrections every", freq, "steps, reinit after", until, flush=True)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM, rnormal=abi.RNORMAL_LEVEL_SET,
                                                          reinit_frequency=4, shocks=(0.29, 1.71))
    s.mp_ls_correction_frequency = freq
    s.mp_corrections_until_reinit = until
    gpu, o = _pair(g, mats, nmat, s, V, ID, _planted(phis), ls)
    for step in range(10):
        gpu.run_steps(1)
        o.run_steps(1)
        _same(gpu, o, 2, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] == (6 if until == 1 else 4)
    assert cg["id_changed_cells"] == co["id_changed"] > 0


def test_levelset_corrections_overlap_bitwise():
    """Overlapping bubbles with the corrections on every step: part I's
    counted branch runs in the plane sweep, next to its optional branches."""
    print("overlapping bubbles with corrections", flush=True)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM,
                                                          centres=((0.8, 0.5, 0.5), (1.15, 0.5, 0.5)),
                                                          radii=(0.2, 0.2), rnormal=abi.RNORMAL_LEVEL_SET,
                                                          reinit_frequency=4)
    s.mp_ls_correction_frequency = 1
    gpu, o = _pair(g, mats, nmat, s, V, ID, phis, ls)
    for step in range(6):
        gpu.run_steps(1)
        o.run_steps(1)
        _same(gpu, o, 2, "step %d" % step)
    cg, co = gpu.counters(), o.counters()
    assert cg["reinit_calls"] == co["reinit_calls"] >= 2


def test_levelset_corrections_one_levelset_bitwise():
    """One level set: part I does not run, part II does (the reference calls
    ResolveConflictsInLevelSets for any number of level sets)."""
    from m2c_amd import runtime
    print("one level set with corrections", flush=True)
    g, mats, nmat, s, V, ID, phis, ls, union = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM, shocks=(0.29, 1.71))
    mats = abi.materials(mats[0], mats[1])
    s.mp_ls_correction_frequency = 3
    ID = np.where(ID == 2, 1, ID).astype(ID.dtype)
    phi = _planted([union])[0]

    def start(s):
        o = ol.Oracle(g, mats, 2, s)
        o.set_state(V, ID, phi)
        o.apply_bc()
        o.update_ghost_ids()
        o.ls_apply_bc()
        V0, ID0, P0 = o.get_state()
        gpu = runtime.Solver(g, mats, 2, s)
        gpu.set_state(V0, ID0, P0)
        return gpu, o

    # the step-3 correction forces a reinitialization: without a
    # reinitializer that is an error, as in the reference
    # (LevelSetOperator.cpp:1341-1345)
    gpu, o = start(s)
    o.run_steps(2)
    gpu.run_steps(2)
    with pytest.raises(RuntimeError, match="not specified"):
        o.run_steps(1)
    with pytest.raises(runtime.M2CError, match="not specified"):
        gpu.run_steps(1)
    s.reinit_frequency = 5
    gpu, o = start(s)
    u = _used(ID.shape)
    for step in range(6):
        gpu.run_steps(1)
        o.run_steps(1)
        Vg, IDg, Pg = gpu.get_state()
        Vo, IDo, Po = o.get_state()
        assert np.array_equal(Pg[u], Po[u]) and np.array_equal(IDg[u], IDo[u]) and np.array_equal(Vg[u], Vo[u]), step
    # forced at step 3, at its frequency at step 5
    assert gpu.counters()["reinit_calls"] == o.counters()["reinit_calls"] == 2


def test_vote_in_the_reference_order_bitwise():
    """two materials without a level set: the device's in-order vote against
    the oracle on the line whose result depends on the order"""
    from m2c_amd import runtime
    import test_oracle_levelsets as tol
    g, mats, s, V, ID, phi, phi_tmp = tol._vote_line()
    o = ol.Oracle(g, mats, 3, s)
    o.set_state(V, ID, phi)
    o.set_levelset(0, phi_tmp, tmp=True)
    o.update_material_ids()
    gpu = runtime.Solver(g, mats, 3, s)
    gpu.set_state(V, ID, phi)
    gpu.update_material_ids([phi_tmp])
    IDg, IDo = gpu.get_state()[1], o.get_state()[1]
    assert IDg[1, 1, 1:-1].tolist() == [1, 1, 1, 1, 0, 0, 0, 0]
    assert np.array_equal(IDg, IDo)
