"""GPU parity of the finite-volume level set in a narrow band against the
oracle, bit for bit: the band-edge reads of the kept reconstruction buffers
on a line (zero before a first write, the last active call's values after),
the residual of a 3D band whose fluxes are summed in the reference's
useful-node order, and shock-bubble steps with the band updated at every
reinitialization (one level set, and two level sets with bands of their
own)."""
import numpy as np
import pytest

import oracle_lib as ol
import levelset_cases as lc
from m2c_amd import abi, configs

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


def _line_pair():
    from m2c_amd import runtime
    g = abi.grid((10, 1, 1), (0.0, 0.0, 0.0), (10.0, 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, ls_materialid=1, ls_solver=abi.LSSOLVER_FINITE_VOLUME,
                           ls_flux=abi.LSFLUX_ROE, reinit_frequency=1, ls_bandwidth=2)
    return runtime.Solver(g, mats, 2, s), ol.Oracle(g, mats, 2, s)


def _line_state(gpu, o, u, x0):
    x = np.broadcast_to(configs._axis(10, 0.0, 10.0)[None, None, :], o.shape)
    V = np.zeros(o.shape + (5,))
    V[..., 0] = 1.0
    V[..., 1] = u
    V[..., 4] = 1.0
    phi = x - x0
    ID = np.where(phi < 0, 1, 0).astype(np.int32)
    for s in (gpu, o):
        s.set_state(V, ID, phi)
        s.construct_narrow_band()


def test_band_edge_reads_bitwise():
    print("finite-volume band on a line", flush=True)
    gpu, o = _line_pair()
    _line_state(gpu, o, 1.0, 4.75)
    Rg, Ro = gpu.levelset_residual(), o.ls_residual()
    assert Rg[1, 1, 5] == Ro[1, 1, 5] == -0.4375 and Rg[1, 1, 6] == Ro[1, 1, 6] == -0.6875
    _line_state(gpu, o, -1.0, 3.75)
    Rg, Ro = gpu.levelset_residual(), o.ls_residual()
    assert Rg[1, 1, 4] == Ro[1, 1, 4] == 0.8125
    lg, ug, ag, _ = gpu.narrow_band()
    assert np.array_equal(Rg[ug.astype(bool)], Ro[ug.astype(bool)])


def test_band_residual_3d_bitwise():
    """the summation order follows the useful-node list: construction (index
    order, then the layers) and an update after a reinitialization (the old
    list's order, then the layers)"""
    from m2c_amd import runtime
    print("finite-volume band residual in 3D", flush=True)
    n = 16
    g = abi.grid((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, ls_materialid=1, ls_solver=abi.LSSOLVER_FINITE_VOLUME,
                           ls_flux=abi.LSFLUX_ROE, reinit_frequency=2, ls_bandwidth=5)
    a = np.array(configs._axis(n, 0.0, 1.0))
    Z, Y, X = np.meshgrid(a, a, a, indexing="ij")
    V = np.zeros((n + 2,) * 3 + (5,))
    V[..., 0] = 1.0
    V[..., 1] = np.sin(np.pi * X) * np.cos(np.pi * Y)
    V[..., 2] = -np.cos(np.pi * X) * np.sin(np.pi * Y) + 0.3 * Z
    V[..., 3] = 0.2 * np.sin(2 * np.pi * Y)
    V[..., 4] = 1.0
    phi = 1.3 * (np.sqrt((X - 0.45) ** 2 + (Y - 0.5) ** 2 + (Z - 0.55) ** 2) - 0.27)
    ID = np.where(phi < 0, 1, 0).astype(np.int32)
    gpu, o = runtime.Solver(g, mats, 2, s), ol.Oracle(g, mats, 2, s)
    for x in (gpu, o):
        x.set_state(V, ID, phi)
        x.construct_narrow_band()
    for rnd in range(2):
        Rg, Ro = gpu.levelset_residual(), o.ls_residual()
        lg, ug, ag, bg = gpu.narrow_band()
        lo, uo, ao, bo = o.narrow_band()
        assert np.array_equal(lg, lo) and bg == bo, rnd
        u = ug.astype(bool)
        assert np.array_equal(Rg[u], Ro[u]), (rnd, np.abs(Rg[u] - Ro[u]).max())
        # a reinitialization moves the band (UpdateNarrowBand)
        gpu.reinitialize()
        o.reinitialize()
        assert np.array_equal(gpu.get_state()[2], o.get_state()[2]), rnd


@pytest.mark.parametrize("ti", [abi.TI_FE, abi.TI_RK2, abi.TI_RK3])
def test_fv_band_steps_bitwise(ti):
    from m2c_amd import runtime
    print("finite-volume band steps, integrator", ti, flush=True)
    g, mats, nmat, s, V, ID, phis, ls, union = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM, integrator=ti,
                                                              reinit_frequency=2, bandwidth=4, shocks=(0.29, 1.71))
    mats = abi.materials(mats[0], mats[1])
    s.ls_solver = abi.LSSOLVER_FINITE_VOLUME
    ID = np.where(ID == 2, 1, ID).astype(ID.dtype)
    o = ol.Oracle(g, mats, 2, s)
    o.set_state(V, ID, union)
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    V0, ID0, P0 = o.get_state()
    o.construct_narrow_band()
    gpu = runtime.Solver(g, mats, 2, s)
    gpu.set_state(V0, ID0, P0)
    gpu.construct_narrow_band()
    u = _used(ID0.shape)
    for step in range(9):
        gpu.run_steps(1)
        o.run_steps(1)
        Vg, IDg, Pg = gpu.get_state()
        Vo, IDo, Po = o.get_state()
        assert np.array_equal(Pg[u], Po[u]), (step, np.abs(Pg[u] - Po[u]).max())
        assert np.array_equal(IDg[u], IDo[u]) and np.array_equal(Vg[u], Vo[u]), step
    assert gpu.counters()["reinit_calls"] == o.counters()["reinit_calls"] == 4
    assert gpu.counters()["id_changed_cells"] == o.counters()["id_changed"] > 0


def test_fv_bands_of_two_levelsets_bitwise():
    from m2c_amd import runtime
    print("finite-volume bands of two level sets", flush=True)
    g, mats, nmat, s, V, ID, phis, ls, _ = lc.two_bubbles(n=(32, 12, 12), mat2=HELIUM, rnormal=abi.RNORMAL_LEVEL_SET,
                                                          reinit_frequency=2, bandwidth=4, shocks=(0.29, 1.71))
    ls[1].bandwidth = 5
    for l in ls:
        l.solver = abi.LSSOLVER_FINITE_VOLUME
    ls[1].flux = abi.LSFLUX_UPWIND
    o = ol.Oracle(g, mats, nmat, s, levelsets=ls)
    o.set_state(V, ID, phis[0])
    o.set_levelset(1, phis[1])
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    V0, ID0, _ = o.get_state()
    P0 = [o.levelset(q) for q in range(2)]
    o.construct_narrow_band()
    gpu = runtime.Solver(g, mats, nmat, s, levelsets=ls)
    gpu.set_state(V0, ID0, P0[0])
    gpu.upload_levelset(1, P0[1])
    gpu.construct_narrow_band()
    u = _used(ID0.shape)
    for step in range(9):
        gpu.run_steps(1)
        o.run_steps(1)
        for q in range(2):
            assert np.array_equal(gpu.download_levelset(q)[u], o.levelset(q)[u]), (step, q)
        Vg, IDg, _ = gpu.get_state()
        Vo, IDo, _ = o.get_state()
        assert np.array_equal(IDg[u], IDo[u]) and np.array_equal(Vg[u], Vo[u]), step
    assert gpu.counters()["reinit_calls"] == o.counters()["reinit_calls"] == 8
