"""Flux = Godunov (FluxFcnGodunov.h:37-60) on the GPU against the oracle, bit
for bit: the batched flux, a single-material 3D tube stepped with it, and a
two-material shock-bubble whose bulk faces use it, with the interface flux
Exact or Numerical (the numerical flux of each side against its star state
is then itself a Godunov flux)."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi, configs

pytestmark = pytest.mark.gpu


def _rt():
    from m2c_amd import runtime
    return runtime


@pytest.mark.parametrize("kind", ["air", "nasg"])
def test_godunov_flux_batch_bitwise(kind):
    rng = np.random.default_rng(5)
    n = 600
    if kind == "air":
        mats = abi.materials(abi.material("SG", gamma=1.4))
        rho, p = 1.0, 1.0
    else:
        mats = abi.materials(abi.material("NASG", gamma=1.19, pc=6.2178e8, b=6.7212e-4, q=-1.177788e6))
        rho, p = 1000.0, 1e5
    Vm, Vp = np.zeros((n, 5)), np.zeros((n, 5))
    for V in (Vm, Vp):
        V[:, 0] = rho * np.exp(rng.uniform(-0.3, 0.3, n))
        V[:, 4] = p * np.exp(rng.uniform(-1.0, 2.0, n))
        V[:, 1:4] = rng.uniform(-0.5, 0.5, (n, 3)) * np.sqrt(V[:, 4] / V[:, 0])[:, None]
    Vp[:10] = Vm[:10]  # uniform states
    s = abi.default_scheme(flux=abi.FLUX_GODUNOV, riem_tol_main=1e-10)
    sol = _rt().Solver(abi.grid((4, 4, 4)), mats, 1, s)
    for d in range(3):
        Fg = sol.numerical_flux_batch(d, Vm, Vp, np.zeros(n))
        Fo = ol.flux_batch(mats, 1, s, d, Vm, Vp, np.zeros(n))
        assert np.array_equal(Fg, Fo), (d, np.max(np.abs(Fg - Fo)))
    sol.close()


def test_godunov_tube_steps_bitwise():
    g, mats, nmat, s, V, ID, Phi = configs.c1_sod((24, 8, 6))
    s.flux = abi.FLUX_GODUNOV
    s.riem_tol_main = 1e-10
    # transverse velocities so that the y and z faces carry flux
    z, y, x = np.meshgrid(np.arange(V.shape[0]), np.arange(V.shape[1]), np.arange(V.shape[2]), indexing="ij")
    V[..., 2] = 0.1 * np.sin(0.7 * x + 0.3 * z)
    V[..., 3] = -0.05 * np.cos(0.5 * y + 0.2 * x)
    gpu = _rt().Solver(g, mats, nmat, s)
    gpu.set_state(V, ID, None)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, None)
    for step in range(5):
        gpu.run_steps(1)
        o.run_steps(1)
        Vg = gpu.get_state()[0]
        Vo = o.get_state()[0]
        assert np.array_equal(Vg[1:-1, 1:-1, 1:-1], Vo[1:-1, 1:-1, 1:-1]), (step, np.max(np.abs(Vg - Vo)))
    gpu.close()


@pytest.mark.parametrize("mpflux", [abi.MPFLUX_EXACT, abi.MPFLUX_NUMERICAL])
def test_godunov_shock_bubble_steps_bitwise(mpflux):
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(16)
    s.flux = abi.FLUX_GODUNOV
    s.mp_flux = mpflux
    gpu = _rt().Solver(g, mats, nmat, s)
    gpu.set_state(V, ID, Phi)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    for step in range(3):
        gpu.run_steps(1)
        o.run_steps(1)
        Vg, IDg, Pg = gpu.get_state()
        Vo, IDo, Po = o.get_state()
        assert np.array_equal(IDg[1:-1, 1:-1, 1:-1], IDo[1:-1, 1:-1, 1:-1]), step
        assert np.array_equal(Pg[1:-1, 1:-1, 1:-1], Po[1:-1, 1:-1, 1:-1]), step
        assert np.array_equal(Vg[1:-1, 1:-1, 1:-1], Vo[1:-1, 1:-1, 1:-1]), (step, np.max(np.abs(Vg - Vo)))
    cg, co = gpu.counters(), o.counters()
    assert cg["riemann_calls"] == co["riemann_calls"] > 0
    assert cg["riemann_fevals"] == co["riemann_fevals"]
    gpu.close()
