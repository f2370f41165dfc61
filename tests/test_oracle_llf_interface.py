"""Pins of the oracle's two-material LLF flux (MultiPhase.Flux =
LocalLaxFriedrichs, FluxFcnLLF.h:119-185).  No reference test deck runs it,
so it is pinned by what the mathematics fixes:

* with two identical material tables it is the single-material LLF flux,
  which the reference's Test11 / Test12 line fixtures pin;
* at a stationary contact (equal pressures, zero velocity) it has the closed
  form (-a/2 (rho_p - rho_m), p, 0, 0, -a/2 (E_p - E_m)) with
  a = max(c_m, c_p) and E = rho e, each side from its own EOS (a flux that
  used one side's material for both fails it);
* mirror symmetry: swapping the sides and flipping the normal velocity
  negates the mass, tangential-momentum and energy fluxes and keeps the
  normal-momentum flux, bit for bit (a dropped or sign-flipped term fails)."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi


def _states(n, seed, rho, p, spread=0.5):
    rng = np.random.default_rng(seed)
    V = np.zeros((n, 5))
    V[:, 0] = rho * np.exp(rng.uniform(-spread, spread, n))
    V[:, 4] = p * np.exp(rng.uniform(-1.0, 2.0, n))
    V[:, 1:4] = rng.uniform(-1.0, 1.0, (n, 3)) * np.sqrt(V[:, 4] / V[:, 0])[:, None]
    return V


@pytest.mark.parametrize("d", [0, 1, 2])
def test_identical_tables_give_the_single_material_flux(d):
    air = abi.material("SG", gamma=1.4)
    mats = abi.materials(air, abi.material("SG", gamma=1.4))
    Vm, Vp = _states(200, 1 + d, 1.0, 1.0), _states(200, 11 + d, 0.5, 0.8)
    n = len(Vm)
    F2 = ol.llf_interface_batch(mats, 2, d, Vm, np.zeros(n), Vp, np.ones(n))
    s = abi.default_scheme(flux=abi.FLUX_LLF)
    F1 = ol.flux_batch(abi.materials(air), 1, s, d, Vm, Vp, np.zeros(n))
    assert np.array_equal(F1, F2)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_stationary_contact_closed_form(d):
    g1, pc1, g2, pc2 = 6.12, 3.43e8, 1.4, 0.0
    mats = abi.materials(abi.material("SG", gamma=g1, pc=pc1), abi.material("SG", gamma=g2, pc=pc2))
    p = np.array([1e5, 3e6, 2e7])
    rm, rp = np.array([1000.0, 1020.0, 990.0]), np.array([1.2, 30.0, 150.0])
    n = len(p)
    Vm = np.zeros((n, 5)); Vm[:, 0] = rm; Vm[:, 4] = p
    Vp = np.zeros((n, 5)); Vp[:, 0] = rp; Vp[:, 4] = p
    F = ol.llf_interface_batch(mats, 2, d, Vm, np.zeros(n), Vp, np.ones(n))
    a = np.maximum(np.sqrt(g1 * (p + pc1) / rm), np.sqrt(g2 * (p + pc2) / rp))
    Em, Ep = (p + g1 * pc1) / (g1 - 1.0), (p + g2 * pc2) / (g2 - 1.0)
    want = np.zeros((n, 5))
    want[:, 0] = -0.5 * a * (rp - rm)
    want[:, 1 + d] = p
    want[:, 4] = -0.5 * a * (Ep - Em)
    scale = np.abs(want).max(axis=1, keepdims=True)
    assert np.all(np.abs(F - want) <= 1e-13 * scale)
    # the same contact with the air table on both sides is far off
    mats_wrong = abi.materials(abi.material("SG", gamma=g2, pc=pc2), abi.material("SG", gamma=g2, pc=pc2))
    Fw = ol.llf_interface_batch(mats_wrong, 2, d, Vm, np.zeros(n), Vp, np.ones(n))
    assert np.all(np.abs(Fw[:, 4] - want[:, 4]) > 1e-3 * np.abs(want[:, 4]))


@pytest.mark.parametrize("d", [0, 1, 2])
def test_mirror_symmetry_bitwise(d):
    mats = abi.materials(abi.material("NASG", gamma=1.19, pc=6.2178e8, b=6.7212e-4, q=-1.177788e6),
                         abi.material("SG", gamma=1.4))
    # water densities below 1/b = 1488 kg/m^3, where the NASG state is valid
    Vm, Vp = _states(300, 21 + d, 1000.0, 1e5, spread=0.3), _states(300, 31 + d, 1.2, 1e5)
    n = len(Vm)
    F = ol.llf_interface_batch(mats, 2, d, Vm, np.zeros(n), Vp, np.ones(n))
    assert np.all(np.isfinite(F))
    flip = lambda V: np.where(np.arange(5) == 1 + d, -V, V)
    Fm = ol.llf_interface_batch(mats, 2, d, flip(Vp), np.ones(n), flip(Vm), np.zeros(n))
    sign = -np.ones(5)
    sign[1 + d] = 1.0
    assert np.array_equal(Fm, sign * F)
