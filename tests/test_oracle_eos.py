"""Closed-form pins of the oracle's equations of state (CPU).

Each EOS is written here once, as the published pressure law p(rho, e)
(stiffened gas; Noble-Abel stiffened gas, Le Metayer & Saurel 2016; the
Mie-Gruneisen form with a linear Us-up Hugoniot reference curve), in mpmath at
50 digits.  Everything else is derived from that one formula numerically:
e(rho, p) by a root solve, dp/drho at constant e and Gamma = (1/rho) dp/de at
constant rho by mpmath differentiation, and c^2 = dp/drho|_s =
dp/drho|_e + (p/rho^2) dp/de|_rho.  The oracle's values (VarFcnSG.h:64-108,
VarFcnNASG.h:55-86, VarFcnMG.h:70-99, VarFcnBase.h:174-179) must agree to a
few ulps.  For NASG the sound speed also has the textbook closed form
c^2 = gamma (p + p_inf) / (rho (1 - b rho)).
"""
import random

import mpmath as mp
import pytest

import oracle_lib as ol
from m2c_amd import abi

mp.mp.dps = 50

CASES = {
    "SG water": (abi.material("SG", gamma=6.12, pc=3.43e8), (800.0, 1200.0), (1e3, 1e9)),
    "SG air": (abi.material("SG", gamma=1.4), (0.1, 10.0), (1e3, 1e7)),
    "NASG water": (abi.material("NASG", gamma=1.19, pc=6.2178e8, b=6.7212e-4, q=-1.177788e6), (900.0, 1400.0),
                   (1e3, 5e9)),
    "NASG order one": (abi.material("NASG", gamma=1.6, pc=2.0, b=0.1, q=-0.5), (0.5, 5.0), (0.1, 10.0)),
    "MG copper": (abi.material("MG", rho0=8.96e-3, c0=3.933e6, Gamma0=1.99, s=1.5, e0=0.0), (8.0e-3, 1.1e-2),
                  (1e5, 1e11)),
    "MG order one": (abi.material("MG", rho0=1.0, c0=2.0, Gamma0=2.0, s=1.5, e0=0.3), (0.9, 1.3), (0.5, 5.0)),
}


def _pressure_law(m):
    g, pc, b, q = mp.mpf(m.gamma), mp.mpf(m.pc), mp.mpf(m.b), mp.mpf(m.q)
    r0, c0, G0, s, e0 = mp.mpf(m.rho0), mp.mpf(m.c0), mp.mpf(m.Gamma0), mp.mpf(m.s), mp.mpf(m.e0)
    if m.eos == abi.EOS_SG:
        return lambda rho, e: (g - 1) * rho * e - g * pc
    if m.eos == abi.EOS_NASG:
        return lambda rho, e: (g - 1) * (e - q) / (1 / rho - b) - g * pc

    def mg(rho, e):
        eta = 1 - r0 / rho
        return r0 * c0 ** 2 * eta * (1 - G0 * eta / 2) / (1 - s * eta) ** 2 + G0 * r0 * (e - e0)
    return mg


@pytest.mark.parametrize("name", sorted(CASES))
def test_eos_against_its_pressure_law(name):
    mat, (rlo, rhi), (plo, phi) = CASES[name]
    P = _pressure_law(mat)
    rnd = random.Random(hash(name) & 0xFFFF)
    worst = [0.0] * 5
    for _ in range(40):
        rho = rnd.uniform(rlo, rhi)
        p = plo * (phi / plo) ** rnd.random()
        e_o, p_o, dpdr_o, G_o, c2_o = ol.eos_eval(mat, rho, p)
        R = mp.mpf(rho)
        e = mp.findroot(lambda x: P(R, x) - p, mp.mpf(e_o))
        dpdr = mp.diff(lambda r: P(r, e), R)
        dpde = mp.diff(lambda x: P(R, x), e)
        G = dpde / R
        c2 = dpdr + P(R, e) / R ** 2 * dpde
        for q, (mine, ex) in enumerate(((e_o, e), (p_o, mp.mpf(p)), (dpdr_o, dpdr), (G_o, G), (c2_o, c2))):
            worst[q] = max(worst[q], float(abs((mp.mpf(mine) - ex) / ex)))
        if mat.eos == abi.EOS_NASG:
            c2_closed = mat.gamma * (mp.mpf(p) + mat.pc) / (R * (1 - mat.b * R))
            assert abs((c2_o - c2_closed) / c2_closed) < 1e-13
    # e and p lose digits to the cancellation of the pressure constants
    # (p + gamma p_c for water); the derivatives are ratios of well-conditioned terms
    assert worst[0] < 1e-12 and worst[1] < 1e-8 and max(worst[2:]) < 1e-12, (name, worst)
