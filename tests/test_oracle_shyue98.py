"""The reference's two-dimensional explosion deck (Tests/ExplosionShyue98:
500 x 500 x 1 cells, stiffened-gas liquid around an ideal-gas bubble, Roe
flux, conservative-characteristic GMM 1.2, level set FD3 with
LinearExtrapolation boundaries, PhaseChange = Extrapolation, RK3, CFL 0.4,
Farfield and Symmetry boundaries) replayed on the oracle (CPU).

This is the only multi-dimensional fixture of the path: the x and y face
fluxes, their accumulation and the 2D time step all enter.  The first steps'
dt are compared with log.out in the default suite; the full 326-step replay
(about 6 minutes on one core) with the line plot of results/line_0001.txt
runs when M2C_LONG_FIXTURES=1.  Full replay here: every dt within 2.8e-7 of
the printed values, line plot within 4e-9 (density), 8e-10 (u), 2.6e-7 (v,
which is nearly zero on the line), 1.8e-9 (pressure), 1.7e-9 (phi) of the
column maxima, and 1.5e-10 on the material-ID column (the probe interpolates
IDs between cell centres, so it is not integer next to the interface)."""
import json
import os

import numpy as np
import pytest

import deck as dk
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# column -> tolerance relative to max|column| (the files print 9 digits)
SHYUE_TOL = {1: 2e-8, 2: 5e-9, 3: 1e-6, 5: 1e-8, 7: 1e-9, 8: 1e-8}


def fixture():
    with open(os.path.join(ROOT, "tests", "golden", "shyue98.json")) as f:
        return json.load(f)


def initial(fx):
    g, mats, nmat, s, V, ID, Phi = dk.setup(fx["deck"])
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    return g, mats, nmat, s, o


def check_line(fx, g, V, ID, Phi):
    rows = np.array(fx["line_rows"])
    mine = dk.probe_rows(fx["deck"], g, V, ID, Phi)
    for col, tol in SHYUE_TOL.items():
        ref = rows[:, col]
        err = np.max(np.abs(mine[:, col] - ref))
        assert err <= tol * np.max(np.abs(ref)), (col, err / np.max(np.abs(ref)))


def test_shyue98_first_time_steps():
    fx = fixture()
    g, mats, nmat, s, o = initial(fx)
    assert (g.nx, g.ny, g.nz) == (500, 500, 1) and s.riem_tol_main == 1e-9
    ref = np.array([x[2] for x in fx["log_steps"][:8]])
    dts = []
    for _ in range(len(ref)):
        o.run_steps(1)
        dts.append(o.time()[1])
    assert np.max(np.abs(np.array(dts) - ref) / ref) < 6e-7  # %e prints 7 digits
    assert o.counters()["riemann_calls"] > 0


@pytest.mark.skipif(os.environ.get("M2C_LONG_FIXTURES") != "1",
                    reason="about 6 minutes on one core; set M2C_LONG_FIXTURES=1")
def test_shyue98_full_replay():
    fx = fixture()
    g, mats, nmat, s, o = initial(fx)
    steps = fx["log_steps"]
    dts = []
    for _ in range(len(steps) + 3):
        o.run_steps(1)
        t, dt, st = o.time()
        dts.append(dt)
        if t >= s.max_time:
            break
    assert st == len(steps)
    ref = np.array([x[2] for x in steps])
    assert np.max(np.abs(np.array(dts[:-1]) - ref[:-1]) / ref[:-1]) < 6e-7
    check_line(fx, g, *o.get_state())
