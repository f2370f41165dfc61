"""The accuracy rule of the learned p* guess, shared by its CPU and GPU tests.

The learned guess only moves the starting pressure of the exact solver; the
FP64 iteration after it keeps the reference's stopping rule
(ExactRiemannSolverBase.cpp:290-300: bracket width below tol * pscale or
|f(p)| below tol * cmax, pscale = max|p + rho u_n^2 / 2|, cmax = max c).  How
close its result is to the plain (acoustic-start) result depends on the wave
curves:

* Shock-shock problems.  Both wave curves are Hugoniot roots found by
  TOMS748 to 1e-12, the same function whichever pressures were tried before,
  so both starts converge to the same root within the stopping rule.  Rule:
  |dp*| <= 10 tol pscale, |du*| <= 10 tol cmax, |drho*| <= 10 tol rho* against
  the solution converged to tol 1e-13.
* Problems with a rarefaction.  The rarefaction curve is integrated by
  adaptive Runge-Kutta from a cached integration path (:1213-1420) that
  depends on the pressures evaluated before, to ToleranceRarefaction (1e-8)
  per step with the reference's Cash-Karp abscissa; f(p) is therefore known
  only to that integration error, and a different start ends at a different
  point of that error band.  Rule: against the fine solution (tol 1e-13,
  ToleranceRarefaction 1e-11), the learned path is as accurate as the plain
  one: the 99th percentile and the maximum of |dp*|/pscale and of
  |du*|/cmax are at most 1.25 times the plain path's plus 10 tol.
Measured on 4,000 held-out water/air problems: shock-shock learned against
converged |dp*|/pscale <= 4.7e-10, |du*|/cmax <= 2.2e-13 (tol 1e-10); with a
rarefaction both paths are within 3.7e-4 (p*) and 3.7e-4 (u*) of the fine
solution, the learned path's percentiles equal to the plain path's.
"""
import numpy as np

from m2c_amd import abi, configs, guess

TOL = 1e-10


def water_air():
    return abi.materials(abi.material("SG", **configs.WATER_SG), abi.material("SG", **configs.AIR_SG))


def problems(mats, n=4000, seed=99):
    """Held-out problems: a seed the training data does not use."""
    return guess.sample_problems(mats, n, seed=seed)


def features(mats, dirs, Vm, idm, Vp, idp):
    """The network's inputs (include/m2c_amd.h: densities and p + p_c scaled
    by the larger of each pair, normal velocities by sqrt(P/R), the two
    gammas) and the scales that turn its output into a pressure."""
    pc = np.array([mats[i].pc for i in range(2)])
    gam = np.array([mats[i].gamma for i in range(2)])
    ul = (Vm[:, 1:4] * dirs).sum(1)
    ur = (Vp[:, 1:4] * dirs).sum(1)
    ql, qr = Vm[:, 4] + pc[idm], Vp[:, 4] + pc[idp]
    P = np.maximum(ql, qr)
    R = np.maximum(Vm[:, 0], Vp[:, 0])
    C = np.sqrt(P / R)
    X = np.stack([np.log(Vm[:, 0] / R), np.log(Vp[:, 0] / R), ul / C, ur / C, np.log(ql / P), np.log(qr / P),
                  gam[idm], gam[idp]], 1).astype(np.float32)
    return X, P, np.minimum(pc[idm], pc[idp])


def scales(mats, dirs, Vm, idm, Vp, idp):
    unm, unp = (Vm[:, 1:4] * dirs).sum(1), (Vp[:, 1:4] * dirs).sum(1)
    pscale = np.maximum(np.abs(Vm[:, 4] + 0.5 * Vm[:, 0] * unm ** 2), np.abs(Vp[:, 4] + 0.5 * Vp[:, 0] * unp ** 2))
    cmax = np.maximum(guess.sound_speeds(mats, idm, Vm), guess.sound_speeds(mats, idp, Vp))
    return pscale, cmax


def references(mats, prob):
    """The converged (tol 1e-13) and fine (also ToleranceRarefaction 1e-11)
    solutions, on the oracle."""
    import oracle_lib as ol
    conv = ol.riemann_batch(mats, 2, abi.default_scheme(riem_tol_main=1e-13), *prob)
    fine_s = abi.default_scheme(riem_tol_main=1e-13)
    fine_s.riem_tol_rarefaction = 1e-11
    fine = ol.riemann_batch(mats, 2, fine_s, *prob)
    return conv, fine


def _dev(a, b, dirs, pscale, cmax, sel):
    dp = np.abs(a["Vsm"][:, 4] - b["Vsm"][:, 4])[sel] / pscale[sel]
    du = np.abs((a["Vsm"][:, 1:4] * dirs).sum(1) - (b["Vsm"][:, 1:4] * dirs).sum(1))[sel] / cmax[sel]
    dr = np.maximum(np.abs(a["Vsm"][:, 0] - b["Vsm"][:, 0]) / np.abs(b["Vsm"][:, 0]),
                    np.abs(a["Vsp"][:, 0] - b["Vsp"][:, 0]) / np.abs(b["Vsp"][:, 0]))[sel]
    return dp, du, dr


def check(mats, prob, plain, learned, conv, fine, tol=TOL):
    """Asserts the rule; returns the figures it was checked on."""
    dirs, Vm, idm, Vp, idp = prob
    pscale, cmax = scales(mats, *prob)
    ok = (plain["err"] == 0) & (learned["err"] == 0) & (conv["err"] == 0) & (fine["err"] == 0)
    assert ok.mean() > 0.99
    ps = conv["Vsm"][:, 4]
    shock2 = ok & (ps > np.maximum(Vm[:, 4], Vp[:, 4]))
    rare = ok & ~shock2
    assert shock2.sum() > 200 and rare.sum() > 200
    out = {}
    for name, run in (("plain", plain), ("learned", learned)):
        dp, du, dr = _dev(run, conv, dirs, pscale, cmax, shock2)
        assert dp.max() <= 10 * tol and du.max() <= 10 * tol and dr.max() <= 10 * tol, (name, dp.max(), du.max(),
                                                                                     dr.max())
        out[name + "_shock_shock"] = (dp.max(), du.max(), dr.max())
    pd, pu, _ = _dev(plain, fine, dirs, pscale, cmax, rare)
    ld, lu, _ = _dev(learned, fine, dirs, pscale, cmax, rare)
    for a, b, what in ((ld, pd, "p*"), (lu, pu, "u*")):
        for q in (0.99, 1.0):
            la, pa = np.quantile(a, q), np.quantile(b, q)
            assert la <= 1.25 * pa + 10 * tol, (what, q, la, pa)
    out["rarefaction_p_p99_max"] = (np.quantile(ld, 0.99), ld.max(), np.quantile(pd, 0.99), pd.max())
    out["rarefaction_u_p99_max"] = (np.quantile(lu, 0.99), lu.max(), np.quantile(pu, 0.99), pu.max())
    return out
