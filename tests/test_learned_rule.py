"""The learned-guess accuracy rule (tests/learned_rule.py) on the CPU: the
oracle's exact solver started from the network's predictions (the float64
restatement of the network in oracle/mlp_ref.py, with the shipped
parameters), against the plain start and the converged and fine solutions.
The GPU half of this check is tests/test_gpu_mlp.py."""
import os
import sys

import numpy as np

import learned_rule as lr
import oracle_lib as ol
from m2c_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref  # noqa: E402


def test_learned_guess_meets_the_accuracy_rule_on_the_oracle():
    mats = lr.water_air()
    prob = lr.problems(mats)
    theta = np.load(os.path.join(ROOT, "m2c_amd", "data", "guess_params.npy"))
    X, P, pcmin = lr.features(mats, *prob)
    y, _ = mlp_ref.forward(theta, X, bf16_mode=True)
    g = np.exp(np.asarray(y, dtype=np.float64)) * P - pcmin
    s = abi.default_scheme(riem_tol_main=lr.TOL)
    plain = ol.riemann_batch(mats, 2, s, *prob)
    learned = ol.riemann_batch(mats, 2, s, *prob, guess=g)
    conv, fine = lr.references(mats, prob)
    figs = lr.check(mats, prob, plain, learned, conv, fine)
    print(figs)
    # the guess is used, and it saves f(p) evaluations
    assert not np.array_equal(plain["iters"], learned["iters"])
    assert learned["fevals"].mean() < plain["fevals"].mean()


def test_rule_fails_for_a_sloppy_solve():
    """The rule is not vacuous: a solve stopped at tolerance 1e-6 breaks it."""
    mats = lr.water_air()
    prob = lr.problems(mats, n=1500, seed=98)
    s = abi.default_scheme(riem_tol_main=lr.TOL)
    plain = ol.riemann_batch(mats, 2, s, *prob)
    sloppy = ol.riemann_batch(mats, 2, abi.default_scheme(riem_tol_main=1e-6), *prob)
    conv, fine = lr.references(mats, prob)
    try:
        lr.check(mats, prob, plain, sloppy, conv, fine)
    except AssertionError:
        return
    raise AssertionError("a 1e-6 solve passed the 1e-10 rule")
