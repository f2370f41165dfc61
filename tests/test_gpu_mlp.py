"""The learned p* guess on the GPU: MFMA forward pass and training gradient
against the CPU oracle (oracle/mlp_ref.py), training, and the ablation of the
learned guess against the plain acoustic one on held-out problems and on the
reference's two-material Sod deck (Tests/RiemannProblems/Test1)."""
import os
import sys

import numpy as np
import pytest

from m2c_amd import abi, configs, guess

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def _solver(mats=None, tol=1e-10, learned=False):
    from m2c_amd import runtime
    if mats is None:
        mats = abi.materials(abi.material("SG", **configs.WATER_SG), abi.material("SG", **configs.AIR_SG))
    s = abi.default_scheme(riem_tol_main=tol, riem_guess=abi.GUESS_MLP if learned else abi.GUESS_ACOUSTIC)
    return runtime.Solver(abi.grid((4, 4, 4)), mats, 2, s), mats


def test_mfma_forward_matches_oracle():
    """Checks the MFMA operand layout and the whole forward pass: random
    parameters and inputs, GPU BF16/FP32 against the float64 oracle with the
    same BF16 roundings."""
    sol, _ = _solver()
    rng = np.random.default_rng(11)
    theta = mlp_ref.init_params(4)
    theta[512:576] = rng.normal(0, 0.2, 64)
    theta[4672:4736] = rng.normal(0, 0.2, 64)
    theta[8960] = -0.7
    sol.mlp_set_params(theta)
    X = rng.normal(size=(1000, 8)).astype(np.float32)
    y = sol.mlp_predict(X)
    yr, _ = mlp_ref.forward(theta, X, bf16_mode=True)
    scale = np.max(np.abs(yr))
    assert np.max(np.abs(y - yr)) <= 2e-2 * scale, np.max(np.abs(y - yr))
    # a transposed weight operand (the layout mistake this pins) would not come close
    bad = theta.copy()
    bad[576:4672] = theta[576:4672].reshape(64, 64).T.ravel()
    ybad, _ = mlp_ref.forward(bad, X, bf16_mode=True)
    assert np.max(np.abs(y - ybad)) > 1e-1 * scale


def test_training_gradient_matches_oracle():
    sol, _ = _solver()
    rng = np.random.default_rng(12)
    theta = mlp_ref.init_params(7)
    theta[4736 + 4096:4736 + 4160] = rng.normal(0, 0.1, 64)
    sol.mlp_set_params(theta)
    X = rng.normal(size=(4096, 8)).astype(np.float32)
    T = rng.normal(size=4096).astype(np.float32)
    loss, g = sol.mlp_loss_grad(X, T)
    lr, gr = mlp_ref.loss_grad(theta, X, T, bf16_mode=True)
    assert abs(loss - lr) <= 1e-3 * lr
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert rel <= 2e-2, rel
    for name, a, b in (("W1", 0, 512), ("W2", 576, 4672), ("W3", 4736, 8832), ("W4", 8896, 8960)):
        r = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        assert r <= 5e-2, (name, r)


def _air_pair():
    return abi.materials(abi.material("SG", gamma=1.4), abi.material("SG", gamma=1.4))


@pytest.fixture(scope="module")
def trained():
    """Water/air problems (the shock-bubble pair) and ideal-gas/ideal-gas
    problems (the two-material Sod pair), labelled by the plain FP64 solver."""
    sol, mats = _solver()
    X1, T1, _ = guess.labelled_features(sol, *guess.sample_problems(mats, 40000, seed=1))
    sol2, mats2 = _solver(_air_pair())
    X2, T2, _ = guess.labelled_features(sol2, *guess.sample_problems(mats2, 20000, seed=2, ref_density=(1.0, 1.0),
                                                                      p_level=(0.1, 10.0)))
    X, T = np.concatenate([X1, X2]), np.concatenate([T1, T2])
    hist = guess.train(sol, X, T, epochs=25, batch=2048, lr=2e-3, seed=0)
    return sol, mats, hist, sol.mlp_get_params()


def test_training_reduces_loss(trained):
    _, _, hist, _ = trained
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.2 * hist[0], hist


SHIPPED = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "m2c_amd", "data",
                       "guess_params.npy")


def test_learned_guess_ablation_held_out():
    """Held-out problems (a seed the training set does not use), with the
    shipped parameters: the learned guess must need fewer f(p) evaluations
    than the acoustic one, and the GPU's plain and learned results must both
    meet the accuracy rule of tests/learned_rule.py against the oracle's
    converged and fine solutions (shock-shock problems within 10 tol of the
    converged root; with a rarefaction, as accurate as the plain path against
    the fine solution)."""
    import learned_rule as lr
    theta = np.load(SHIPPED)
    mats = lr.water_air()
    plain, _ = _solver(mats, tol=lr.TOL)
    learned, _ = _solver(mats, tol=lr.TOL, learned=True)
    learned.mlp_set_params(theta)
    prob = lr.problems(mats)
    a = plain.riemann_batch(*prob)
    b = learned.riemann_batch(*prob)
    conv, fine = lr.references(mats, prob)
    figs = lr.check(mats, prob, a, b, conv, fine)
    ok = (a["err"] == 0) & (b["err"] == 0)
    it_a, it_b = a["iters"][ok].mean(), b["iters"][ok].mean()
    fe_a, fe_b = a["fevals"][ok].mean(), b["fevals"][ok].mean()
    print("held-out: %s" % figs)
    print("held-out: iterations/face plain %.2f learned %.2f; f-evals/face plain %.2f learned %.2f"
          % (it_a, it_b, fe_a, fe_b))
    assert fe_b < fe_a


def test_learned_guess_with_level_set_normal():
    """With the Riemann normal taken from the level set, the network is fed
    the problem along that normal (the one the solver solves): the guesses
    are accepted and the learned run follows the plain one.  The shock starts
    at the bubble, so that about 2,000 of the interface problems of these
    steps have acoustic wave strengths in the range where the guess is used."""
    from m2c_amd import runtime
    theta = np.load(SHIPPED)
    g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(20, radius=0.3, mach=2.5, shock_x=0.3)
    s.riemann_normal = abi.RNORMAL_LEVEL_SET
    out = {}
    for mode in (abi.GUESS_ACOUSTIC, abi.GUESS_MLP):
        s.riem_guess = mode
        sol = runtime.Solver(g, mats, nmat, s)
        if mode == abi.GUESS_MLP:
            sol.mlp_set_params(theta)
        sol.set_state(V, ID, Phi)
        sol.run_steps(6)
        out[mode] = (sol.get_state(), sol.counters())
    (Va, IDa, _), ca = out[abi.GUESS_ACOUSTIC]
    (Vb, IDb, _), cb = out[abi.GUESS_MLP]
    assert cb["riemann_calls"] == ca["riemann_calls"] > 0
    assert cb["riemann_guess_fallbacks"] == 0
    assert cb["riemann_fevals"] != ca["riemann_fevals"]  # the guesses were used
    inner = (slice(1, -1),) * 3
    assert np.sum(IDa[inner] != IDb[inner]) <= 4
    scale = np.max(np.abs(Va[inner]), axis=(0, 1, 2))
    assert np.all(np.max(np.abs(Vb[inner] - Va[inner]), axis=(0, 1, 2)) <= 1e-3 * scale)


def test_learned_guess_on_two_material_sod(fixtures, trained):
    """The reference's Test1 deck (two ideal gases, level set at x = 0.5):
    learned and plain runs agree to the Riemann tolerance; report the
    iterations per interface face of both."""
    import deck as dk
    from m2c_amd import runtime
    _, _, _, theta = trained
    import oracle_lib as ol
    d = fixtures["Test1"]["deck"]
    g, mats, nmat, s, V, ID, Phi = dk.setup(d)
    o = ol.Oracle(g, mats, nmat, s)   # boundary fill of the initial state
    o.set_state(V, ID, Phi)
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    V0, ID0, Phi0 = o.get_state()
    res = {}
    for mode in (abi.GUESS_ACOUSTIC, abi.GUESS_MLP):
        s.riem_guess = mode
        sol = runtime.Solver(g, mats, nmat, s)
        if mode == abi.GUESS_MLP:
            sol.mlp_set_params(theta)
        sol.set_state(V0, ID0, Phi0)
        sol.run_steps(len(fixtures["Test1"]["log_steps"]))
        res[mode] = (sol.get_state(), sol.counters())
    (Va, IDa, _), ca = res[abi.GUESS_ACOUSTIC]
    (Vb, IDb, _), cb = res[abi.GUESS_MLP]
    print("Test1: iterations/face plain %.2f learned %.2f (faces %d), guess fallbacks %d"
          % (ca["riemann_main_iters"] / ca["riemann_calls"], cb["riemann_main_iters"] / cb["riemann_calls"],
             cb["riemann_calls"], cb["riemann_guess_fallbacks"]))
    assert np.array_equal(IDa, IDb)
    rows = np.array(fixtures["Test1"]["line_rows"])
    assert np.max(np.abs(Vb[1, 1, 1:-1, 0] - rows[:, 1])) <= 1e-4 * np.max(rows[:, 1])
