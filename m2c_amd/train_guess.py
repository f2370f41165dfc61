"""Train the learned p* guess on the GPU and store its parameters.

    python -m m2c_amd.train_guess [--problems 2000000] [--epochs 40] [--out PATH]

Problems (seeded) are drawn for three material pairs: the stiffened-gas water /
ideal-gas air pair of the shock-bubble workloads, NASG water / air of the
bubble cloud and the ideal-gas pair of the two-material shock tube; the library's FP64 exact solver labels them, the
library's BF16-MFMA kernels train the network (Adam, mini-batches of 4096, the
rate halved three times over the second half); 5 % of the problems are held
out and scored.  The parameters are written as float32 (.npy, 8961 values).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from m2c_amd import abi, configs, guess  # noqa: E402

DEFAULT_OUT = os.path.join(ROOT, "m2c_amd", "data", "guess_params.npy")


def solver_for(mats):
    from m2c_amd import runtime
    s = abi.default_scheme(riem_tol_main=1e-10)
    return runtime.Solver(abi.grid((4, 4, 4)), mats, 2, s)


def build_training_set(n, seed):
    """n labelled problems: half stiffened-gas water / air, a quarter NASG
    water / air (the bubble-cloud pair), a quarter air / air."""
    wa = abi.materials(abi.material("SG", **configs.WATER_SG), abi.material("SG", **configs.AIR_SG))
    na = abi.materials(abi.material("NASG", **configs.WATER_NASG), abi.material("SG", **configs.AIR_SG))
    aa = abi.materials(abi.material("SG", gamma=1.4), abi.material("SG", gamma=1.4))
    sol_wa, sol_na, sol_aa = solver_for(wa), solver_for(na), solver_for(aa)
    parts = [(sol_wa, wa, n // 2, {}), (sol_na, na, n // 4, {}),
             (sol_aa, aa, n - n // 2 - n // 4, dict(ref_density=(1.0, 1.0), p_level=(0.1, 10.0)))]
    Xs, Ts = [], []
    chunk = 250000
    for pi, (sol, mats, m_total, kw) in enumerate(parts):
        for k, start in enumerate(range(0, m_total, chunk)):
            m = min(chunk, m_total - start)
            X, T, _ = guess.labelled_features(sol, *guess.sample_problems(mats, m, seed=seed + 1000 * pi + k, **kw))
            Xs.append(X)
            Ts.append(T)
    return sol_wa, np.concatenate(Xs), np.concatenate(Ts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=2000000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    t0 = time.time()
    sol, X, T = build_training_set(a.problems, a.seed)
    # 5 % held out (seeded split)
    perm = np.random.default_rng(a.seed).permutation(len(X))
    nh = len(X) // 20
    Xh, Th, Xt, Tt = X[perm[:nh]], T[perm[:nh]], X[perm[nh:]], T[perm[nh:]]
    print("training set: %d labelled problems (%d held out) in %.1f s" % (len(X), nh, time.time() - t0), flush=True)
    t1 = time.time()
    # step schedule: the full rate for half the epochs, then halvings
    e1 = a.epochs // 2
    rest = a.epochs - e1
    sched = [(a.lr, e1), (a.lr / 2, rest // 3), (a.lr / 4, rest // 3), (a.lr / 8, rest - 2 * (rest // 3))]
    hist = list(guess.train(sol, Xt, Tt, epochs=sched[0][1], batch=a.batch, lr=sched[0][0], seed=a.seed))
    for k, (lr, ep) in enumerate(sched[1:]):
        if ep > 0:
            hist += list(sol.mlp_train(Xt, Tt, ep, a.batch, lr, a.seed + 1 + k))
    yh = sol.mlp_predict(Xh)
    print("trained %d epochs in %.1f s; loss %s" % (a.epochs, time.time() - t1, " ".join("%.3g" % v for v in hist)),
          flush=True)
    print("held-out mean squared error of log((p*+pc)/P): %.4f (target variance %.3f)"
          % (float(np.mean((yh - Th) ** 2)), float(np.var(Th))), flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.save(a.out, sol.mlp_get_params())
    print("wrote", a.out)


if __name__ == "__main__":
    main()
