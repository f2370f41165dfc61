"""Train the learned p* guess on the GPU and store its parameters.

    python -m m2c_amd.train_guess [--problems 2000000] [--epochs 20] [--out PATH]

Problems (seeded) are drawn for two material pairs: the stiffened-gas water /
ideal-gas air pair of the shock-bubble workloads and the ideal-gas pair of the
two-material shock tube; the library's FP64 exact solver labels them, the
library's BF16-MFMA kernels train the network (Adam, mini-batches of 4096).
The parameters are written as float32 (.npy, 8961 values).
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
    wa = abi.materials(abi.material("SG", **configs.WATER_SG), abi.material("SG", **configs.AIR_SG))
    aa = abi.materials(abi.material("SG", gamma=1.4), abi.material("SG", gamma=1.4))
    sol_wa, sol_aa = solver_for(wa), solver_for(aa)
    n1 = int(n * 0.75)
    Xs, Ts = [], []
    chunk = 250000
    for k, start in enumerate(range(0, n1, chunk)):
        m = min(chunk, n1 - start)
        X, T, _ = guess.labelled_features(sol_wa, *guess.sample_problems(wa, m, seed=seed + k))
        Xs.append(X)
        Ts.append(T)
    for k, start in enumerate(range(0, n - n1, chunk)):
        m = min(chunk, n - n1 - start)
        X, T, _ = guess.labelled_features(sol_aa, *guess.sample_problems(aa, m, seed=seed + 1000 + k,
                                                                          ref_density=(1.0, 1.0),
                                                                          p_level=(0.1, 10.0)))
        Xs.append(X)
        Ts.append(T)
    return sol_wa, np.concatenate(Xs), np.concatenate(Ts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=2000000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    t0 = time.time()
    sol, X, T = build_training_set(a.problems, a.seed)
    print("training set: %d labelled problems in %.1f s" % (len(X), time.time() - t0), flush=True)
    t1 = time.time()
    hist = guess.train(sol, X, T, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    print("trained %d epochs in %.1f s; loss %s" % (a.epochs, time.time() - t1,
                                                     " ".join("%.3g" % v for v in hist)), flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.save(a.out, sol.mlp_get_params())
    print("wrote", a.out)


if __name__ == "__main__":
    main()
