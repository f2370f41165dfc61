"""Training data and training of the learned p* guess.

Problems are drawn from a seed the way the benchmark description fixes them
(pressure ratio log-uniform in [1e-3, 1e4], density ratio log-uniform in
[1e-4, 1e1] in either orientation of the two materials, normal velocities
uniform in +-2 c_min), solved by the library's own FP64 exact solver on the
GPU (the labels are its p*), turned into features on the GPU and used to train
the network with the library's BF16-MFMA training kernels.
"""
import numpy as np


def _sound_speed(mat, rho, p):
    """c of a stiffened gas (SG and NASG with b = 0 behave alike here)."""
    gam, pc = mat.gamma, mat.pc
    return np.sqrt(np.maximum(gam * (p + pc) / rho, 1e-300))


def sound_speeds(mats, ids, V):
    """c of each state V[q] (rho, u, v, w, p) of material ids[q]."""
    return np.array([_sound_speed(mats[int(i)], v[0], v[4]) for i, v in zip(ids, V)])


def sample_problems(mats, n, seed, ref_density=None, p_level=(1e4, 1e6)):
    """n two-material Riemann problems between materials 0 and 1 (random
    orientation).  Returns dirs, Vm, idm, Vp, idp (float64 / int32)."""
    rng = np.random.default_rng(seed)
    if ref_density is None:  # water-like for stiff materials, air-like otherwise
        ref_density = [1000.0 if mats[i].pc > 0 else 1.2 for i in range(2)]
    # density of material 0 around its reference, material 1 = ratio x that,
    # ratio log-uniform in [1e-4, 1e1]; then a random side for each
    rho_a = ref_density[0] * 10 ** rng.uniform(-0.5, 0.5, n)
    rho_b = rho_a * 10 ** rng.uniform(-4, 1, n)
    flip = rng.random(n) < 0.5
    idm = np.where(flip, 1, 0).astype(np.int32)
    idp = 1 - idm
    rhom = np.where(flip, rho_b, rho_a)
    rhop = np.where(flip, rho_a, rho_b)
    pm = 10 ** rng.uniform(np.log10(p_level[0]), np.log10(p_level[1]), n)
    pp = pm * 10 ** rng.uniform(-3, 4, n)
    swap = rng.random(n) < 0.5
    pm, pp = np.where(swap, pp, pm), np.where(swap, pm, pp)
    cm = np.array([_sound_speed(mats[i], r, p) for i, r, p in zip(idm, rhom, pm)])
    cp = np.array([_sound_speed(mats[i], r, p) for i, r, p in zip(idp, rhop, pp)])
    cmin = np.minimum(cm, cp)
    Vm = np.zeros((n, 5))
    Vp = np.zeros((n, 5))
    Vm[:, 0], Vp[:, 0] = rhom, rhop
    Vm[:, 4], Vp[:, 4] = pm, pp
    d = rng.integers(0, 3, n)
    dirs = np.zeros((n, 3))
    dirs[np.arange(n), d] = 1.0
    Vm[np.arange(n), 1 + d] = rng.uniform(-2, 2, n) * cmin
    Vp[np.arange(n), 1 + d] = rng.uniform(-2, 2, n) * cmin
    # tangential velocities do not enter the normal problem; keep some
    for V in (Vm, Vp):
        t = (d + 1) % 3
        V[np.arange(n), 1 + t] = rng.uniform(-0.1, 0.1, n) * cmin
    return dirs, Vm, idm, Vp, idp


def labelled_features(solver, dirs, Vm, idm, Vp, idp):
    """Features X and targets T of the problems the plain FP64 solver solves
    (err == 0), with the solver's per-problem iteration counts."""
    r = solver.riemann_batch(dirs, Vm, idm, Vp, idp)
    good = r["err"] == 0
    pstar = r["Vsm"][:, 4]
    X, T, ok = solver.mlp_features(dirs, Vm, idm, Vp, idp, pstar)
    keep = good & ok
    return X[keep], T[keep], r


def train(solver, X, T, epochs=30, batch=4096, lr=2e-3, seed=0):
    """Initialise (seeded) and train on the GPU; returns the loss history."""
    theta = _init(seed)
    solver.mlp_set_params(theta)
    return solver.mlp_train(X, T, epochs, batch, lr, seed)


def _init(seed):
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / 8), (8, 64)), np.zeros(64),
             rng.normal(0, np.sqrt(2.0 / 64), (64, 64)), np.zeros(64),
             rng.normal(0, np.sqrt(2.0 / 64), (64, 64)), np.zeros(64),
             rng.normal(0, np.sqrt(1.0 / 64), 64), np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)
