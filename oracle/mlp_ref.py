"""CPU oracle of the learned p* guess (test infrastructure only).

A float64 numpy restatement of the network the library trains and evaluates
on the GPU (8 -> 64 -> 64 -> 64 -> 1, SiLU), its mean-squared-error loss and
the gradient of that loss with respect to the flat parameter vector
(W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4, W[in][out]).

`bf16=True` rounds, like the GPU kernels, the inputs, weight matrices, hidden
activations and back-propagated pre-activation gradients to bfloat16
(round to nearest even) while accumulating in float64; `bf16=False` is the
exact network, whose gradient is checked against finite differences.  The
network has no counterpart in the reference (it replaces the acoustic-theory
starting pressure of ExactRiemannSolverBase.cpp:1107-1158).
"""
import numpy as np

NPARAM = 8961
_SHAPES = [("W1", (8, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
           ("W4", (64,)), ("b4", (1,))]


def unpack(theta):
    out, o = {}, 0
    for name, shp in _SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    assert o == NPARAM
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).ravel() for name, _ in _SHAPES])


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even), as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed):
    """He-style initialisation (the library takes any parameter vector)."""
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, np.sqrt(2.0 / 8), (8, 64)), "b1": np.zeros(64),
         "W2": rng.normal(0, np.sqrt(2.0 / 64), (64, 64)), "b2": np.zeros(64),
         "W3": rng.normal(0, np.sqrt(2.0 / 64), (64, 64)), "b3": np.zeros(64),
         "W4": rng.normal(0, np.sqrt(1.0 / 64), 64), "b4": np.zeros(1)}
    return pack(p).astype(np.float32)


def forward(theta, X, bf16_mode=True):
    p = unpack(theta)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(X)
    W1, W2, W3, W4 = (r(p[k]) for k in ("W1", "W2", "W3", "W4"))
    z1 = x @ W1 + p["b1"]
    a1 = r(silu(z1))
    z2 = a1 @ W2 + p["b2"]
    a2 = r(silu(z2))
    z3 = a2 @ W3 + p["b3"]
    a3 = r(silu(z3))
    y = a3 @ W4 + p["b4"][0]
    return y, (x, z1, a1, z2, a2, z3, a3, W2, W3, W4)


def loss_grad(theta, X, T, bf16_mode=True):
    """Mean squared error over the rows of X and its gradient (NPARAM)."""
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    y, (x, z1, a1, z2, a2, z3, a3, W2, W3, W4) = forward(theta, X, bf16_mode)
    n = len(X)
    e = y - np.asarray(T, dtype=np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = {"W4": a3.T @ dy, "b4": np.array([dy.sum()])}
    dz3 = r(np.outer(dy, W4) * dsilu(z3))
    g["b3"] = dz3.sum(axis=0)
    g["W3"] = a2.T @ dz3
    dz2 = r((dz3 @ W3.T) * dsilu(z2))
    g["b2"] = dz2.sum(axis=0)
    g["W2"] = a1.T @ dz2
    dz1 = r((dz2 @ W2.T) * dsilu(z1))
    g["b1"] = dz1.sum(axis=0)
    g["W1"] = x.T @ dz1
    return loss, pack(g)
