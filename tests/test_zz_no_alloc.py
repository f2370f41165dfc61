"""m2cx_riemann_batch_noalloc (device arrays, fixed scratch, no allocation)
against m2cx_riemann_batch on the same problems, bit for bit, on two calls in
a row (the second smaller, so that stale scratch contents would show), with
the plain and the learned starting guess."""
import ctypes as C
import os

import numpy as np
import pytest

from m2c_amd import abi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class _Hip:
    """Device memory through the HIP runtime that libm2c_amd.so already
    uses (opened by its soname, so the loaded copy is the one bound).  torch
    is not used here: loaded after the library, it brings a second HIP
    runtime into the process."""

    def __init__(self):
        from m2c_amd import runtime
        runtime.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _problems(n, seed):
    rng = np.random.default_rng(seed)
    Vm = np.zeros((n, 5))
    Vp = np.zeros((n, 5))
    for V, r0 in ((Vm, 1000.0), (Vp, 1.2)):
        V[:, 0] = r0 * np.exp(rng.uniform(-0.5, 0.5, n))
        V[:, 4] = 1e5 * np.exp(rng.uniform(-1.0, 4.0, n))
        cs = np.sqrt(V[:, 4] / V[:, 0])
        V[:, 1:4] = rng.uniform(-0.5, 0.5, (n, 3)) * cs[:, None]
    dirs = np.zeros((n, 3))
    dirs[np.arange(n), rng.integers(0, 3, n)] = 1.0
    Vp[:4, 4] = Vm[:4, 4]
    Vp[:4, 1:4] = Vm[:4, 1:4]
    return dirs, Vm, np.zeros(n, np.int32), Vp, np.ones(n, np.int32)


@pytest.mark.parametrize("learned", [False, True])
def test_noalloc_batch_matches_batch(learned):
    from m2c_amd import runtime
    mats = abi.materials(abi.material("SG", gamma=6.12, pc=3.43e8), abi.material("SG", gamma=1.4))
    s = abi.default_scheme(riem_tol_main=1e-10, riem_guess=abi.GUESS_MLP if learned else abi.GUESS_ACOUSTIC)
    sol = runtime.Solver(abi.grid((4, 4, 4)), mats, 2, s)
    if learned:
        sol.mlp_set_params(np.load(os.path.join(ROOT, "m2c_amd", "data", "guess_params.npy")))
    L = runtime.lib()
    hip = _Hip()
    try:
        for call, n in enumerate((3000, 1100)):
            prob = _problems(n, 41 + call)
            ref = sol.riemann_batch(*prob)
            dirs, Vm, idm, Vp, idp = (hip.put(x) for x in prob)
            Vs, Vsm, Vsp = (hip.put(np.full((n, 5), -7.0)) for _ in range(3))
            mid, err, it, fe = (hip.put(np.full(n, -7, np.int32)) for _ in range(4))
            rc = L.m2cx_riemann_batch_noalloc(sol.h, n, dirs, Vm, idm, Vp, idp, Vs, mid, Vsm, Vsp, err, it, fe)
            assert rc == 0, L.m2cx_last_error(sol.h)
            sol.synchronize()
            got = dict(Vs=hip.get(Vs, (n, 5), np.float64), id=hip.get(mid, n, np.int32),
                       Vsm=hip.get(Vsm, (n, 5), np.float64), Vsp=hip.get(Vsp, (n, 5), np.float64),
                       err=hip.get(err, n, np.int32), iters=hip.get(it, n, np.int32),
                       fevals=hip.get(fe, n, np.int32))
            for key in ("Vs", "Vsm", "Vsp", "id", "err", "iters", "fevals"):
                a, b = got[key], ref[key]
                same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else (a == b)
                assert np.all(same), (call, key, np.nonzero(~same)[0][:5])
            # and without the optional count arrays
            rc = L.m2cx_riemann_batch_noalloc(sol.h, n, dirs, Vm, idm, Vp, idp, Vs, mid, Vsm, Vsp, err, None, None)
            assert rc == 0
            sol.synchronize()
            vsm = hip.get(Vsm, (n, 5), np.float64)
            ok = ~np.isnan(ref["Vsm"])
            assert np.array_equal(vsm[ok], ref["Vsm"][ok])
            hip.free()
    finally:
        hip.free()
        sol.close()


def test_noalloc_batch_refuses_too_many_problems():
    from m2c_amd import runtime
    mats = abi.materials(abi.material("SG", gamma=6.12, pc=3.43e8), abi.material("SG", gamma=1.4))
    sol = runtime.Solver(abi.grid((4, 4, 4)), mats, 2, abi.default_scheme())
    L = runtime.lib()
    rc = L.m2cx_riemann_batch_noalloc(sol.h, 65537, *([C.c_void_p(8)] * 12))
    assert rc == abi.E_ARG
    assert b"at most 65536" in L.m2cx_last_error(sol.h)
