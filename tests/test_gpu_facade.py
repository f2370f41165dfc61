"""The reference's time loop (Main.cpp:653-712) compiled in C++ against the
façades of include/m2c_amd_facade.hpp (SpaceVariable3D with GetDataPointer /
RestoreDataPointerAndInsert, SpaceOperator::ComputeTimeStepSize and
ComputeResidual, TimeIntegratorRK2/RK3::AdvanceOneTimeStep), driven on the
reference's fixture decks and compared with the oracle bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import deck as dk
import oracle_lib as ol
from m2c_amd import abi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def driver():
    path = os.path.join(ROOT, "tests", "libfacade_driver.so")
    if not os.path.exists(path):
        raise RuntimeError("façade test driver not built: run m2c_amd/build.sh")
    L = C.CDLL(path)
    L.facade_run.restype = C.c_int
    L.facade_run.argtypes = [C.POINTER(abi.Grid), C.POINTER(abi.Material), C.c_int, C.POINTER(abi.Scheme), D, I, D,
                             C.c_int, D, D, D, D, D]
    return L


def _dp(a):
    return a.ctypes.data_as(D)


@pytest.mark.parametrize("name", ["Test1", "Test4", "Test9"])
def test_reference_time_loop_through_facades(fixtures, name):
    d = fixtures[name]["deck"]
    g, mats, nmat, s, V, ID, Phi = dk.setup(d)
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()
    V0, ID0, P0 = o.get_state()
    o2 = ol.Oracle(g, mats, nmat, s)
    o2.set_state(V0, ID0, P0)
    R0 = o2.residual(capture=True)
    nsteps = len(fixtures[name]["log_steps"]) + 3
    Vf = np.zeros_like(V0)
    IDf = np.zeros(ID0.shape)
    Pf = np.zeros_like(P0)
    Rf = np.zeros_like(V0)
    t = C.c_double()
    V0c, ID0c, P0c = (np.ascontiguousarray(x) for x in (V0, ID0.astype(np.int32), P0))
    st = driver().facade_run(C.byref(g), mats, nmat, C.byref(s), _dp(V0c), ID0c.ctypes.data_as(I), _dp(P0c),
                             nsteps, _dp(Vf), _dp(IDf), _dp(Pf), _dp(Rf), C.byref(t))
    assert st == len(fixtures[name]["log_steps"]), st
    inner = (slice(1, -1),) * 3
    assert np.array_equal(Rf[inner], R0[inner])
    o.run_steps(st)
    Vo, IDo, Po = o.get_state()
    assert t.value == o.time()[0]
    assert np.array_equal(Vf[inner], Vo[inner])
    assert np.array_equal(IDf[inner], IDo[inner].astype(np.float64))
    assert np.array_equal(Pf[inner], Po[inner])
