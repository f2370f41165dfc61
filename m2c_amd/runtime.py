"""Python binding of libm2c_amd.so through its C ABI (include/m2c_amd.h).

This mirrors the reference's operator interface for the explicit step:
`Solver.compute_residual` is SpaceOperator::ComputeResidual
(SpaceOperator.h:144-149), `Solver.compute_time_step` is
ComputeTimeStepSize (SpaceOperator.h:135-136), `Solver.advance_one_time_step`
is TimeIntegratorBase::AdvanceOneTimeStep (TimeIntegrator.h:77-88) and
`Solver.run_steps` is the driver loop of Main.cpp:653-712.  Fields cross the
boundary in the reference's SpaceVariable3D layout (ghosted, AoS).

The library must be built (`python -c "import __graft_entry__ as g; g.build()"`);
there is no fallback implementation: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libm2c_amd.so")
_lib = None
P = C.c_void_p
D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)


class M2CError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise M2CError("libm2c_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.m2cx_create.argtypes = [C.POINTER(abi.Grid), C.POINTER(abi.Material), C.c_int, C.POINTER(abi.Scheme),
                                  C.c_int, C.POINTER(P)]
        L.m2cx_destroy.argtypes = [P]
        L.m2cx_last_error.restype = C.c_char_p
        L.m2cx_last_error.argtypes = [P]
        L.m2cx_upload.argtypes = [P, C.c_int, P]
        L.m2cx_download.argtypes = [P, C.c_int, P]
        L.m2cx_set_time.argtypes = [P, C.c_double, C.c_int]
        L.m2cx_get_time.argtypes = [P, D, D, C.POINTER(C.c_int)]
        L.m2cx_compute_residual.argtypes = [P, C.c_int, C.c_int]
        L.m2cx_compute_time_step.argtypes = [P, D, D]
        L.m2cx_apply_boundary_conditions.argtypes = [P, C.c_int]
        L.m2cx_clip.argtypes = [P, C.c_int]
        L.m2cx_levelset_residual.argtypes = [P, C.c_int]
        L.m2cx_download_ls_residual.argtypes = [P, D]
        L.m2cx_levelset_reinitialize.argtypes = [P, C.c_int]
        L.m2cx_levelset_construct_narrow_band.argtypes = [P]
        L.m2cx_levelset_construct_narrow_band_ls.argtypes = [P, C.c_int]
        L.m2cx_levelset_reinitialize_ls.argtypes = [P, C.c_int, C.c_int]
        L.m2cx_update_material_id_by_levelset.argtypes = [P, D]
        L.m2cx_get_narrow_band.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_uint8),
                                           C.POINTER(C.c_uint8), D]
        L.m2cx_create_levelsets.argtypes = [C.POINTER(abi.Grid), C.POINTER(abi.Material), C.c_int,
                                            C.POINTER(abi.Scheme), C.POINTER(abi.Levelset), C.c_int, C.c_int,
                                            C.POINTER(P)]
        L.m2cx_upload_levelset.argtypes = [P, C.c_int, D]
        L.m2cx_download_levelset.argtypes = [P, C.c_int, C.c_int, D]
        L.m2cx_set_overlap.argtypes = [P, C.c_int]
        L.m2cx_advance_one_time_step.argtypes = [P, C.c_double, C.c_double, C.c_int]
        L.m2cx_run_steps.argtypes = [P, C.c_int]
        L.m2cx_get_counters.argtypes = [P, C.POINTER(abi.Counters)]
        L.m2cx_reset_counters.argtypes = [P]
        L.m2cx_get_riemann_histogram.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.m2cx_numerical_flux_batch.argtypes = [P, C.c_int, C.c_int, D, D, I32, D]
        L.m2cx_riemann_batch.argtypes = [P, C.c_int, D, D, I32, D, I32, D, I32, D, D, I32, I32, I32]
        # device pointers (as integers) for the entry point that allocates nothing
        L.m2cx_riemann_batch_noalloc.argtypes = [P, C.c_int] + [P] * 12
        L.m2cx_profile.argtypes = [P, C.c_int]
        L.m2cx_kernel_time.argtypes = [P, C.c_char_p, D, C.POINTER(C.c_int64)]
        L.m2cx_synchronize.argtypes = [P]
        F = C.POINTER(C.c_float)
        L.m2cx_mlp_set_params.argtypes = [P, F]
        L.m2cx_mlp_get_params.argtypes = [P, F]
        L.m2cx_mlp_features.argtypes = [P, C.c_int, D, D, I32, D, I32, D, F, F, I32]
        L.m2cx_mlp_predict.argtypes = [P, C.c_int, F, F]
        L.m2cx_mlp_loss_grad.argtypes = [P, C.c_int, F, F, F, F]
        L.m2cx_mlp_train.argtypes = [P, C.c_int, F, F, C.c_int, C.c_int, C.c_float, C.c_uint32, F]
        L.m2cx_autotune.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.m2cx_set_tile.argtypes = [P, C.c_int]
        L.m2cx_initialize.argtypes = [P, D, C.c_int, C.POINTER(abi.Geometry), C.c_int]
        L.m2cx_ensemble_create.argtypes = [C.POINTER(P), C.c_int, C.POINTER(P)]
        L.m2cx_ensemble_run_steps.argtypes = [P, C.c_int]
        L.m2cx_ensemble_last_error.restype = C.c_char_p
        L.m2cx_ensemble_last_error.argtypes = [P]
        L.m2cx_ensemble_destroy.argtypes = [P]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _d(a):
    return a.ctypes.data_as(D) if a is not None else None


def _i(a):
    return a.ctypes.data_as(I32)


class Solver:
    """One device-resident explicit-step solver (one GPU)."""

    def __init__(self, grid, mats, nmat, scheme, device=0, levelsets=None):
        """levelsets: a sequence of abi.Levelset (several level sets,
        m2cx_create_levelsets), or None for the scheme's level set 0"""
        self.L = lib()
        self.grid, self.mats, self.nmat, self.scheme = grid, mats, nmat, scheme
        h = P()
        if levelsets is None:
            rc = self.L.m2cx_create(C.byref(grid), mats, nmat, C.byref(scheme), device, C.byref(h))
        else:
            self._ls = abi.levelsets(*levelsets)
            rc = self.L.m2cx_create_levelsets(C.byref(grid), mats, nmat, C.byref(scheme), self._ls, len(levelsets),
                                              device, C.byref(h))
        self.h = h
        self._check(rc)
        self.shape = (grid.nz + 2, grid.ny + 2, grid.nx + 2)

    def close(self):
        if getattr(self, "h", None):
            self.L.m2cx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc < 0:
            msg = self.L.m2cx_last_error(self.h).decode() if self.h else ""
            raise M2CError("m2cx error %d: %s" % (rc, msg))
        return rc

    # ---- state transfer (SpaceVariable3D layout) ----
    def upload(self, field, arr):
        if field == abi.FIELD_ID:
            a = np.ascontiguousarray(arr, dtype=np.int32)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self.L.m2cx_upload(self.h, field, a.ctypes.data_as(P)))

    def download(self, field):
        if field == abi.FIELD_ID:
            out = np.zeros(self.shape, dtype=np.int32)
        elif field == abi.FIELD_PHI:
            out = np.zeros(self.shape)
        else:
            out = np.zeros(self.shape + (5,))
        self._check(self.L.m2cx_download(self.h, field, out.ctypes.data_as(P)))
        return out

    def set_state(self, V, ID, Phi=None):
        self.upload(abi.FIELD_V, V)
        self.upload(abi.FIELD_ID, ID)
        if Phi is not None:
            self.upload(abi.FIELD_PHI, Phi)

    def get_state(self):
        V = self.download(abi.FIELD_V)
        ID = self.download(abi.FIELD_ID)
        Phi = self.download(abi.FIELD_PHI) if self.scheme.n_levelsets else None
        return V, ID, Phi

    def set_time(self, t, step):
        self._check(self.L.m2cx_set_time(self.h, t, step))

    def time(self):
        t, dt, st = C.c_double(), C.c_double(), C.c_int()
        self._check(self.L.m2cx_get_time(self.h, C.byref(t), C.byref(dt), C.byref(st)))
        return t.value, dt.value, st.value

    # ---- SpaceOperator ----
    def compute_residual(self, v_field=abi.FIELD_V, capture=False):
        self._check(self.L.m2cx_compute_residual(self.h, v_field, int(capture)))
        return self.download(abi.FIELD_R)

    def compute_time_step(self):
        dt, cfl = C.c_double(), C.c_double()
        self._check(self.L.m2cx_compute_time_step(self.h, C.byref(dt), C.byref(cfl)))
        return dt.value, cfl.value

    def apply_boundary_conditions(self, v_field=abi.FIELD_V):
        self._check(self.L.m2cx_apply_boundary_conditions(self.h, v_field))

    def clip(self, v_field=abi.FIELD_V):
        return self._check(self.L.m2cx_clip(self.h, v_field))

    def levelset_residual(self, v_field=abi.FIELD_V):
        self._check(self.L.m2cx_levelset_residual(self.h, v_field))
        out = np.zeros(self.shape)
        self._check(self.L.m2cx_download_ls_residual(self.h, _d(out)))
        return out

    def set_overlap(self, on):
        """Stage kernel tiles without interface faces during the Riemann
        solve (on, the default) or one launch per stage after it."""
        return self._check(self.L.m2cx_set_overlap(self.h, int(bool(on))))

    def reinitialize(self, special_maxits=0):
        """LevelSetOperator::Reinitialize with must_do = true: level set 0
        back to a signed distance (scheme.reinit_* settings)."""
        return self._check(self.L.m2cx_levelset_reinitialize(self.h, special_maxits))

    def construct_narrow_band(self):
        """LevelSetOperator::ConstructNarrowBandInReinitializer: the band of
        level set 0 around its zero level set, phi cut off outside it (after
        set_state; initialize() builds it itself)."""
        return self._check(self.L.m2cx_levelset_construct_narrow_band(self.h))

    def upload_levelset(self, q, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        self._check(self.L.m2cx_upload_levelset(self.h, q, _d(phi)))

    def update_material_ids(self, phi_tmp):
        """MultiPhaseOperator::UpdateMaterialIDByLevelSet with Phi_tmp = the
        given level sets (one ghosted field per level set)"""
        arr = np.ascontiguousarray(np.stack([np.asarray(p, dtype=np.float64) for p in phi_tmp]))
        self._check(self.L.m2cx_update_material_id_by_levelset(self.h, _d(arr)))

    def download_levelset(self, q, what=abi.LS_PHI):
        out = np.zeros(self.shape)
        self._check(self.L.m2cx_download_levelset(self.h, q, what, _d(out)))
        return out

    def narrow_band(self, ls=0):
        """level (int32), useful and active (bool) per ghosted cell, and
        (phi_max, phi_min, phi_out_pos, phi_out_neg) of level set `ls`"""
        level = np.zeros(self.shape, dtype=np.int32)
        useful = np.zeros(self.shape, dtype=np.uint8)
        active = np.zeros(self.shape, dtype=np.uint8)
        b = np.zeros(4)
        U8 = C.POINTER(C.c_uint8)
        self._check(self.L.m2cx_get_narrow_band(self.h, ls, level.ctypes.data_as(C.POINTER(C.c_int32)),
                                                useful.ctypes.data_as(U8), active.ctypes.data_as(U8), _d(b)))
        return level, useful.astype(bool), active.astype(bool), tuple(b)

    # ---- TimeIntegrator / driver ----
    def advance_one_time_step(self, time, dt, step):
        return self._check(self.L.m2cx_advance_one_time_step(self.h, time, dt, step))

    def run_steps(self, n):
        return self._check(self.L.m2cx_run_steps(self.h, n))

    def synchronize(self):
        self._check(self.L.m2cx_synchronize(self.h))

    def counters(self):
        c = abi.Counters()
        self._check(self.L.m2cx_get_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in abi.Counters._fields_}

    def riemann_histogram(self):
        """(iters, fevals): faces per count of secant iterations and of f(p)
        evaluations since the last reset (last bin: that count or more)."""
        it = np.zeros(abi.RIEMANN_HIST_BINS, dtype=np.uint64)
        fe = np.zeros(abi.RIEMANN_HIST_BINS, dtype=np.uint64)
        self._check(self.L.m2cx_get_riemann_histogram(self.h, it.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                      fe.ctypes.data_as(C.POINTER(C.c_uint64))))
        return it, fe

    def reset_counters(self):
        self._check(self.L.m2cx_reset_counters(self.h))

    # ---- point kernels ----
    def numerical_flux_batch(self, d, Vm, Vp, ids):
        Vm = np.ascontiguousarray(Vm, dtype=np.float64)
        Vp = np.ascontiguousarray(Vp, dtype=np.float64)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        F = np.zeros((len(Vm), 5))
        self._check(self.L.m2cx_numerical_flux_batch(self.h, len(Vm), d, _d(Vm), _d(Vp), _i(ids), _d(F)))
        return F

    def riemann_batch(self, dirs, Vm, idm, Vp, idp):
        n = len(Vm)
        dirs, Vm, Vp = (np.ascontiguousarray(x, dtype=np.float64) for x in (dirs, Vm, Vp))
        idm, idp = (np.ascontiguousarray(x, dtype=np.int32) for x in (idm, idp))
        Vs, Vsm, Vsp = np.zeros((n, 5)), np.zeros((n, 5)), np.zeros((n, 5))
        mid, err, it, fe = (np.zeros(n, dtype=np.int32) for _ in range(4))
        self._check(self.L.m2cx_riemann_batch(self.h, n, _d(dirs), _d(Vm), _i(idm), _d(Vp), _i(idp), _d(Vs),
                                              _i(mid), _d(Vsm), _d(Vsp), _i(err), _i(it), _i(fe)))
        return dict(Vs=Vs, id=mid, Vsm=Vsm, Vsp=Vsp, err=err, iters=it, fevals=fe)

    def autotune(self):
        v = C.c_int()
        ms = (C.c_float * 4)()
        self._check(self.L.m2cx_autotune(self.h, C.byref(v), ms))
        return v.value, list(ms)

    def set_tile(self, variant):
        self._check(self.L.m2cx_set_tile(self.h, variant))

    def initialize(self, default_state, default_id, geoms):
        """Initial condition from planes and spheres on the device, then the
        boundary fill (SpaceInitializer rules)."""
        st = np.ascontiguousarray(default_state, dtype=np.float64)
        arr = (abi.Geometry * max(len(geoms), 1))(*geoms)
        self._check(self.L.m2cx_initialize(self.h, _d(st), default_id, arr, len(geoms)))

    # ---- learned p* guess ----
    NPARAM = 8961

    def mlp_set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        assert theta.size == self.NPARAM
        self._check(self.L.m2cx_mlp_set_params(self.h, _f(theta)))

    def mlp_get_params(self):
        out = np.zeros(self.NPARAM, dtype=np.float32)
        self._check(self.L.m2cx_mlp_get_params(self.h, _f(out)))
        return out

    def mlp_features(self, dirs, Vm, idm, Vp, idp, pstar=None):
        n = len(Vm)
        dirs, Vm, Vp = (np.ascontiguousarray(x, dtype=np.float64) for x in (dirs, Vm, Vp))
        idm, idp = (np.ascontiguousarray(x, dtype=np.int32) for x in (idm, idp))
        X = np.zeros((n, 8), dtype=np.float32)
        T = np.zeros(n, dtype=np.float32)
        ok = np.zeros(n, dtype=np.int32)
        ps = None if pstar is None else np.ascontiguousarray(pstar, dtype=np.float64)
        self._check(self.L.m2cx_mlp_features(self.h, n, _d(dirs), _d(Vm), _i(idm), _d(Vp), _i(idp), _d(ps), _f(X),
                                             _f(T), _i(ok)))
        return X, (T if pstar is not None else None), ok.astype(bool)

    def mlp_predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(len(X), dtype=np.float32)
        self._check(self.L.m2cx_mlp_predict(self.h, len(X), _f(X), _f(y)))
        return y

    def mlp_loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.NPARAM, dtype=np.float32)
        loss = C.c_float()
        self._check(self.L.m2cx_mlp_loss_grad(self.h, len(X), _f(X), _f(T), _f(g), C.byref(loss)))
        return loss.value, g

    def mlp_train(self, X, T, epochs, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        self._check(self.L.m2cx_mlp_train(self.h, len(X), _f(X), _f(T), epochs, batch, lr, seed, _f(hist)))
        return hist

    # ---- measurement ----
    def profile(self, on=True):
        self._check(self.L.m2cx_profile(self.h, int(on)))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int64()
        self._check(self.L.m2cx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Ensemble:
    """Independent solvers advanced in lockstep, one Riemann launch per RK
    stage for all of them, one captured HIP graph per time step."""

    def __init__(self, solvers):
        self.L = lib()
        self.members = list(solvers)
        arr = (C.c_void_p * len(self.members))(*[s.h for s in self.members])
        h = C.c_void_p()
        rc = self.L.m2cx_ensemble_create(arr, len(self.members), C.byref(h))
        self.h = h
        if rc < 0:
            msg = self.L.m2cx_ensemble_last_error(self.h).decode() if self.h else ""
            self.close()
            raise M2CError("m2cx_ensemble_create error %d: %s" % (rc, msg))

    def run_steps(self, n):
        rc = self.L.m2cx_ensemble_run_steps(self.h, n)
        if rc < 0:
            raise M2CError("m2cx_ensemble_run_steps error %d: %s"
                           % (rc, self.L.m2cx_ensemble_last_error(self.h).decode()))
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.m2cx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
