"""ctypes binding of libmlfma.so (include/mlfma.h).

The product path has no CPU fallback: if the HIP library is missing or no GPU is
present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .geometry import Mesh

_LIB_PATH = Path(__file__).resolve().parent / "libmlfma.so"

EFIE, CFIE, PMCHWT = 0, 1, 2
FP64, FP32, BF16 = 0, 1, 2
PRECOND_NONE, PRECOND_BLOCK_DIAG, PRECOND_POLY, PRECOND_INNER = 0, 1, 2, 3
PRECOND_FEATURES = 20
GMRES, CGS = 0, 1
TF_FULL_ORDER, TF_COMPLEX_ARG = 1, 2   # mlfma_params.tf_mode bits (include/mlfma.h)
STAGE_AGGREGATE, STAGE_M2M, STAGE_M2L_L2L, STAGE_FAR, STAGE_NEAR = range(5)

ERRORS = {1: "MLFMA_ERR_ARG", 2: "MLFMA_ERR_HIP", 3: "MLFMA_ERR_OOM", 4: "MLFMA_ERR_NONFINITE",
          5: "MLFMA_ERR_NOT_CONVERGED", 6: "MLFMA_ERR_MESH", 7: "MLFMA_ERR_STATE"}


class Params(C.Structure):
    _fields_ = [("integral_equation", C.c_int), ("gauss_points", C.c_int), ("frequency_hz", C.c_double),
                ("eps_r_re", C.c_double), ("eps_r_im", C.c_double), ("mu_r_re", C.c_double),
                ("mu_r_im", C.c_double), ("cfie_alpha", C.c_double), ("leaf_coeff", C.c_double),
                ("far_precision", C.c_int), ("near_precision", C.c_int), ("tf_mode", C.c_int),
                ("max_rhs", C.c_int),
                ("device", C.c_int), ("verbose", C.c_int), ("precond_precision", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("itr_max", C.c_int), ("restart", C.c_int), ("tol_rel", C.c_double),
                ("tol_abs", C.c_double), ("precond", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("matvecs", C.c_int), ("converged", C.c_int),
                ("rel_residual", C.c_double), ("seconds", C.c_double), ("fp64_matvecs", C.c_int),
                ("precond_fallbacks", C.c_int)]


_lib = None


def library() -> C.CDLL:
    """Load libmlfma.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} not built: run python -m mlfma_pmchwt_amd.build")
        lib = C.CDLL(str(_LIB_PATH))
        vp = C.c_void_p
        lib.mlfma_params_default.argtypes = [C.POINTER(Params)]
        lib.mlfma_create.argtypes = [C.POINTER(Params), C.c_int, vp, vp, C.c_int, vp, vp, C.POINTER(vp)]
        lib.mlfma_create_nas.argtypes = [C.POINTER(Params), C.c_char_p, C.POINTER(vp)]
        lib.mlfma_destroy.argtypes = [vp]
        lib.mlfma_destroy.restype = None
        lib.mlfma_last_error.argtypes = [vp]
        lib.mlfma_last_error.restype = C.c_char_p
        lib.mlfma_dims.argtypes = [vp, vp]
        lib.mlfma_apply.argtypes = [vp, C.c_int, vp, vp]
        lib.mlfma_apply_batch.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.mlfma_apply_device.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.mlfma_set_apply_precision.argtypes = [vp, C.c_int]
        lib.mlfma_set_graphs.argtypes = [vp, C.c_int]
        lib.mlfma_autotune.argtypes = [vp, C.c_int, C.c_int, vp]
        lib.mlfma_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        lib.mlfma_precond_features.argtypes = [vp, vp]
        lib.mlfma_precond_statistics.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
        lib.mlfma_set_precond_poly.argtypes = [vp, C.c_int, vp]
        lib.mlfma_set_precond_inner.argtypes = [vp, C.c_int]
        lib.mlfma_precond_apply.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.mlfma_precond_apply_batch.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.mlfma_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint, C.POINTER(vp)]
        lib.mlfma_mlp_destroy.argtypes = [vp]
        lib.mlfma_mlp_last_error.argtypes = [vp]
        lib.mlfma_mlp_last_error.restype = C.c_char_p
        lib.mlfma_mlp_num_params.argtypes = [vp, C.POINTER(C.c_longlong)]
        lib.mlfma_mlp_get_params.argtypes = [vp, vp]
        lib.mlfma_mlp_set_params.argtypes = [vp, vp]
        lib.mlfma_mlp_set_normalization.argtypes = [vp, vp, vp]
        lib.mlfma_mlp_predict.argtypes = [vp, C.c_int, vp, vp]
        lib.mlfma_mlp_loss_grad.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.POINTER(C.c_double), vp]
        lib.mlfma_mlp_train.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_float, vp]
        lib.mlfma_solve.argtypes = [vp, C.POINTER(SolveParams), C.c_int, vp, vp, C.POINTER(SolveStats)]
        lib.mlfma_solve_batch.argtypes = [vp, C.POINTER(SolveParams), C.c_int, C.c_int, vp, vp, vp]
        lib.mlfma_solve_multi.argtypes = [C.c_int, vp, C.POINTER(SolveParams), vp, vp, vp, vp]
        lib.mlfma_excitation.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp]
        lib.mlfma_rcs.argtypes = [vp, vp, C.c_int, vp, vp, C.c_double, vp]
        lib.mlfma_excitation_batch.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        lib.mlfma_monostatic.argtypes = [vp, vp, C.c_double, C.c_int, vp, vp, vp, vp]
        lib.mlfma_scatter_grid.argtypes = [C.c_double] * 5 + [vp, vp, vp]
        lib.mlfma_batch_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        lib.mlfma_batch_load_text.argtypes = [C.c_char_p, C.POINTER(vp)]
        lib.mlfma_batch_destroy.argtypes = [vp]
        lib.mlfma_batch_destroy.restype = None
        lib.mlfma_batch_count.argtypes = [vp, vp, vp]
        lib.mlfma_batch_task.argtypes = [vp, C.c_int, vp, vp, C.c_char_p, C.c_int]
        lib.mlfma_batch_grid.argtypes = [vp, vp, vp]
        lib.mlfma_batch_run.argtypes = [vp, vp, vp]
        lib.mlfma_batch_result.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.mlfma_batch_set_concurrency.argtypes = [vp, C.c_int]
        lib.mlfma_stage.argtypes = [vp, C.c_int, vp, vp]
        lib.mlfma_get_workspace.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        lib.mlfma_set_profiling.argtypes = [vp, C.c_int]
        lib.mlfma_get_profile.argtypes = [vp, vp, vp, vp]
        lib.mlfma_get_rwg.argtypes = [vp, vp, vp, vp]
        lib.mlfma_get_level.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        lib.mlfma_get_level_list.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        lib.mlfma_get_drvec.argtypes = [vp, C.c_int, vp, vp]
        lib.mlfma_get_grid.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        lib.mlfma_get_interp.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        lib.mlfma_get_tf.argtypes = [vp, C.c_int, C.c_int, vp]
        lib.mlfma_get_pattern.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
        lib.mlfma_get_near_row.argtypes = [vp, C.c_int, vp, vp, vp]
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def default_params(**kw) -> Params:
    p = Params()
    library().mlfma_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class MlfmaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Operator:
    """One MLFMA operator on the GPU (a mlfma_handle)."""

    def __init__(self, mesh: Mesh | str | os.PathLike, **params):
        lib = library()
        self.params = default_params(**params)
        h = C.c_void_p()
        if isinstance(mesh, (str, os.PathLike)):
            rc = lib.mlfma_create_nas(C.byref(self.params), str(mesh).encode(), C.byref(h))
        else:
            rc = lib.mlfma_create(C.byref(self.params), mesh.n_points, _p(mesh.point_ids), _p(mesh.xyz),
                                  mesh.n_faces, _p(mesh.face_ids), _p(mesh.faces), C.byref(h))
        if rc != 0:
            raise MlfmaError(rc, lib.mlfma_last_error(None).decode())
        self._h = h
        d = np.zeros(6, dtype=np.int64)
        lib.mlfma_dims(self._h, _p(d))
        self.n_rwg, self.n, self.max_level, self.n_leaves, self.near_entries, self.device_bytes = map(int, d)

    def close(self):
        if getattr(self, "_h", None):
            library().mlfma_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MlfmaError(rc, library().mlfma_last_error(self._h).decode())

    # ---------------------------------------------------------------- compute
    def apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.ndim == 1:
            y = np.empty_like(x)
            self._check(library().mlfma_apply(self._h, x.shape[0], _p(x), _p(y)))
            return y
        xb = np.asfortranarray(x)
        y = np.empty_like(xb)
        self._check(library().mlfma_apply_batch(self._h, x.shape[0], x.shape[1], _p(xb), _p(y)))
        return y

    def apply_device(self, x_ptr: int, y_ptr: int, nrhs: int = 1):
        self._check(library().mlfma_apply_device(self._h, self.n, nrhs, C.c_void_p(x_ptr), C.c_void_p(y_ptr)))

    def set_graphs(self, on: bool = True):
        """Record batched matvecs as hipGraphs and replay them."""
        self._check(library().mlfma_set_graphs(self._h, int(on)))

    def autotune(self, nrhs: int, precision: int = FP32) -> dict:
        out = np.zeros(4)
        self._check(library().mlfma_autotune(self._h, nrhs, precision, _p(out)))
        return {"rhs_per_pass": int(out[0]), "matrix_core_near": bool(int(out[1]) & 1),
                "matrix_core_disaggregation": bool(int(out[1]) & 2), "cluster_translation": bool(int(out[1]) & 4),
                "bits": int(out[1]), "ms": float(out[2]), "ms_slowest": float(out[3])}

    def set_tuning(self, precision: int = FP32, rhs_per_pass: int = 0, bits: int = 7):
        """Launch configuration without timing: bit 0 matrix-core near field,
        bit 1 matrix-core disaggregation, bit 2 sibling-cluster translation."""
        self._check(library().mlfma_set_tuning(self._h, int(precision), int(rhs_per_pass), int(bits)))

    def set_apply_precision(self, precision: int):
        """FP64 or FP32 far field for apply* and the Krylov iterations of solve."""
        self._check(library().mlfma_set_apply_precision(self._h, int(precision)))

    def solve(self, rhs: np.ndarray, x0: np.ndarray | None = None, solver=GMRES, itr_max=None, restart=30,
              tol_rel=1e-3, tol_abs=0.0, precond=0, check=False):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(rhs) if x0 is None else np.array(x0, dtype=np.complex128)
        sp = SolveParams(solver, itr_max if itr_max is not None else 4 * self.n, restart, tol_rel, tol_abs, precond)
        st = SolveStats()
        rc = library().mlfma_solve(self._h, C.byref(sp), self.n, _p(rhs), _p(x), C.byref(st))
        if rc not in (0, 5) or (check and rc != 0):
            self._check(rc)
        return x, {"iterations": st.iterations, "matvecs": st.matvecs, "converged": bool(st.converged),
                   "rel_residual": st.rel_residual, "seconds": st.seconds, "fp64_matvecs": st.fp64_matvecs,
                   "precond_fallbacks": st.precond_fallbacks}

    def solve_batch(self, rhs: np.ndarray, x0: np.ndarray | None = None, itr_max=None, restart=30, tol_rel=1e-3,
                    tol_abs=0.0, precond=0):
        """GMRES for the columns of rhs (n x R) in lockstep with batched matvecs."""
        rhs = np.asfortranarray(rhs, dtype=np.complex128)
        if rhs.ndim != 2 or rhs.shape[0] != self.n:
            raise ValueError("rhs must be n x R")
        R = rhs.shape[1]
        x = np.zeros_like(rhs, order="F") if x0 is None else np.asfortranarray(x0, dtype=np.complex128).copy(order="F")
        sp = SolveParams(GMRES, itr_max if itr_max is not None else 4 * self.n, restart, tol_rel, tol_abs, precond)
        st = (SolveStats * R)()
        rc = library().mlfma_solve_batch(self._h, C.byref(sp), self.n, R, rhs.ctypes.data_as(C.c_void_p),
                                         x.ctypes.data_as(C.c_void_p), st)
        if rc not in (0, 5):
            self._check(rc)
        return x, [{"iterations": s.iterations, "matvecs": s.matvecs, "converged": bool(s.converged),
                    "rel_residual": s.rel_residual, "seconds": s.seconds, "fp64_matvecs": s.fp64_matvecs,
                    "precond_fallbacks": s.precond_fallbacks} for s in st]

    # ---------------------------------------------------------------- learned preconditioner
    def precond_features(self) -> np.ndarray:
        f = np.zeros(PRECOND_FEATURES, dtype=np.float32)
        self._check(library().mlfma_precond_features(self._h, _p(f)))
        return f

    def precond_statistics(self, rhs: np.ndarray, steps: int = 30, degree: int = 6):
        """(G, h, c) of the polynomial coefficients from `steps` Krylov pairs started at rhs."""
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        m = degree + 1
        G = np.zeros((m, m), dtype=np.complex128)
        h = np.zeros(m, dtype=np.complex128)
        c = np.zeros(1)
        self._check(library().mlfma_precond_statistics(self._h, _p(rhs), steps, degree, _p(G), _p(h), _p(c)))
        return G, h, float(c[0])

    def set_precond_inner(self, steps: int = 10):
        """GMRES steps on the near field per application of preconditioner 3."""
        self._check(library().mlfma_set_precond_inner(self._h, int(steps)))

    def set_precond_poly(self, alpha):
        a = np.ascontiguousarray(alpha, dtype=np.complex128)
        self._check(library().mlfma_set_precond_poly(self._h, a.size - 1, _p(a)))

    def precond_apply(self, x: np.ndarray, kind: int = 2) -> np.ndarray:
        """M^-1 x; a 2-D x (n x nrhs) goes through the batched polynomial path (kind 2)."""
        if np.ndim(x) == 2:
            if kind != PRECOND_POLY:
                raise ValueError("batched preconditioner apply is polynomial only")
            xt = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).T)
            yt = np.empty_like(xt)
            self._check(library().mlfma_precond_apply_batch(self._h, self.n, xt.shape[0], _p(xt), _p(yt)))
            return yt.T
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        self._check(library().mlfma_precond_apply(self._h, kind, x.size, _p(x), _p(y)))
        return y

    def excitation(self, pol=0.0, theta=0.0, phi=0.0) -> np.ndarray:
        v = np.empty(self.n, dtype=np.complex128)
        self._check(library().mlfma_excitation(self._h, pol, theta, phi, _p(v)))
        return v

    def rcs(self, currents: np.ndarray, theta, phi, pol=0.0) -> np.ndarray:
        I = np.ascontiguousarray(currents, dtype=np.complex128)
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(phi), th.shape), dtype=np.float64)
        out = np.empty((th.size, 3))
        self._check(library().mlfma_rcs(self._h, _p(I), th.size, _p(th), _p(ph), pol, _p(out)))
        return out

    def excitation_batch(self, pol, theta, phi) -> np.ndarray:
        """Plane-wave right-hand sides, one column per (pol, theta, phi)."""
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        pl = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(pol), th.shape), dtype=np.float64)
        ph = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(phi), th.shape), dtype=np.float64)
        v = np.empty((th.size, self.n), dtype=np.complex128)   # column-major n x m
        self._check(library().mlfma_excitation_batch(self._h, th.size, _p(pl), _p(th), _p(ph), _p(v)))
        return v.T

    def monostatic(self, theta, phi, pol=0.0, restart=50, tol_rel=1e-4, tol_abs=1e-8, itr_max=None, precond=0):
        """Monostatic RCS (total, V, H dBsm per direction) and per-direction solve statistics."""
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(phi), th.shape), dtype=np.float64)
        sp = SolveParams(GMRES, itr_max if itr_max is not None else self.n, restart, tol_rel, tol_abs, precond)
        out = np.empty((th.size, 3))
        st = (SolveStats * th.size)()
        rc = library().mlfma_monostatic(self._h, C.byref(sp), pol, th.size, _p(th), _p(ph), _p(out), st)
        if rc not in (0, 5):
            self._check(rc)
        return out, [{"iterations": s.iterations, "converged": bool(s.converged), "rel_residual": s.rel_residual}
                     for s in st]

    def stage(self, which: int, x: np.ndarray | None = None) -> np.ndarray:
        y = np.zeros(self.n, dtype=np.complex128)
        xp = None if x is None else np.ascontiguousarray(x, dtype=np.complex128)
        self._check(library().mlfma_stage(self._h, which, None if xp is None else _p(xp), _p(y)))
        return y

    # ---------------------------------------------------------------- profiling
    PROFILE_GROUPS = ("aggregate", "interpolate", "translate_anterpolate", "near", "disaggregate")

    def set_profiling(self, on: bool = True):
        self._check(library().mlfma_set_profiling(self._h, int(on)))

    def profile(self) -> dict:
        ms = np.zeros(5)
        launches = np.zeros(5, dtype=np.int64)
        n = C.c_int()
        self._check(library().mlfma_get_profile(self._h, _p(ms), _p(launches), C.byref(n)))
        return {"matvecs": n.value,
                "groups": {g: {"ms": float(ms[i]), "launches": int(launches[i])}
                           for i, g in enumerate(self.PROFILE_GROUPS)}}

    # ---------------------------------------------------------------- inspection
    def grid(self, wn: int, level: int):
        L = C.c_int()
        library().mlfma_get_grid(self._h, wn, level, C.byref(L), None, None, None, None)
        th = np.empty(L.value)
        ph = np.empty(2 * L.value)
        w = np.empty(L.value)
        dphi = np.zeros(1)
        self._check(library().mlfma_get_grid(self._h, wn, level, C.byref(L), _p(th), _p(ph), _p(w), _p(dphi)))
        return L.value, th, ph, w, float(dphi[0])

    def workspace(self, wn, cur, which, level, node):
        L = self.grid(wn, level)[0]
        out = np.empty((2 * L * L, 3), dtype=np.complex128)
        self._check(library().mlfma_get_workspace(self._h, wn, cur, which, level, node, _p(out)))
        return out

    def rwg(self):
        ints = np.empty((self.n_rwg, 6), dtype=np.int32)
        ln = np.empty(self.n_rwg)
        mid = np.empty((self.n_rwg, 3))
        self._check(library().mlfma_get_rwg(self._h, _p(ints), _p(ln), _p(mid)))
        return ints, ln, mid

    def level(self, level: int):
        n = C.c_int()
        self._check(library().mlfma_get_level(self._h, level, C.byref(n), None, None, None, None))
        codes = np.empty(n.value, dtype=np.uint64)
        cen = np.empty((n.value, 3))
        ln = np.empty(n.value)
        par = np.empty(n.value, dtype=np.int32)
        self._check(library().mlfma_get_level(self._h, level, C.byref(n), _p(codes), _p(cen), _p(ln), _p(par)))
        return codes, cen, ln, par

    def level_list(self, level: int, which: int):
        tot = C.c_int()
        self._check(library().mlfma_get_level_list(self._h, level, which, C.byref(tot), None, None))
        n = self.level(level)[0].size
        ptr = np.empty(n + 1, dtype=np.int32)
        idx = np.empty(max(tot.value, 1), dtype=np.int32)
        self._check(library().mlfma_get_level_list(self._h, level, which, C.byref(tot), _p(ptr), _p(idx)))
        return ptr, idx[:tot.value]

    def drvec(self, level: int):
        n = C.c_int()
        self._check(library().mlfma_get_drvec(self._h, level, C.byref(n), None))
        out = np.empty((n.value, 3))
        self._check(library().mlfma_get_drvec(self._h, level, C.byref(n), _p(out)))
        return out

    def interp(self, wn: int, level: int):
        L = self.grid(wn, level)[0]
        taps = np.empty((2 * L * L, 9), dtype=np.int32)
        w = np.empty((2 * L * L, 9))
        self._check(library().mlfma_get_interp(self._h, wn, level, _p(taps), _p(w)))
        return taps, w

    def tf(self, wn: int, level: int):
        L = self.grid(wn, level)[0]
        nv = self.drvec(level).shape[0]
        out = np.empty((nv, 2 * L * L), dtype=np.complex128)
        self._check(library().mlfma_get_tf(self._h, wn, level, _p(out)))
        return out

    def pattern(self, wn: int, which: int, rwg: int):
        L = self.grid(wn, self.max_level)[0]
        out = np.empty((2 * L * L, 3), dtype=np.complex128)
        self._check(library().mlfma_get_pattern(self._h, wn, which, rwg, _p(out)))
        return out

    def near_row(self, row: int):
        n = C.c_int()
        self._check(library().mlfma_get_near_row(self._h, row, C.byref(n), None, None))
        cols = np.empty(n.value, dtype=np.int32)
        vals = np.empty(n.value, dtype=np.complex128)
        self._check(library().mlfma_get_near_row(self._h, row, C.byref(n), _p(cols), _p(vals)))
        return cols, vals


class Mlp:
    """The coefficient network (mlfma_mlp): n_in -> width -> width -> n_out, tanh, FP32, on the GPU."""

    def __init__(self, n_in: int = PRECOND_FEATURES, width: int = 256, n_out: int = 14, seed: int = 0):
        lib = library()
        h = C.c_void_p()
        rc = lib.mlfma_mlp_create(n_in, width, n_out, seed, C.byref(h))
        if rc != 0:
            raise MlfmaError(rc, lib.mlfma_mlp_last_error(None).decode())
        self._h = h
        self.n_in, self.width, self.n_out = n_in, width, n_out
        cnt = C.c_longlong()
        lib.mlfma_mlp_num_params(self._h, C.byref(cnt))
        self.n_params = int(cnt.value)

    def close(self):
        if getattr(self, "_h", None):
            library().mlfma_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MlfmaError(rc, library().mlfma_mlp_last_error(self._h).decode())

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(library().mlfma_mlp_get_params(self._h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(library().mlfma_mlp_set_params(self._h, _p(p)))

    def set_normalization(self, mean, std):
        mean = np.ascontiguousarray(mean, dtype=np.float32)
        std = np.ascontiguousarray(std, dtype=np.float32)
        self._check(library().mlfma_mlp_set_normalization(self._h, _p(mean), _p(std)))

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self._check(library().mlfma_mlp_predict(self._h, x.shape[0], _p(x), _p(y)))
        return y

    def loss_grad(self, x, G, h, c):
        x = np.ascontiguousarray(x, dtype=np.float32)
        G = np.ascontiguousarray(G, dtype=np.complex128)
        h = np.ascontiguousarray(h, dtype=np.complex128)
        c = np.ascontiguousarray(c, dtype=np.float64)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        self._check(library().mlfma_mlp_loss_grad(self._h, x.shape[0], _p(x), _p(G), _p(h), _p(c), C.byref(loss),
                                                  _p(g)))
        return float(loss.value), g

    def train(self, x, G, h, c, steps: int = 500, lr: float = 1e-3) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        G = np.ascontiguousarray(G, dtype=np.complex128)
        h = np.ascontiguousarray(h, dtype=np.complex128)
        c = np.ascontiguousarray(c, dtype=np.float64)
        hist = np.zeros(max(steps, 1), dtype=np.float32)
        self._check(library().mlfma_mlp_train(self._h, x.shape[0], _p(x), _p(G), _p(h), _p(c), steps, lr,
                                              _p(hist)))
        return hist[:steps]


def scatter_grid(theta_start, theta_end, phi_start, phi_end, step):
    """The reference's scattering-angle grid: (theta, phi) arrays."""
    n = C.c_int(0)
    rc = library().mlfma_scatter_grid(theta_start, theta_end, phi_start, phi_end, step, C.byref(n), None, None)
    if rc != 0:
        raise MlfmaError(rc, "bad scattering grid")
    th, ph = np.empty(n.value), np.empty(n.value)
    library().mlfma_scatter_grid(theta_start, theta_end, phi_start, phi_end, step, C.byref(n), _p(th), _p(ph))
    return th, ph


def _stats_dicts(st):
    return [{"iterations": s.iterations, "matvecs": s.matvecs, "converged": bool(s.converged),
             "rel_residual": s.rel_residual, "seconds": s.seconds, "fp64_matvecs": s.fp64_matvecs,
             "precond_fallbacks": s.precond_fallbacks} for s in st]


def solve_multi(ops, rhs_list, x0_list=None, itr_max=None, restart=30, tol_rel=1e-3, tol_abs=0.0, precond=0):
    """Batched GMRES of several operators in lockstep (mlfma_solve_multi): every Arnoldi step
    runs all still-iterating objects' matvecs as one recorded graph.  rhs_list[k] is n_k x R_k;
    returns ([X_k], [[stats per right-hand side] per object]).  Graphs are switched on."""
    K = len(ops)
    if K < 1 or len(rhs_list) != K:
        raise ValueError("one right-hand-side block per operator")
    rhs = [np.asfortranarray(b, dtype=np.complex128) for b in rhs_list]
    for op, b in zip(ops, rhs):
        if b.ndim != 2 or b.shape[0] != op.n:
            raise ValueError("rhs must be n x R for its operator")
        op.set_graphs(True)
    xs = [np.zeros_like(b, order="F") if x0_list is None or x0_list[k] is None
          else np.asfortranarray(x0_list[k], dtype=np.complex128).copy(order="F") for k, b in enumerate(rhs)]
    sp = SolveParams(GMRES, itr_max if itr_max is not None else 4 * max(op.n for op in ops), restart, tol_rel,
                     tol_abs, precond)
    sts = [(SolveStats * b.shape[1])() for b in rhs]
    hs = (C.c_void_p * K)(*[op._h for op in ops])
    nr = (C.c_int * K)(*[b.shape[1] for b in rhs])
    bp = (C.c_void_p * K)(*[b.ctypes.data for b in rhs])
    xp = (C.c_void_p * K)(*[x.ctypes.data for x in xs])
    stp = (C.c_void_p * K)(*[C.cast(s, C.c_void_p) for s in sts])
    rc = library().mlfma_solve_multi(K, hs, C.byref(sp), nr, bp, xp, stp)
    if rc not in (0, 5):
        ops[0]._check(rc)
    return xs, [_stats_dicts(s) for s in sts]


class Batch:
    """The reference's batch configuration (Code/batch_config.txt format), run on the GPU."""

    def __init__(self, path=None, text=None):
        h = C.c_void_p()
        if text is not None:
            rc = library().mlfma_batch_load_text(text.encode(), C.byref(h))
        else:
            rc = library().mlfma_batch_load(str(path).encode(), C.byref(h))
        if rc != 0:
            raise MlfmaError(rc, "cannot read batch configuration")
        self._h = h
        nt, npt = C.c_int(0), C.c_int(0)
        library().mlfma_batch_count(self._h, C.byref(nt), C.byref(npt))
        self.n_tasks, self.n_points = nt.value, npt.value

    def close(self):
        if getattr(self, "_h", None):
            library().mlfma_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def task(self, i: int) -> dict:
        d = np.zeros(7)
        iv = np.zeros(5, dtype=np.int32)
        buf = C.create_string_buffer(4096)
        rc = library().mlfma_batch_task(self._h, i, _p(d), _p(iv), buf, 4096)
        if rc != 0:
            raise MlfmaError(rc, "bad task index")
        return {"nas": buf.value.decode(), "frequency": d[0], "theta": d[1], "phi": d[2], "eps_r": complex(d[3], d[4]),
                "mu_r": complex(d[5], d[6]), "algorithm": int(iv[0]), "integral_equation": int(iv[1]),
                "solver": int(iv[2]), "gauss_points": int(iv[3]), "mono": bool(iv[4])}

    def grid(self):
        th, ph = np.empty(self.n_points), np.empty(self.n_points)
        library().mlfma_batch_grid(self._h, _p(th), _p(ph))
        return th, ph

    def run(self, solve: SolveParams | None = None, concurrency: int = 1, **params) -> int:
        """Runs every task on the GPU; params override operator defaults (not equation/frequency/medium);
        concurrency: objects processed at once by that many host threads."""
        rc = library().mlfma_batch_set_concurrency(self._h, concurrency)
        if rc != 0:
            raise MlfmaError(rc, "concurrency must be 1..16")
        base = default_params(**params) if params else None
        return library().mlfma_batch_run(self._h, None if base is None else C.byref(base),
                                         None if solve is None else C.byref(solve))

    def result(self, task: int, pol: int):
        """(status, rcs[n_points, 3], stats) of run (task, pol); pol 0 = "h", 1 = "v"."""
        out = np.zeros((self.n_points, 3))
        st = SolveStats()
        rc = library().mlfma_batch_result(self._h, task, pol, _p(out), C.byref(st))
        return rc, out, {"iterations": st.iterations, "converged": bool(st.converged), "rel_residual": st.rel_residual}
