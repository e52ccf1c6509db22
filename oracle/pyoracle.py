"""ORACLE -- test infrastructure only.

ctypes view of oracle/_build/liborc.so, the CPU FP64 restatement of the
reference pipeline.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, always as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liborc.so"
HARNESS = HERE / "_ref" / "ref_harness"

_lib = None


def build(ref: bool = True) -> None:
    """make the oracle (and, where the reference sources exist, the harness)."""
    subprocess.run(["make", "-C", str(HERE), "-j8", "all"], check=True, capture_output=True)
    if ref and Path("/root/reference/Code").is_dir():
        subprocess.run(["make", "-C", str(HERE), "ref"], check=True, capture_output=True)


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build(ref=False)
        lib = C.CDLL(str(LIB))
        vp, ci, cd = C.c_void_p, C.c_int, C.c_double
        lib.orc_create.restype = vp
        lib.orc_create.argtypes = [ci, vp, vp, ci, vp, vp, ci, cd, cd, cd, cd, cd, cd, cd, ci, ci, ci]
        lib.orc_create_nas.restype = vp
        lib.orc_create_nas.argtypes = [C.c_char_p, ci, cd, cd, cd, cd, cd, cd, cd, ci, ci, ci]
        lib.orc_destroy.argtypes = [vp]
        for name in ("orc_dims", "orc_wave"):
            getattr(lib, name).argtypes = [vp, vp]
        lib.orc_mesh_geometry.argtypes = [vp, vp, vp]
        lib.orc_rwg.argtypes = [vp, vp, vp, vp]
        lib.orc_level_count.argtypes = [vp, ci]
        lib.orc_level_nodes.argtypes = [vp, ci, vp, vp, vp, vp]
        lib.orc_level_list_total.argtypes = [vp, ci, ci]
        lib.orc_level_list.argtypes = [vp, ci, ci, vp, vp]
        lib.orc_drvec_count.argtypes = [vp, ci]
        lib.orc_drvec.argtypes = [vp, ci, vp]
        lib.orc_grid_L.argtypes = [vp, ci, ci]
        lib.orc_grid.argtypes = [vp, ci, ci, vp, vp, vp, vp, vp]
        lib.orc_interp.argtypes = [vp, ci, ci, vp, vp]
        lib.orc_tf.argtypes = [vp, ci, ci, vp]
        lib.orc_pattern.argtypes = [vp, ci, ci, ci, vp]
        lib.orc_near_row_len.argtypes = [vp, ci]
        lib.orc_near_row.argtypes = [vp, ci, vp, vp]
        lib.orc_apply.argtypes = [vp, vp, vp]
        lib.orc_stage_pass.argtypes = [vp, ci, ci, vp]
        lib.orc_get_ws.argtypes = [vp, ci, ci, ci, ci, ci, vp]
        lib.orc_gmres.argtypes = [vp, vp, vp, ci, ci, cd, cd, vp, vp]
        lib.orc_cgs.argtypes = [vp, vp, vp, ci, cd, vp, vp]
        lib.orc_rhs.argtypes = [vp, cd, cd, cd, vp]
        lib.orc_rcs.argtypes = [vp, vp, cd, cd, cd, vp]
        lib.orc_dense.argtypes = [vp, vp]
        lib.orc_op.argtypes = [vp, ci, ci, ci, vp]
        lib.orc_gauss.argtypes = [ci, vp, vp, vp, vp]
        lib.orc_gauss_legendre.argtypes = [ci, vp, vp]
        lib.orc_sph_jy.argtypes = [ci, cd, vp, vp]
        lib.orc_legendre_p.argtypes = [ci, cd]
        lib.orc_legendre_p.restype = cd
        lib.orc_legendre_p_prime.argtypes = [ci, cd]
        lib.orc_legendre_p_prime.restype = cd
        lib.orc_plane_wave.argtypes = [cd, cd, cd, vp]
        lib.orc_sing.argtypes = [vp, vp, vp, vp]
        lib.orc_threads.restype = ci
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """mode: 1 patterns, 2 near field, 4 synthetic values (timing only).

    tf_mode: 0 = the reference's translation series (parity); bit 1 = all L terms;
    bit 2 = Hankel functions at the complex kR (accuracy mode for lossy media)."""

    def __init__(self, mesh, ie=2, freq=299792458.0, eps=(1.0, 0.0), mu=(1.0, 0.0), alpha=0.2, leaf=0.3,
                 gauss=7, tf_mode=0, mode=3):
        lib = library()
        if isinstance(mesh, (str, Path)):
            h = lib.orc_create_nas(str(mesh).encode(), ie, freq, eps[0], eps[1], mu[0], mu[1], alpha, leaf,
                                   gauss, tf_mode, mode)
        else:
            h = lib.orc_create(mesh.n_points, _p(mesh.point_ids), _p(mesh.xyz), mesh.n_faces, _p(mesh.face_ids),
                               _p(mesh.faces), ie, freq, eps[0], eps[1], mu[0], mu[1], alpha, leaf, gauss,
                               tf_mode, mode)
        if not h:
            raise RuntimeError("oracle construction failed")
        self._h = C.c_void_p(h)
        self.ie = ie
        d = np.zeros(5, dtype=np.int32)
        lib.orc_dims(self._h, _p(d))
        self.n_rwg, self.n, self.max_level, self.level_span, self.n_leaves = map(int, d)

    def __del__(self):
        if getattr(self, "_h", None):
            library().orc_destroy(self._h)
            self._h = None

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        library().orc_apply(self._h, _p(x), _p(y))
        return y

    def gmres(self, rhs, itr_max=None, restart=30, tol_rel=1e-3, tol_abs=0.0):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(rhs)
        it = C.c_int()
        res = C.c_double()
        conv = library().orc_gmres(self._h, _p(rhs), _p(x), itr_max or 4 * self.n, restart, tol_abs, tol_rel,
                                   C.byref(it), C.byref(res))
        return x, {"iterations": it.value, "rel_residual": res.value, "converged": bool(conv)}

    def cgs(self, rhs, itr_max=None, tol=1e-3):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(rhs)
        it = C.c_int()
        res = C.c_double()
        conv = library().orc_cgs(self._h, _p(rhs), _p(x), itr_max or 4 * self.n, tol, C.byref(it), C.byref(res))
        return x, {"iterations": it.value, "rel_residual": res.value, "converged": bool(conv)}

    def rhs(self, pol=0.0, theta=0.0, phi=0.0):
        v = np.empty(self.n, dtype=np.complex128)
        library().orc_rhs(self._h, pol, theta, phi, _p(v))
        return v

    def rcs(self, currents, theta, phi, pol=0.0):
        I = np.ascontiguousarray(currents, dtype=np.complex128)
        out = np.empty(9)
        library().orc_rcs(self._h, _p(I), pol, theta, phi, _p(out))
        return out[:3], out[3::2] + 1j * out[4::2]

    def wave(self):
        o = np.empty(5, dtype=np.complex128)
        library().orc_wave(self._h, _p(o))
        return dict(zip(["k1", "k2", "eta1", "eta2", "etai"], o))

    def rwg(self):
        ints = np.empty((self.n_rwg, 6), dtype=np.int32)
        ln = np.empty(self.n_rwg)
        mid = np.empty((self.n_rwg, 3))
        library().orc_rwg(self._h, _p(ints), _p(ln), _p(mid))
        return ints, ln, mid

    def level(self, level):
        n = library().orc_level_count(self._h, level)
        codes = np.empty(n, dtype=np.uint64)
        cen = np.empty((n, 3))
        ln = np.empty(n)
        par = np.empty(n, dtype=np.int32)
        library().orc_level_nodes(self._h, level, _p(codes), _p(cen), _p(ln), _p(par))
        return codes, cen, ln, par

    def level_list(self, level, which):
        n = library().orc_level_count(self._h, level)
        tot = library().orc_level_list_total(self._h, level, which)
        ptr = np.empty(n + 1, dtype=np.int32)
        idx = np.empty(max(tot, 1), dtype=np.int32)
        library().orc_level_list(self._h, level, which, _p(ptr), _p(idx))
        return ptr, idx[:tot]

    def drvec(self, level):
        n = library().orc_drvec_count(self._h, level)
        out = np.empty((n, 3))
        library().orc_drvec(self._h, level, _p(out))
        return out

    def grid(self, wn, level):
        rev = self.max_level - level
        L = library().orc_grid_L(self._h, wn, rev)
        th, ph, w = np.empty(L), np.empty(2 * L), np.empty(L)
        kp = np.empty((2 * L * L, 3))
        dphi = np.zeros(1)
        library().orc_grid(self._h, wn, rev, _p(th), _p(ph), _p(w), _p(dphi), _p(kp))
        return L, th, ph, w, float(dphi[0]), kp

    def interp(self, wn, level):
        L = self.grid(wn, level)[0]
        taps = np.empty((2 * L * L, 9), dtype=np.int32)
        w = np.empty((2 * L * L, 9))
        library().orc_interp(self._h, wn, level - 2, _p(taps), _p(w))
        return taps, w

    def tf(self, wn, level):
        L = self.grid(wn, level)[0]
        nv = self.drvec(level).shape[0]
        out = np.empty((nv, 2 * L * L), dtype=np.complex128)
        library().orc_tf(self._h, wn, level - 2, _p(out))
        return out

    def pattern(self, wn, which, rwg):
        L = self.grid(wn, self.max_level)[0]
        out = np.empty((2 * L * L, 3), dtype=np.complex128)
        library().orc_pattern(self._h, wn, which, rwg, _p(out))
        return out

    def near_row(self, row):
        n = library().orc_near_row_len(self._h, row)
        vals = np.empty(n, dtype=np.complex128)
        ids = np.empty(n, dtype=np.int32)
        library().orc_near_row(self._h, row, _p(vals), _p(ids))
        return ids, vals

    def stage_pass(self, wn, cur, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        library().orc_stage_pass(self._h, wn, cur, _p(x))

    def workspace(self, wn, cur, which, level, node):
        L = self.grid(wn, level)[0]
        out = np.empty((2 * L * L, 3), dtype=np.complex128)
        library().orc_get_ws(self._h, wn, cur, which, level - 2, node, _p(out))
        return out

    def dense(self):
        Z = np.empty((self.n, self.n), dtype=np.complex128)
        library().orc_dense(self._h, _p(Z))
        return Z

    def op(self, which, f, s):
        out = np.empty(2)
        library().orc_op(self._h, which, f, s, _p(out))
        return out[0] + 1j * out[1]


def gauss(n):
    l1, l2, l3, w = (np.empty(n) for _ in range(4))
    library().orc_gauss(n, _p(l1), _p(l2), _p(l3), _p(w))
    return l1, l2, l3, w


def gauss_legendre(n):
    x, w = np.empty(n), np.empty(n)
    library().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def sph_jy(L, x):
    j, y = np.empty(L + 1), np.empty(L + 1)
    library().orc_sph_jy(L, x, _p(j), _p(y))
    return j, y


def plane_wave(pol, th, ph):
    out = np.empty(15)
    library().orc_plane_wave(pol, th, ph, _p(out))
    return out.reshape(5, 3)


def sing(tri, normal, r):
    """Analytic integrals over triangle `tri` (3x3) seen from r: Ivec, Isca, Ivec3, Isca3, rho."""
    t = np.ascontiguousarray(tri, dtype=np.float64).reshape(9)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    p = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(11)
    library().orc_sing(_p(t), _p(n), _p(p), _p(out))
    return out[0:3], out[3], out[4:7], out[7], out[8:11]


def harness(*args) -> dict:
    """Run the reference harness (only present where it was built)."""
    import json
    r = subprocess.run([str(HARNESS), *map(str, args)], capture_output=True, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])
