"""Several objects in one lockstep solve (mlfma_solve_multi): every Arnoldi step runs all
still-iterating objects' preconditioned matvecs as one recorded graph with a branch per object.

* FP64, no preconditioner: members of the seeded C4 mix (cylinder, cube, ellipsoids at a quarter
  of their size, 8 incidences each) solved together equal their own batched solves, and the two
  smallest members' solutions meet the oracle's operator to the solver tolerance.
* Mixed precision with the polynomial preconditioner: every object reaches a true FP64 residual
  of 1e-3 together as alone, with iteration counts close to the single-object solves.
* Argument checks of the C entry point.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from mlfma_pmchwt_amd import api, configs, learned  # noqa: E402
import pyoracle  # noqa: E402

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _objects(indices, **kw):
    objs = configs.c4_objects(scale=0.25)
    inc = configs.c4_incidences()
    out = []
    for i in indices:
        o = objs[i]
        op = api.Operator(o["mesh"], integral_equation=api.PMCHWT, eps_r_re=o["eps_r"], max_rhs=8, **kw)
        B = op.excitation_batch([a[0] for a in inc], [a[1] for a in inc], [a[2] for a in inc])
        print(f"[multi] object {i}: {o['kind']} eps {o['eps_r']:.3f}, {op.n} unknowns, {op.max_level + 1} levels",
              flush=True)
        out.append((o, op, B))
    return out


def test_multi_equals_single_fp64_and_oracle():
    objs = _objects((0, 2, 3, 5))
    singles = []
    for o, op, B in objs:
        op.set_graphs(True)
        singles.append(op.solve_batch(B, restart=50, tol_rel=1e-4, itr_max=3000))
    Xm, stm = api.solve_multi([op for _, op, _ in objs], [B for _, _, B in objs], restart=50, tol_rel=1e-4,
                              itr_max=3000)
    for k, (o, op, B) in enumerate(objs):
        its_s = [s["iterations"] for s in singles[k][1]]
        its_m = [s["iterations"] for s in stm[k]]
        e = relerr(Xm[k], singles[k][0])
        print(f"[multi] object {o['index']}: iterations alone {min(its_s)}..{max(its_s)}, together "
              f"{min(its_m)}..{max(its_m)}, solutions differ {e:.2e}", flush=True)
        assert all(s["converged"] and s["rel_residual"] <= 1e-4 for s in stm[k])
        assert max(abs(a - b) for a, b in zip(its_s, its_m)) <= 2
        assert e <= 1e-6
    # the two smallest members against the oracle's operator
    for o, op, B in sorted(objs, key=lambda t: t[1].n)[:2]:
        k = [t[1] for t in objs].index(op)
        orc = pyoracle.Oracle(o["mesh"], ie=2, eps=(o["eps_r"], 0.0))
        for j in (0, 7):
            r = relerr(orc.apply(Xm[k][:, j]), B[:, j])
            print(f"[multi] object {o['index']} incidence {j}: oracle residual {r:.2e}")
            assert r <= 1.001e-4
    for _, op, _ in objs:
        op.close()


def test_multi_mixed_polynomial():
    objs = _objects((1, 4, 6), far_precision=api.FP32, near_precision=api.BF16, precond_precision=api.BF16)
    singles = []
    for o, op, B in objs:
        G, h, c = op.precond_statistics(B[:, 0], steps=30, degree=learned.DEGREE)
        op.set_precond_poly(learned.fit_coefficients(G, h, c))
        op.set_apply_precision(api.FP32)
        op.set_graphs(True)
        singles.append(op.solve_batch(B, restart=50, tol_rel=1e-3, itr_max=3000, precond=api.PRECOND_POLY))
    Xm, stm = api.solve_multi([op for _, op, _ in objs], [B for _, _, B in objs], restart=50, tol_rel=1e-3,
                              itr_max=3000, precond=api.PRECOND_POLY)
    for k, (o, op, B) in enumerate(objs):
        its_s = [s["iterations"] for s in singles[k][1]]
        its_m = [s["iterations"] for s in stm[k]]
        print(f"[multi mixed] object {o['index']} ({op.n} unknowns): iterations alone {min(its_s)}..{max(its_s)}, "
              f"together {min(its_m)}..{max(its_m)}, fallbacks {stm[k][0]['precond_fallbacks']}", flush=True)
        assert all(s["converged"] and s["rel_residual"] <= 1e-3 for s in stm[k])
        assert max(its_m) <= max(its_s) + 10
        op.set_apply_precision(api.FP64)
        R = op.apply(Xm[k]) - B
        assert np.linalg.norm(R, axis=0).max() / np.linalg.norm(B, axis=0).min() <= 1.2e-3
    for _, op, _ in objs:
        op.close()


def test_multi_argument_checks():
    objs = _objects((2,))
    _, op, B = objs[0]
    lib = api.library()
    sp = api.SolveParams(api.GMRES, 100, 30, 1e-3, 0.0, 0)
    n = (C.c_int * 2)(8, 8)
    Bf = np.asfortranarray(B)
    X = np.zeros_like(Bf)
    hs = (C.c_void_p * 2)(op._h, op._h)
    bp = (C.c_void_p * 2)(Bf.ctypes.data, Bf.ctypes.data)
    xp = (C.c_void_p * 2)(X.ctypes.data, X.ctypes.data)
    op.set_graphs(True)
    assert lib.mlfma_solve_multi(2, hs, C.byref(sp), n, bp, xp, None) == 1        # the same handle twice
    op.set_graphs(False)
    assert lib.mlfma_solve_multi(1, hs, C.byref(sp), n, bp, xp, None) == 7        # graphs off: MLFMA_ERR_STATE
    assert lib.mlfma_solve_multi(0, hs, C.byref(sp), n, bp, xp, None) == 1
    op.close()
