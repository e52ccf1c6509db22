"""The near-field block-diagonal preconditioner in batched solves, and the fallback chain.

* mlfma_solve_batch with precond = 1 (block-diagonal) equals one single solve per right-hand
  side with precond = 1 (same recurrences per right-hand side), so iteration counts agree and
  the solutions agree to far below the tolerance; graphs on or off give the same result.
* A polynomial preconditioner that stalls is replaced by the block-diagonal one (and a stalling
  block-diagonal one by none): with all coefficients zero (M^-1 = 0, so a cycle cannot move x) on a CFIE sphere,
  the solve still converges and reports the replacement in precond_fallbacks.
"""
import numpy as np
import pytest

from conftest import relerr
from mlfma_pmchwt_amd import api, geometry

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ie,eps,graphs,tol", [(api.CFIE, (1.0, 0.0), False, 1e-8), (api.PMCHWT, (4.0, -0.2), True, 1e-4)])
def test_batched_block_diag_matches_single(ie, eps, graphs, tol):
    mesh = geometry.icosphere(2, 1.0)
    print(f"[blockdiag] ie={ie} graphs={graphs}", flush=True)
    op = api.Operator(mesh, integral_equation=ie, eps_r_re=eps[0], eps_r_im=eps[1], max_rhs=16)
    op.set_graphs(graphs)
    th = [0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0, 45.0, 135.0, 10.0, 170.0]
    B = op.excitation_batch([0.0] * len(th), th, [0.0] * len(th))
    X, st = op.solve_batch(B, restart=30, tol_rel=tol, itr_max=3000, precond=api.PRECOND_BLOCK_DIAG)
    for j in range(len(th)):
        x1, s1 = op.solve(np.ascontiguousarray(B[:, j]), restart=30, tol_rel=tol, itr_max=3000,
                          precond=api.PRECOND_BLOCK_DIAG)
        print(f"[blockdiag] rhs {j}: batched {st[j]['iterations']} it {st[j]['rel_residual']:.2e} "
              f"fallbacks {st[j]['precond_fallbacks']}; single {s1['iterations']} it {s1['rel_residual']:.2e} "
              f"fallbacks {s1['precond_fallbacks']}")
        assert st[j]["converged"] and s1["converged"]
        assert abs(st[j]["iterations"] - s1["iterations"]) <= 2
        assert relerr(X[:, j], x1) <= 100 * tol
    op.close()


@pytest.mark.parametrize("batched", [False, True])
def test_stalling_polynomial_falls_back(batched):
    mesh = geometry.icosphere(2, 1.0)
    op = api.Operator(mesh, integral_equation=api.CFIE, max_rhs=4)
    op.set_precond_poly(np.zeros(7))   # M^-1 = 0: the first cycle cannot reduce the residual
    b = op.excitation(0.0, 0.0, 0.0)
    if batched:
        B = np.stack([b, op.excitation(90.0, 60.0, 0.0)], axis=1)
        X, st = op.solve_batch(B, restart=30, tol_rel=1e-6, itr_max=4000, precond=api.PRECOND_POLY)
        s = st[0]
        print(f"[fallback] batched: {[(t['iterations'], t['rel_residual'], t['precond_fallbacks']) for t in st]}")
        assert all(t["converged"] and t["rel_residual"] < 1e-6 for t in st)
    else:
        x, s = op.solve(b, restart=30, tol_rel=1e-6, itr_max=4000, precond=api.PRECOND_POLY)
        print(f"[fallback] single: {s}")
        assert s["converged"] and s["rel_residual"] < 1e-6
    print(f"[fallback] batched={batched}: {s['iterations']} iterations, fallbacks {s['precond_fallbacks']}")
    assert s["precond_fallbacks"] >= 1
    op.close()


@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
def test_inner_near_solve_batched_matches_single(prec):
    """Preconditioner 3 (inner GMRES on the stored near field): the batched solve runs each
    right-hand side's own inner and outer recurrences, so it equals one single solve per
    right-hand side; with BF16 preconditioner products (FP32 mode) the true FP64 residual
    still reaches the tolerance, in fewer outer iterations than without a preconditioner."""
    mesh = geometry.icosphere(2, 1.0)
    op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=4.0, far_precision=api.FP32,
                      near_precision=api.FP32, precond_precision=api.BF16, max_rhs=8)
    op.set_apply_precision(prec)
    op.set_precond_inner(8)
    th = [0.0, 45.0, 90.0, 135.0, 180.0]
    B = op.excitation_batch([0.0, 90.0, 0.0, 90.0, 0.0], th, [0.0] * len(th))
    X, st = op.solve_batch(B, restart=30, tol_rel=1e-6, itr_max=2000, precond=api.PRECOND_INNER)
    for j in range(len(th)):
        x1, s1 = op.solve(np.ascontiguousarray(B[:, j]), restart=30, tol_rel=1e-6, itr_max=2000,
                          precond=api.PRECOND_INNER)
        x0, s0 = op.solve(np.ascontiguousarray(B[:, j]), restart=30, tol_rel=1e-6, itr_max=2000)
        print(f"[inner] prec {prec} rhs {j}: batched {st[j]['iterations']} it {st[j]['rel_residual']:.2e}; single "
              f"{s1['iterations']} it; unpreconditioned {s0['iterations']} it")
        assert st[j]["converged"] and s1["converged"] and st[j]["rel_residual"] < 1e-6
        assert abs(st[j]["iterations"] - s1["iterations"]) <= 2
        assert relerr(X[:, j], x1) <= 1e-4
        assert st[j]["iterations"] < s0["iterations"]
    op.close()
