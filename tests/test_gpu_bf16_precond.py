"""BF16 near-field products of the polynomial preconditioner (precond_precision = BF16).

* kernel: one polynomial step t_1 = t_0 - D^-1 Z_near t_0 (t_0 = D^-1 r) with
  the BF16 copy against a numpy restatement that rounds Z and t_0 to BF16 the
  same way (round to nearest even of the FP32 value) and accumulates in FP64;
  only the FP32 accumulation differs, so they agree to ~1e-6, while a wrong
  fragment layout, sign or plane gives O(1) and an unrounded product differs
  by the BF16 rounding (~1e-3);
* solves: with the BF16 preconditioner (flexible GMRES), single and batched
  FP32-mode solves reach the 1e-3 true residual of the oracle's operator in
  about as many iterations as with the FP32 preconditioner.
"""
import numpy as np
import pytest

from conftest import relerr, seeded_x
from mlfma_pmchwt_amd import api, geometry, learned
import pyoracle

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def cbf16(z):
    return bf16(z.real) + 1j * bf16(z.imag)


def dense_near(orc):
    n = orc.n
    A = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        ids, v = orc.near_row(i)
        A[i, ids] = v
    return A


CASES = {
    "pmchwt_1lvl": ((2, 0.5), api.PMCHWT, (4.0, -0.2)),
    "cfie_ml": ((2, 1.5), api.CFIE, (1.0, 0.0)),
}
_cache = {}


def case(name):
    if name not in _cache:
        (s, r), ie, eps = CASES[name]
        mesh = geometry.icosphere(s, r)
        print(f"[case] {name}: icosphere s={s} r={r} ie={ie} eps={eps} bf16 preconditioner", flush=True)
        op = api.Operator(mesh, integral_equation=ie, eps_r_re=eps[0], eps_r_im=eps[1], far_precision=api.FP32,
                          near_precision=api.FP32, precond_precision=api.BF16, max_rhs=40)
        orc = pyoracle.Oracle(mesh, ie=ie, eps=eps)
        _cache[name] = (op, orc, dense_near(orc))
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_bf16_step_matches_rounded_restatement(name):
    op, orc, An = case(name)
    d = np.diag(An)
    r = seeded_x(op.n, 31)
    op.set_precond_poly(np.array([0.0, 1.0]))
    op.set_apply_precision(api.FP32)
    y = op.precond_apply(r, kind=api.PRECOND_POLY)
    op.set_apply_precision(api.FP64)
    y64 = op.precond_apply(r, kind=api.PRECOND_POLY)
    t0 = r / d
    yq = t0 - (cbf16(An) @ cbf16(t0)) / d
    yx = t0 - (An @ t0) / d
    e_q, e_x, e64 = relerr(y, yq), relerr(y, yx), relerr(y64, yx)
    print(f"[bf16] {name}: vs rounded restatement {e_q:.2e}, vs exact {e_x:.2e}, FP64 path vs exact {e64:.2e}")
    assert e64 <= 1e-12
    assert e_q <= 2e-5
    assert 1e-5 < e_x <= 2e-2          # the BF16 rounding is there, and small


@pytest.mark.parametrize("R", [3, 12, 20, 40])
def test_bf16_batched_step(R):
    """The batched pass (8, 16, 32 and 64 right-hand sides per workgroup) gives the
    single-vector results column by column."""
    op, orc, An = case("pmchwt_1lvl")
    d = np.diag(An)
    Rm = np.stack([seeded_x(op.n, 400 + j) for j in range(R)], axis=1)
    op.set_precond_poly(np.array([0.0, 1.0]))
    op.set_apply_precision(api.FP32)
    Y = op.precond_apply(Rm, kind=api.PRECOND_POLY)
    T0 = Rm / d[:, None]
    Yq = T0 - (cbf16(An) @ cbf16(T0)) / d[:, None]
    for j in range(R):
        y1 = op.precond_apply(Rm[:, j], kind=api.PRECOND_POLY)
        assert relerr(Y[:, j], y1) <= 1e-6
        assert relerr(Y[:, j], Yq[:, j]) <= 2e-5
    op.set_apply_precision(api.FP64)


def test_bf16_preconditioned_solves():
    op, orc, An = case("pmchwt_1lvl")
    b0 = op.excitation(0.0, 0.0, 0.0)
    op.set_apply_precision(api.FP64)
    G, h, c = op.precond_statistics(b0, steps=30, degree=learned.DEGREE)
    op.set_precond_poly(learned.fit_coefficients(G, h, c))
    op.set_apply_precision(api.FP32)
    inc = [(0.0, 20.0, 0.0), (90.0, 60.0, 0.0), (0.0, 120.0, 45.0)]
    B = np.stack([op.excitation(*a) for a in inc], axis=1)
    X, st = op.solve_batch(B, tol_rel=1e-3, restart=30, itr_max=2000, precond=api.PRECOND_POLY)
    for j in range(B.shape[1]):
        x1, s1 = op.solve(B[:, j], tol_rel=1e-3, restart=30, itr_max=2000, precond=api.PRECOND_POLY)
        print(f"[bf16 solve] {inc[j]} batch it={st[j]['iterations']} single it={s1['iterations']} "
              f"fallbacks {st[j]['precond_fallbacks']}")
        assert st[j]["converged"] and s1["converged"]
        assert st[j]["precond_fallbacks"] == 0
        assert relerr(orc.apply(X[:, j]), B[:, j]) < 1e-3
        assert relerr(orc.apply(x1), B[:, j]) < 1e-3
        assert abs(st[j]["iterations"] - s1["iterations"]) <= 2
    # the same solves with the FP32 preconditioner products take about as many iterations
    ref = api.Operator(geometry.icosphere(2, 0.5), integral_equation=api.PMCHWT, eps_r_re=4.0, eps_r_im=-0.2,
                       far_precision=api.FP32, near_precision=api.FP32, max_rhs=8)
    ref.set_apply_precision(api.FP64)
    G, h, c = ref.precond_statistics(ref.excitation(0.0, 0.0, 0.0), steps=30, degree=learned.DEGREE)
    ref.set_precond_poly(learned.fit_coefficients(G, h, c))
    ref.set_apply_precision(api.FP32)
    _, st32 = ref.solve_batch(B, tol_rel=1e-3, restart=30, itr_max=2000, precond=api.PRECOND_POLY)
    for j in range(B.shape[1]):
        print(f"[bf16 solve] {inc[j]} bf16 it={st[j]['iterations']} fp32 it={st32[j]['iterations']}")
        assert st[j]["iterations"] <= 1.15 * st32[j]["iterations"] + 3
    ref.close()


def test_bf16_graph_replay_matches_eager():
    """The recorded (hipGraph) preconditioned batched matvec with BF16 products
    gives the eager solve's iterations and solutions (every Arnoldi step is a
    replay of the same recording)."""
    op, orc, An = case("pmchwt_1lvl")
    op.set_apply_precision(api.FP64)
    G, h, c = op.precond_statistics(op.excitation(0.0, 0.0, 0.0), steps=30, degree=learned.DEGREE)
    op.set_precond_poly(learned.fit_coefficients(G, h, c))
    op.set_apply_precision(api.FP32)
    B = np.stack([op.excitation(0.0, th, 0.0) for th in (15.0, 75.0, 135.0, 165.0)], axis=1)
    runs = {}
    for graphs in (False, True):
        op.set_graphs(graphs)
        runs[graphs] = op.solve_batch(B, tol_rel=1e-4, restart=30, itr_max=2000, precond=api.PRECOND_POLY)
    op.set_graphs(False)
    (Xe, se), (Xg, sg) = runs[False], runs[True]
    for j in range(B.shape[1]):
        print(f"[bf16 graph] column {j}: eager it={se[j]['iterations']} graph it={sg[j]['iterations']}")
        assert se[j]["converged"] and sg[j]["converged"]
        assert se[j]["iterations"] == sg[j]["iterations"]
    assert relerr(Xg, Xe) <= 1e-12
