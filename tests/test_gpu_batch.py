"""Batched right-hand sides: the batched matvec and the lockstep GMRES against
the single-vector path and the oracle.

The batched kernels change only the summation order of the near-field column
groups, so a batched matvec matches the single-vector one to rounding
(1e-13) and the oracle to the FP64 parity bar (1e-10).  Lockstep GMRES runs
each right-hand side's own recurrences, so its iteration counts equal the
single solves' and its solutions agree to far below the tolerance.
"""
import numpy as np
import pytest

from conftest import relerr, seeded_x
from mlfma_pmchwt_amd import api, geometry
import pyoracle

pytestmark = pytest.mark.gpu

CASES = {
    "efie_ml": ((2, 1.5), api.EFIE, (1.0, 0.0)),
    "cfie_ml": ((2, 1.5), api.CFIE, (1.0, 0.0)),
    "pmchwt_lossy_ml": ((2, 1.5), api.PMCHWT, (2.5, -0.1)),
    "pmchwt_1lvl": ((2, 0.5), api.PMCHWT, (4.0, 0.0)),
}
_cache = {}


def pair(name):
    if name not in _cache:
        (s, r), ie, eps = CASES[name]
        mesh = geometry.icosphere(s, r)
        print(f"[case] {name}: icosphere s={s} r={r} ie={ie} eps={eps} batched", flush=True)
        op = api.Operator(mesh, integral_equation=ie, eps_r_re=eps[0], eps_r_im=eps[1], max_rhs=64,
                          far_precision=api.FP32, near_precision=api.FP32)
        orc = pyoracle.Oracle(mesh, ie=ie, eps=eps)
        _cache[name] = (mesh, op, orc)
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("R", [1, 3, 8, 13])
def test_batched_matvec(name, R):
    _, op, orc = pair(name)
    X = np.stack([seeded_x(op.n, 100 + j) for j in range(R)], axis=1)
    for prec in (api.FP64, api.FP32):
        op.set_apply_precision(prec)
        Y = op.apply(X)
        for j in range(R):
            y1 = op.apply(np.ascontiguousarray(X[:, j]))
            tol = 1e-13 if prec == api.FP64 else 1e-6
            assert relerr(Y[:, j], y1) <= tol
            if prec == api.FP64 and j in (0, R - 1):
                assert relerr(Y[:, j], orc.apply(X[:, j])) <= 1e-10
    op.set_apply_precision(api.FP64)


@pytest.mark.parametrize("name", ["cfie_ml", "pmchwt_1lvl"])
def test_batched_solve_matches_single(name):
    _, op, orc = pair(name)
    op.set_apply_precision(api.FP64)
    inc = [(0.0, th, 0.0) for th in (0.0, 40.0, 90.0, 150.0)] + [(90.0, 70.0, 0.0)]
    B = np.stack([op.excitation(*a) for a in inc], axis=1)
    Xb, stb = op.solve_batch(B, tol_rel=1e-6, restart=30)
    for j, a in enumerate(inc):
        x1, st1 = op.solve(B[:, j], tol_rel=1e-6, restart=30)
        print(f"[batch solve] {name} inc={a} it batch={stb[j]['iterations']} single={st1['iterations']} "
              f"res={stb[j]['rel_residual']:.2e}")
        assert stb[j]["converged"] and stb[j]["rel_residual"] < 1e-6
        assert abs(stb[j]["iterations"] - st1["iterations"]) <= 1
        assert relerr(Xb[:, j], x1) <= 1e-6
        # true residual against the oracle operator
        assert relerr(orc.apply(Xb[:, j]), B[:, j]) < 1.01e-6


def test_batched_solve_fp32_mode():
    _, op, orc = pair("cfie_ml")
    op.set_apply_precision(api.FP32)
    B = np.stack([op.excitation(0.0, th, 0.0) for th in np.linspace(0, 180, 9)], axis=1)
    Xb, stb = op.solve_batch(B, tol_rel=1e-3, restart=30)
    for j in range(B.shape[1]):
        assert stb[j]["converged"] and stb[j]["rel_residual"] < 1e-3
        assert relerr(orc.apply(Xb[:, j]), B[:, j]) < 1e-3
    op.set_apply_precision(api.FP64)


def test_batched_limits(sphere_small):
    op = api.Operator(sphere_small, integral_equation=api.CFIE, max_rhs=4)
    X = np.ones((op.n, 300), dtype=np.complex128)
    with pytest.raises(RuntimeError):
        op.apply(X)
    B = np.ones((op.n, 2), dtype=np.complex128)
    with pytest.raises(ValueError):
        op.solve_batch(B[:1])


@pytest.mark.parametrize("ie", [api.CFIE, api.PMCHWT])
@pytest.mark.parametrize("R", [2, 21])
@pytest.mark.parametrize("freq", [12, 20])
def test_batched_mfma_near_large_leaves(ie, R, freq):
    """Dense meshes on a small sphere: leaves of 37-148 (freq 12) and about 400
    (freq 20) RWGs exercise the matrix-core near product's multi-tile and
    multi-pass (more than 256 rows) paths (FP32 near)."""
    mesh = geometry.geodesic_sphere(freq, 0.5)
    op = api.Operator(mesh, integral_equation=ie, eps_r_re=4.0, max_rhs=32, far_precision=api.FP32,
                      near_precision=api.FP32)
    X = np.stack([seeded_x(op.n, 300 + j) for j in range(R)], axis=1)
    op.set_apply_precision(api.FP32)
    Y = op.apply(X)
    for j in range(R):
        y1 = op.apply(np.ascontiguousarray(X[:, j]))
        assert relerr(Y[:, j], y1) <= 2e-6
    op.close()


def test_graph_replays_match_eager():
    """Recorded batched matvecs replayed several times, with eager work and
    synchronisations in between, give the eager result every time."""
    mesh = geometry.icosphere(2, 1.5)
    op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=2.5, eps_r_im=-0.1, max_rhs=16,
                      far_precision=api.FP32, near_precision=api.FP32)
    for prec in (api.FP64, api.FP32):
        op.set_apply_precision(prec)
        X = np.stack([seeded_x(op.n, 500 + j) for j in range(11)], axis=1)
        op.set_graphs(False)
        Y0 = op.apply(X)
        op.set_graphs(True)
        for rep in range(4):
            Y = op.apply(X)
            assert np.array_equal(Y, Y0), f"replay {rep} differs"
            op.apply(np.ascontiguousarray(X[:, rep]))   # eager single matvec between replays
            op.solve(X[:, 0], tol_rel=1e-1, restart=5, itr_max=5)
        op.set_graphs(False)
    op.close()


def test_graph_batched_solve_matches_eager():
    _, op, orc = pair("pmchwt_1lvl")
    op.set_apply_precision(api.FP32)
    B = np.stack([op.excitation(0.0, th, 0.0) for th in (0.0, 50.0, 110.0)], axis=1)
    G, h, c = op.precond_statistics(B[:, 0], steps=30, degree=6)
    from mlfma_pmchwt_amd import learned
    op.set_precond_poly(learned.fit_coefficients(G, h, c))
    for pc in (0, api.PRECOND_POLY):
        op.set_graphs(False)
        X0, s0 = op.solve_batch(B, tol_rel=1e-3, restart=30, precond=pc)
        op.set_graphs(True)
        X1, s1 = op.solve_batch(B, tol_rel=1e-3, restart=30, precond=pc)
        X2, s2 = op.solve_batch(B, tol_rel=1e-3, restart=30, precond=pc)
        op.set_graphs(False)
        assert [s["iterations"] for s in s1] == [s["iterations"] for s in s0] == [s["iterations"] for s in s2]
        assert np.array_equal(X1, X0) and np.array_equal(X2, X0)
    op.set_apply_precision(api.FP64)


def test_autotune_keeps_results():
    mesh = geometry.icosphere(2, 1.5)
    op = api.Operator(mesh, integral_equation=api.CFIE, max_rhs=16, far_precision=api.FP32,
                      near_precision=api.FP32)
    X = np.stack([seeded_x(op.n, 700 + j) for j in range(16)], axis=1)
    op.set_apply_precision(api.FP32)
    Y0 = op.apply(X)
    t = op.autotune(16, api.FP32)
    print(f"[autotune] {t}")
    assert t["rhs_per_pass"] in (1, 2, 4, 8) and 0 < t["ms"] <= t["ms_slowest"]
    Y1 = op.apply(X)
    assert relerr(Y1, Y0) <= 2e-6   # another launch configuration: FP32 rounding only
    op.close()


@pytest.mark.parametrize("name", ["cfie_ml", "pmchwt_lossy_ml"])
def test_cluster_translation_matches_per_box(name):
    """Sibling-cluster translation reorders only the sum over distant relatives:
    FP64 results agree with the per-box kernel to rounding, both precisions,
    single and batched, and with the oracle at the FP64 bar."""
    mesh, op, orc = pair(name)
    X = np.stack([seeded_x(op.n, 900 + j) for j in range(9)], axis=1)   # passes of 8 and 1
    out = {}
    for bits in (0, 4, 12, 20):   # per box, sibling clusters, 12 columns per thread, all-sibling LDS-staged (FP32)
        for prec in (api.FP64, api.FP32):
            op.set_tuning(prec, 0, bits)
            op.set_apply_precision(prec)
            out[(bits, prec)] = (op.apply(X[:, 0]), op.apply(X))
    op.set_apply_precision(api.FP64)
    op.set_tuning(api.FP64, 0, 7)
    op.set_tuning(api.FP32, 0, 7)
    for k in range(2):
        e64 = relerr(out[(4, api.FP64)][k], out[(0, api.FP64)][k])
        e32 = relerr(out[(4, api.FP32)][k], out[(0, api.FP32)][k])
        w32 = relerr(out[(12, api.FP32)][k], out[(4, api.FP32)][k])
        s32 = relerr(out[(20, api.FP32)][k], out[(4, api.FP32)][k])
        print(f"[cluster] {name} {'batched' if k else 'single'} fp64 {e64:.2e} fp32 {e32:.2e} wide fp32 {w32:.2e} "
              f"all-sibling fp32 {s32:.2e}")
        assert e64 <= 1e-13 and e32 <= 2e-6
        assert w32 <= 1e-12 and s32 <= 1e-12   # the same sums in the same order (plus exact zeros)
    assert relerr(out[(4, api.FP64)][0], orc.apply(X[:, 0])) <= 1e-10


@pytest.mark.parametrize("name", ["efie_ml", "pmchwt_lossy_ml", "pmchwt_1lvl"])
@pytest.mark.parametrize("R", [5, 13, 21])
def test_all_rhs_far_ends_match_per_pass(name, R):
    """FP32 batched matvecs aggregate all right-hand sides into every pass's
    leaf workspaces in one matrix-core product and disaggregate them in
    another (tuning bit 2 on); with it off each pass aggregates and
    disaggregates its own right-hand sides with the vector kernels.  Passes of
    8, 4 and 1 (R = 5, 13, 21) exercise the pass lookup of the aggregation; the
    two agree to FP32 rounding, and both meet the oracle's FP64 operator at the
    FP32 bar."""
    _, op, orc = pair(name)
    X = np.stack([seeded_x(op.n, 1300 + j) for j in range(R)], axis=1)
    op.set_apply_precision(api.FP32)
    out = {}
    for bits in (7, 5):
        op.set_tuning(api.FP32, 0, bits)
        out[bits] = op.apply(X)
    op.set_tuning(api.FP32, 0, 7)
    op.set_apply_precision(api.FP64)
    e = relerr(out[7], out[5])
    e_or = max(relerr(out[7][:, j], orc.apply(X[:, j])) for j in (0, R - 1))
    print(f"[all-rhs] {name} R={R}: all-RHS vs per-pass {e:.2e}, vs oracle {e_or:.2e}")
    assert e <= 2e-6
    assert e_or <= 1e-5
