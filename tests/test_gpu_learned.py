"""Learned near-field preconditioner on the GPU.

* the network's forward pass and gradients (FP32 kernels) against the FP64
  CPU oracle, whose gradient is pinned by finite differences;
* the polynomial preconditioner, its features and its Krylov statistics
  against the oracle restatements on the oracle's near field;
* the gain test: a network trained on a family of dielectric spheres
  (eps_r from seeds) predicts coefficients for held-out members, and GMRES
  with those coefficients reaches the 1e-3 true residual in fewer iterations
  than with the block-diagonal preconditioner and than unpreconditioned.
"""
import numpy as np
import pytest

from conftest import relerr, seeded_x
from mlfma_pmchwt_amd import api, geometry, learned
import mlp_oracle
import pyoracle

pytestmark = pytest.mark.gpu


def dense_near(orc):
    n = orc.n
    A = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        ids, v = orc.near_row(i)
        A[i, ids] = v
    return A


def test_network_grad_vs_oracle():
    n_in, width, n_out, B = 20, 256, 14, 24
    m = n_out // 2
    net = api.Mlp(n_in, width, n_out, seed=11)
    rng = np.random.default_rng(2)
    p = (rng.standard_normal(net.n_params) * 0.05).astype(np.float32)
    net.set_params(p)
    x = rng.standard_normal((B, n_in)).astype(np.float32)
    T = rng.standard_normal((B, 12, m)) + 1j * rng.standard_normal((B, 12, m))
    tgt = rng.standard_normal((B, 12)) + 1j * rng.standard_normal((B, 12))
    G = np.einsum("spj,spl->sjl", np.conj(T), T)
    h = np.einsum("spj,sp->sj", np.conj(T), tgt)
    c = np.einsum("sp,sp->s", np.conj(tgt), tgt).real
    loss, g = net.loss_grad(x, G, h, c)
    lo, go = mlp_oracle.loss_grad(p.astype(np.float64), x.astype(np.float64), G, h, c, n_in, width, n_out)
    print(f"[mlp] loss gpu {loss:.6e} oracle {lo:.6e} grad rel {relerr(g, go):.2e}")
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert relerr(g, go) <= 1e-4
    y = net.predict(x)
    yo = mlp_oracle.forward(p.astype(np.float64), x.astype(np.float64), n_in, width, n_out)[2]
    assert relerr(y, yo) <= 1e-5


def test_network_training_reaches_optimum():
    """Training on problems whose optimum is known drives the loss to it."""
    n_in, width, m = 20, 64, 3
    net = api.Mlp(n_in, width, 2 * m, seed=1)
    rng = np.random.default_rng(9)
    B = 16
    x = rng.standard_normal((B, n_in)).astype(np.float32)
    a_true = (x[:, :m] * 0.5 + 1.0) + 1j * (x[:, m:2 * m] * 0.3)   # smooth map of the inputs
    T = rng.standard_normal((B, 10, m)) + 1j * rng.standard_normal((B, 10, m))
    tgt = np.einsum("spj,sj->sp", T, a_true)                           # exactly representable
    G = np.einsum("spj,spl->sjl", np.conj(T), T)
    h = np.einsum("spj,sp->sj", np.conj(T), tgt)
    c = np.einsum("sp,sp->s", np.conj(tgt), tgt).real
    hist = net.train(x, G, h, c, steps=1500, lr=2e-3)
    print(f"[mlp] loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.02 * hist[0]


@pytest.fixture(scope="module")
def pm_case():
    mesh = geometry.icosphere(2, 0.5)
    eps = (4.0, -0.2)
    op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps[0], eps_r_im=eps[1])
    orc = pyoracle.Oracle(mesh, ie=2, eps=eps)
    return op, orc, dense_near(orc), eps


def test_poly_apply_vs_oracle(pm_case):
    op, orc, An, _ = pm_case
    rng = np.random.default_rng(4)
    alpha = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    op.set_precond_poly(alpha)
    x = seeded_x(op.n, 5)
    y = op.precond_apply(x, kind=api.PRECOND_POLY)
    yo = mlp_oracle.poly_apply(An, alpha, x)
    assert relerr(y, yo) <= 1e-11


def test_features_vs_oracle(pm_case):
    op, orc, An, eps = pm_case
    f = op.precond_features()
    fo = mlp_oracle.features(An, complex(*eps), True)
    print("[features]", np.round(f, 4))
    assert np.allclose(f, fo, rtol=1e-5, atol=1e-5)


def test_statistics_vs_oracle(pm_case):
    op, orc, An, _ = pm_case
    b = op.excitation(0.0, 0.0, 0.0)
    G, h, c = op.precond_statistics(b, steps=20, degree=4)
    Go, ho, co = mlp_oracle.statistics(orc.apply, An, b, 20, 4)
    assert c == co
    assert relerr(G, Go) <= 1e-9 and relerr(h, ho) <= 1e-9
    # the fitted coefficients reduce the pair residual far below the trivial choice
    a = learned.fit_coefficients(G, h, c)
    assert learned.relative_loss(a, G, h, c) < 0.5 * learned.relative_loss(np.eye(5)[0], G, h, c)


FAMILY_TRAIN = list(range(12))
FAMILY_TEST = [100, 101, 102]


def family_member(seed):
    rng = np.random.default_rng(seed)
    eps = (float(rng.uniform(2.0, 8.0)), -float(rng.uniform(0.0, 0.3)))
    return eps


def test_learned_preconditioner_gain():
    mesh = geometry.icosphere(2, 0.5)
    feats, Gs, hs, cs = [], [], [], []
    for s in FAMILY_TRAIN:
        eps = family_member(s)
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps[0], eps_r_im=eps[1])
        f, G, h, c = learned.collect(op, op.excitation(0.0, 0.0, 0.0), steps=30)
        feats.append(f), Gs.append(G), hs.append(h), cs.append(c)
        op.close()
    lp = learned.LearnedPreconditioner(seed=3)
    hist = lp.fit(np.array(feats), np.array(Gs), np.array(hs), np.array(cs), steps=1500, lr=2e-3)
    print(f"[learned] training loss {hist[0]:.3f} -> {hist[-1]:.3f}")
    assert hist[-1] < hist[0]
    for s in FAMILY_TEST:
        eps = family_member(s)
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps[0], eps_r_im=eps[1], verbose=1)
        orc = pyoracle.Oracle(mesh, ie=2, eps=eps)
        b = op.excitation(0.0, 30.0, 0.0)   # an incidence the training runs did not use
        runs = {}
        for name, pc in (("none", 0), ("block_diag", 1)):
            x, st = op.solve(b, tol_rel=1e-3, restart=30, precond=pc, itr_max=3000)
            runs[name] = (st, x)
        lp.apply_to(op)
        x, st = op.solve(b, tol_rel=1e-3, restart=30, precond=api.PRECOND_POLY, itr_max=3000)
        runs["learned"] = (st, x)
        print(f"[learned] eps={eps[0]:.2f}{eps[1]:+.2f}j iterations: " +
              ", ".join(f"{k}={v[0]['iterations']}" for k, v in runs.items()))
        st, x = runs["learned"]
        assert st["converged"] and st["precond_fallbacks"] == 0
        assert relerr(orc.apply(x), b) < 1e-3          # true residual, oracle operator
        assert st["iterations"] < runs["none"][0]["iterations"]
        assert st["iterations"] < runs["block_diag"][0]["iterations"]
        op.close()


def test_batched_solve_with_polynomial(pm_case):
    """precond = 2 in the lockstep batched GMRES: same iterations as single solves."""
    op, orc, An, _ = pm_case
    b0 = op.excitation(0.0, 0.0, 0.0)
    G, h, c = op.precond_statistics(b0, steps=30, degree=learned.DEGREE)
    op.set_precond_poly(learned.fit_coefficients(G, h, c))
    inc = [(0.0, 20.0, 0.0), (90.0, 60.0, 0.0), (0.0, 120.0, 45.0)]
    B = np.stack([op.excitation(*a) for a in inc], axis=1)
    X, st = op.solve_batch(B, tol_rel=1e-3, restart=30, precond=api.PRECOND_POLY)
    for j in range(B.shape[1]):
        x1, s1 = op.solve(B[:, j], tol_rel=1e-3, restart=30, precond=api.PRECOND_POLY)
        print(f"[poly batch] {inc[j]} batch it={st[j]['iterations']} single it={s1['iterations']}")
        assert st[j]["converged"] and abs(st[j]["iterations"] - s1["iterations"]) <= 1
        assert relerr(orc.apply(X[:, j]), B[:, j]) < 1e-3


C5_TRAIN = range(40)
C5_TEST = [1000, 1001, 1002]


def test_learned_gain_c5_family():
    """The C5 family (ellipsoid (3, 2, 1) lambda scaled by 0.3, geodesic 16: 7,680 RWG; eps_r from
    default_rng(seed) U[2, 8]): the network trained on seeds 0..39 (one seeded incidence each,
    pairs from each problem's own GMRES) against, on held-out seeds 1000..1002 and an incidence
    of its own: no preconditioner, the block-diagonal one, the Neumann series (all coefficients 1)
    and the training-mean polynomial (the network's starting output bias); and the inner
    near-field solve for reference.  The learned polynomial converges (true residual checked on
    the oracle operator) in fewer iterations than none, block-diagonal and Neumann, and is not
    worse than the training mean by more than 10%."""
    from mlfma_pmchwt_amd import configs
    mesh = configs.c5_mesh(16, scale=0.3)
    rng = np.random.default_rng(5)
    feats, Gs, hs, cs = [], [], [], []
    for s in C5_TRAIN:
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=configs.c5_eps(s))
        inc = (float(rng.integers(2)) * 90.0, float(rng.uniform(0.0, 180.0)), 0.0)
        f, G, h, c = learned.collect(op, op.excitation(*inc), steps=30)
        feats.append(f), Gs.append(G), hs.append(h), cs.append(c)
        op.close()
    lp = learned.LearnedPreconditioner(seed=7)
    hist = lp.fit(np.array(feats), np.array(Gs), np.array(hs), np.array(cs), steps=2000, lr=2e-3)
    print(f"[learned C5] training loss {hist[0]:.3f} -> {hist[-1]:.3f}")
    assert hist[-1] < hist[0]
    mean_alpha = np.stack([learned.fit_coefficients(G, h, c) for G, h, c in zip(Gs, hs, cs)]).mean(axis=0)
    for s in C5_TEST:
        eps = configs.c5_eps(s)
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps)
        b = op.excitation(90.0, 37.0, 0.0)
        runs = {}
        for name in ("none", "block_diag", "neumann", "mean_poly", "learned", "inner10"):
            pc = {"none": 0, "block_diag": 1, "inner10": 3}.get(name, 2)
            if name == "neumann":
                op.set_precond_poly(np.ones(learned.DEGREE + 1))
            elif name == "mean_poly":
                op.set_precond_poly(mean_alpha)
            elif name == "learned":
                lp.apply_to(op)
            x, st = op.solve(b, tol_rel=1e-3, restart=30, precond=pc, itr_max=4000)
            runs[name] = (st, x)
        print(f"[learned C5] seed {s} eps {eps:.2f} iterations: " +
              ", ".join(f"{k}={v[0]['iterations']}" for k, v in runs.items()))
        st, x = runs["learned"]
        assert st["converged"] and st["precond_fallbacks"] == 0
        orc = pyoracle.Oracle(mesh, ie=2, eps=(eps, 0.0))
        assert relerr(orc.apply(x), b) < 1.001e-3
        for other in ("none", "block_diag", "neumann"):
            assert st["iterations"] < runs[other][0]["iterations"], other
        assert st["iterations"] <= 1.1 * runs["mean_poly"][0]["iterations"]
        op.close()
