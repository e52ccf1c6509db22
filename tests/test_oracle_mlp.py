"""Pins the CPU oracle of the preconditioner network by central finite
differences of its loss (float64), so that the GPU gradients can be checked
against it."""
import numpy as np

import mlp_oracle


def test_oracle_gradient_finite_differences():
    rng = np.random.default_rng(3)
    n_in, width, n_out, B = 5, 7, 6, 4
    m = n_out // 2
    npar = width * n_in + width + width * width + width + n_out * width + n_out
    p = rng.standard_normal(npar) * 0.5
    x = rng.standard_normal((B, n_in))
    # statistics of actual least-squares data: basis T (pairs x m) and targets e per sample
    T = rng.standard_normal((B, 9, m)) + 1j * rng.standard_normal((B, 9, m))
    tgt = rng.standard_normal((B, 9)) + 1j * rng.standard_normal((B, 9))
    G = np.einsum("spj,spl->sjl", np.conj(T), T)
    h = np.einsum("spj,sp->sj", np.conj(T), tgt)
    c = np.einsum("sp,sp->s", np.conj(tgt), tgt).real
    loss, g = mlp_oracle.loss_grad(p, x, G, h, c, n_in, width, n_out)
    fd = np.zeros_like(p)
    eps = 1e-6
    for k in range(npar):
        e = np.zeros_like(p)
        e[k] = eps
        fd[k] = (mlp_oracle.loss_grad(p + e, x, G, h, c, n_in, width, n_out)[0]
                 - mlp_oracle.loss_grad(p - e, x, G, h, c, n_in, width, n_out)[0]) / (2 * eps)
    assert np.linalg.norm(g - fd) <= 1e-7 * np.linalg.norm(fd)
    # the per-sample loss is the relative residual |T a - e|^2 / |e|^2 of the data
    a = rng.standard_normal((B, m)) + 1j * rng.standard_normal((B, m))
    q = np.einsum("sj,sjl,sl->s", np.conj(a), G, a).real - 2 * np.einsum("sj,sj->s", np.conj(a), h).real
    direct = np.linalg.norm(np.einsum("spj,sj->sp", T, a) - tgt, axis=1) ** 2 / c
    assert np.allclose((q + c) / c, direct, rtol=1e-12)


def test_poly_apply_recurrence():
    """alpha = (1, 1) is the two-term Neumann series 2 D^-1 - D^-1 A D^-1 (t_0 + t_1)."""
    rng = np.random.default_rng(5)
    n = 12
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8 * np.eye(n)
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Dinv = np.diag(1 / np.diag(A))
    expect = 2 * Dinv @ r - Dinv @ A @ Dinv @ r
    assert np.allclose(mlp_oracle.poly_apply(A, [1.0, 1.0], r), expect, rtol=1e-13, atol=1e-13)


def test_probe_is_unit_modulus_and_seeded():
    z = mlp_oracle.probe(1000)
    assert np.allclose(np.abs(z), 1.0)
    assert abs(np.mean(z)) < 0.1   # phases spread over the circle
    assert not np.allclose(z, mlp_oracle.probe(1000, seed=7))
