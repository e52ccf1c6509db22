"""CPU oracle of the preconditioner network and of the polynomial preconditioner.

TEST INFRASTRUCTURE ONLY: imported by tests as the checker of the GPU
kernels, never by the product path.  Neither has a reference counterpart (the
reference solves unpreconditioned, Code/MLFMM.cpp:370-397); both are pinned by
mathematics instead: the network gradient by central finite differences of
the loss, the polynomial by its defining recurrence on a dense near-field
matrix.

Network: x -> tanh(W1 x + b1) -> tanh(W2 h1 + b2) -> W3 h2 + b3 = y, the
parameter vector laid out as W1 (width x n_in), b1, W2, b2, W3 (n_out x width),
b3.  y holds the real parts then the imaginary parts of m = n_out / 2 complex
coefficients a; per sample l = (a^H G a - 2 Re a^H h + c) / c; loss = mean l.
"""
from __future__ import annotations

import numpy as np


def unpack(p, n_in, width, n_out):
    o = 0
    out = []
    for shape in ((width, n_in), (width,), (width, width), (width,), (n_out, width), (n_out,)):
        size = int(np.prod(shape))
        out.append(np.asarray(p[o:o + size], dtype=np.float64).reshape(shape))
        o += size
    assert o == len(p)
    return out


def forward(p, x, n_in, width, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, width, n_out)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h1, h2, h2 @ W3.T + b3


def loss_grad(p, x, G, h, c, n_in, width, n_out):
    """Mean loss and its gradient (float64)."""
    x = np.asarray(x, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, width, n_out)
    h1, h2, y = forward(p, x, n_in, width, n_out)
    m = n_out // 2
    B = x.shape[0]
    a = y[:, :m] + 1j * y[:, m:]
    Ga = np.einsum("sjl,sl->sj", G, a)
    q = np.einsum("sj,sj->s", np.conj(a), Ga).real - 2.0 * np.einsum("sj,sj->s", np.conj(a), h).real
    loss = np.mean((q + c) / c)
    g = Ga - h                      # d l / d conj(a), up to the factor 1 / c
    dy = np.concatenate([2.0 * g.real, 2.0 * g.imag], axis=1) / c[:, None] / B
    dW3 = dy.T @ h2
    db3 = dy.sum(axis=0)
    dz2 = (dy @ W3) * (1.0 - h2 ** 2)
    dW2 = dz2.T @ h1
    db2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (1.0 - h1 ** 2)
    dW1 = dz1.T @ x
    db1 = dz1.sum(axis=0)
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return float(loss), grad


def poly_apply(A_near: np.ndarray, alpha, r: np.ndarray) -> np.ndarray:
    """sum_k alpha_k t_k with t_0 = D^-1 r, t_{k+1} = t_k - D^-1 A_near t_k."""
    d = np.diag(A_near)
    t = r / d
    y = alpha[0] * t
    for a in alpha[1:]:
        t = t - (A_near @ t) / d
        y = y + a * t
    return y


def probe(n: int, seed: int = 12345) -> np.ndarray:
    """The device's unit-modulus probe vector (hashed phases)."""
    i = np.arange(n, dtype=np.uint64)
    h = (i * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)
    h = h ^ np.uint64((seed * 0x9E3779B9) & 0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    ph = 2.0 * np.pi * (h.astype(np.float64) / 4294967296.0)
    return np.cos(ph) + 1j * np.sin(ph)


def features(A_near: np.ndarray, eps_r: complex, pmchwt: bool) -> np.ndarray:
    n = A_near.shape[0]
    d = np.diag(A_near)
    z = probe(n)
    v = z.copy()
    mu = []
    for _ in range(8):
        v = (A_near @ v) / d
        mu.append(np.vdot(z, v) / n)
    mu = np.array(mu)
    density = np.count_nonzero(A_near) / n   # near-field entries per row
    return np.concatenate([mu.real, mu.imag, [eps_r.real, eps_r.imag, 1.0 if pmchwt else 0.0, np.log10(density)]])


def statistics(apply, A_near: np.ndarray, rhs: np.ndarray, steps: int, degree: int):
    """(G, h, c) from `steps` Arnoldi pairs (A v_k, v_k) started at rhs (CGS2);
    apply(v) = A v (the oracle's MLFMA operator)."""
    n = len(rhs)
    V = np.zeros((n, steps + 1), dtype=np.complex128)
    V[:, 0] = rhs / np.linalg.norm(rhs)
    W = np.zeros((n, steps), dtype=np.complex128)
    kept = 0
    for k in range(steps):
        W[:, k] = apply(V[:, k])
        kept += 1
        w = W[:, k].copy()
        for _ in range(2):
            w -= V[:, :k + 1] @ (V[:, :k + 1].conj().T @ w)
        hn = np.linalg.norm(w)
        if not hn > 1e-14:
            break
        V[:, k + 1] = w / hn
    d = np.diag(A_near)
    m = degree + 1
    T = [W[:, :kept] / d[:, None]]
    for _ in range(degree):
        T.append(T[-1] - (A_near @ T[-1]) / d[:, None])
    T = np.stack(T, axis=2)   # n x pairs x m
    G = np.einsum("ipj,ipl->jl", np.conj(T), T)
    h = np.einsum("ipj,ip->j", np.conj(T), V[:, :kept])
    return G, h, float(kept)
