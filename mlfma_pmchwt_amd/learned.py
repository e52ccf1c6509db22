"""Learned near-field preconditioner.

The preconditioner is a polynomial in the point-Jacobi-scaled near field,

    M^-1 r = sum_k alpha_k t_k,   t_0 = D^-1 r,   t_{k+1} = t_k - D^-1 Z_near t_k,

applied on the GPU with one near-field product per degree (solve precond = 2).
Good coefficients depend on the problem (a polynomial fitted for eps_r = 4 can
stall at eps_r = 8), and fitting them needs Krylov data of the problem itself.
The coefficient network predicts them for an unseen problem from features of
its near-field blocks (spectral moments of D^-1 Z_near), having been trained
on residual/correction pairs (r = A v, e = v) that the solver's own GMRES runs
produced on a family of training problems.  The loss of a problem is the
relative least-squares error of the preconditioner on its pairs,
(a^H G a - 2 Re a^H h + c) / c, with (G, h, c) accumulated on the device.

No reference counterpart: the reference solves unpreconditioned
(Code/MLFMM.cpp:370-397).  The plain block-diagonal preconditioner stays
available (precond = 1), and a solve drops a polynomial preconditioner that
stalls and continues unpreconditioned.
"""
from __future__ import annotations

import numpy as np

from . import api

DEGREE = 6


def fit_coefficients(G: np.ndarray, h: np.ndarray, c: float, ridge: float = 1e-10) -> np.ndarray:
    """Least-squares coefficients of one problem: argmin a^H G a - 2 Re a^H h + c."""
    m = G.shape[0]
    lam = ridge * np.trace(G).real / m
    return np.linalg.solve(G + lam * np.eye(m), h)


def relative_loss(alpha: np.ndarray, G: np.ndarray, h: np.ndarray, c: float) -> float:
    a = np.asarray(alpha, dtype=np.complex128)
    return float((np.vdot(a, G @ a).real - 2.0 * np.vdot(a, h).real + c) / c)


def collect(op: api.Operator, rhs: np.ndarray, steps: int = 30, degree: int = DEGREE):
    """Features and coefficient statistics of one training problem."""
    f = op.precond_features()
    G, h, c = op.precond_statistics(rhs, steps=steps, degree=degree)
    return f, G, h, c


def _as_outputs(alpha: np.ndarray) -> np.ndarray:
    alpha = np.asarray(alpha, dtype=np.complex128)
    return np.concatenate([alpha.real, alpha.imag]).astype(np.float32)


def _as_alpha(y: np.ndarray) -> np.ndarray:
    m = y.shape[-1] // 2
    return y[..., :m].astype(np.float64) + 1j * y[..., m:].astype(np.float64)


class LearnedPreconditioner:
    """Coefficient network (features -> alpha) trained on the GPU."""

    def __init__(self, degree: int = DEGREE, width: int = 256, seed: int = 0):
        self.degree = degree
        self.net = api.Mlp(api.PRECOND_FEATURES, width, 2 * (degree + 1), seed)

    def fit(self, features, G, h, c, steps: int = 2000, lr: float = 2e-3) -> np.ndarray:
        """Trains on P problems: features (P x 20), G (P x m x m), h (P x m), c (P)."""
        X = np.asarray(features, dtype=np.float32)
        mean = X.mean(axis=0)
        std = X.std(axis=0)
        std = np.where(std > 1e-6 * (np.abs(mean) + 1.0), std, 1.0).astype(np.float32)
        self.net.set_normalization(mean, std)
        # output bias starts at the mean of the per-problem optimal coefficients
        best = np.stack([fit_coefficients(G[p], h[p], c[p]) for p in range(len(X))])
        params = self.net.params()
        params[-2 * (self.degree + 1):] = _as_outputs(best.mean(axis=0))
        self.net.set_params(params)
        return self.net.train(X, np.asarray(G), np.asarray(h), np.asarray(c, dtype=np.float64), steps=steps, lr=lr)

    def predict(self, features) -> np.ndarray:
        return _as_alpha(self.net.predict(np.atleast_2d(features)))

    def apply_to(self, op: api.Operator) -> np.ndarray:
        """Predicts this operator's coefficients and installs them (solve with precond = 2)."""
        alpha = self.predict(op.precond_features())[0]
        op.set_precond_poly(alpha)
        return alpha

    def state(self) -> dict:
        return {"degree": self.degree, "params": self.net.params()}
