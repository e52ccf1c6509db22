"""Split-pair BF16 near products of the operator (near_precision = BF16).

* kernel: the batched near product of the polynomial step t_1 = t_0 - D^-1 Z t_0
  (t_0 = D^-1 r) against a numpy restatement of the split: f = FP32 value,
  hi = bf16(f), lo = bf16(f - hi) for entries and vectors, product
  Zh Th + Zh Tl + Zl Th accumulated in FP64; only the FP32 accumulation
  differs (~1e-6).  The same restatement without the Zh Tl term, or with the
  single BF16 product, is far outside the tolerance, and so is the exact
  product: the three partial products and the rounding are what runs;
* operator: the FP32-mode batched matvec with split-pair near products stays
  within 1e-4 of the oracle's FP64 operator;
* solves: batched FP32-mode GMRES with these products reaches the oracle's
  1e-3 true residual in about the iterations of the FP32 near field.
"""
import numpy as np
import pytest

from conftest import relerr, seeded_x
from mlfma_pmchwt_amd import api, geometry
import pyoracle

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float32."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split(z):
    """(hi, lo) complex pairs of the FP32 value of z, each part BF16."""
    def one(a):
        f = np.asarray(a, dtype=np.float32)
        h = bf16(f)
        return h.astype(np.float64), bf16(f - h).astype(np.float64)   # f - h is exact in FP32
    hr, lr = one(z.real)
    hi, li = one(z.imag)
    return hr + 1j * hi, lr + 1j * li


def dense_near(orc):
    A = np.zeros((orc.n, orc.n), dtype=np.complex128)
    for i in range(orc.n):
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
        print(f"[case] {name}: icosphere s={s} r={r} ie={ie} eps={eps} split-pair near", flush=True)
        op = api.Operator(mesh, integral_equation=ie, eps_r_re=eps[0], eps_r_im=eps[1], far_precision=api.FP32,
                          near_precision=api.BF16, max_rhs=40)
        orc = pyoracle.Oracle(mesh, ie=ie, eps=eps)
        _cache[name] = (op, orc, dense_near(orc))
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("R", [2, 12, 40])
def test_split_step_matches_restatement(name, R):
    op, orc, An = case(name)
    d = np.diag(An)
    Rm = np.stack([seeded_x(op.n, 700 + j) for j in range(R)], axis=1)
    op.set_precond_poly(np.array([0.0, 1.0]))
    op.set_apply_precision(api.FP32)
    Y = op.precond_apply(Rm, kind=api.PRECOND_POLY)
    op.set_apply_precision(api.FP64)
    T0 = Rm / d[:, None]
    Zh, Zl = split(An)
    Th, Tl = split(T0)
    step = lambda P: T0 - P / d[:, None]
    e = relerr(Y, step(Zh @ Th + Zh @ Tl + Zl @ Th))
    e_no_cross = relerr(Y, step(Zh @ Th + Zl @ Th))
    e_single = relerr(Y, step(Zh @ Th))
    e_exact = relerr(Y, step(An @ T0))
    print(f"[split] {name} R={R}: restatement {e:.2e}, without Zh Tl {e_no_cross:.2e}, "
          f"single BF16 {e_single:.2e}, exact {e_exact:.2e}")
    assert e <= 3e-6
    assert e_no_cross > 10 * 3e-6 and e_single > 10 * 3e-6
    assert 1e-7 < e_exact <= 1e-4     # about 2^-17 of the product, not the FP32 rounding


def test_split_batched_matvec_near_oracle():
    op, orc, An = case("pmchwt_1lvl")
    X = np.stack([seeded_x(op.n, 900 + j) for j in range(5)], axis=1)
    op.set_apply_precision(api.FP32)
    Y = op.apply(X)
    op.set_apply_precision(api.FP64)
    for j in range(X.shape[1]):
        e = relerr(Y[:, j], orc.apply(X[:, j]))
        print(f"[split matvec] column {j}: vs oracle FP64 {e:.2e}")
        assert e <= 1e-4


def test_split_batched_solves():
    op, orc, An = case("pmchwt_1lvl")
    inc = [(0.0, 20.0, 0.0), (90.0, 60.0, 0.0), (0.0, 120.0, 45.0), (90.0, 170.0, 10.0)]
    B = np.stack([op.excitation(*a) for a in inc], axis=1)
    op.set_apply_precision(api.FP32)
    X, st = op.solve_batch(B, tol_rel=1e-3, restart=30, itr_max=2000)
    op.set_apply_precision(api.FP64)
    ref = api.Operator(geometry.icosphere(2, 0.5), integral_equation=api.PMCHWT, eps_r_re=4.0, eps_r_im=-0.2,
                       far_precision=api.FP32, near_precision=api.FP32, max_rhs=8)
    ref.set_apply_precision(api.FP32)
    _, st32 = ref.solve_batch(B, tol_rel=1e-3, restart=30, itr_max=2000)
    ref.close()
    for j in range(B.shape[1]):
        print(f"[split solve] {inc[j]}: split it={st[j]['iterations']} fp32 it={st32[j]['iterations']}")
        assert st[j]["converged"]
        assert relerr(orc.apply(X[:, j]), B[:, j]) < 1e-3
        assert st[j]["iterations"] <= 1.1 * st32[j]["iterations"] + 2


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("R", [5, 21, 40])
def test_split_disaggregation_matches_fp32(name, R):
    """With near_precision = BF16 the all-RHS disaggregation also runs on split
    pairs (patterns split once at set-up, the effective matrices when staged).
    Against the per-pass FP32 vector disaggregation (tuning bit 2 off; the near
    products are the same split-pair ones in both) it differs by about 2^-17 of
    the far-field terms, far below FP32 mode's 1e-4 bar against the oracle."""
    op, orc, An = case(name)
    X = np.stack([seeded_x(op.n, 1500 + j) for j in range(R)], axis=1)
    op.set_apply_precision(api.FP32)
    out = {}
    for bits in (7, 5):
        op.set_tuning(api.FP32, 0, bits)
        out[bits] = op.apply(X)
    op.set_tuning(api.FP32, 0, 7)
    op.set_apply_precision(api.FP64)
    e = relerr(out[7], out[5])
    e_or = max(relerr(out[7][:, j], orc.apply(X[:, j])) for j in (0, R - 1))
    print(f"[split disagg] {name} R={R}: vs per-pass FP32 {e:.2e}, vs oracle {e_or:.2e}")
    assert e <= 5e-5
    assert e_or <= 1e-4
