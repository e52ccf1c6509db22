"""FP32 mode (far field stored and accumulated in FP32, near-field entries
stored in FP32) against the FP64 oracle.

The reduced-precision operator must stay within 1e-3 of the FP64 matvec (the
bar for reduced precision); in practice FP32 tables give about 1e-7.  Solves in
FP32 mode run their Krylov iterations on the FP32 operator and judge (and
refine) with the FP64 operator's true residual.
"""
import numpy as np
import pytest

from conftest import relerr, seeded_x
from mlfma_pmchwt_amd import api, geometry
import pyoracle

pytestmark = pytest.mark.gpu

CASES = {
    "cfie_ml": ((2, 1.5), api.CFIE, (1.0, 0.0)),
    "pmchwt_lossy_ml": ((2, 1.5), api.PMCHWT, (2.5, -0.1)),
    "pmchwt_1lvl": ((2, 0.5), api.PMCHWT, (4.0, 0.0)),
}
_cache = {}


def pair(name):
    if name not in _cache:
        (s, r), ie, eps = CASES[name]
        mesh = geometry.icosphere(s, r)
        print(f"[case] {name}: icosphere s={s} r={r} ie={ie} eps={eps} far FP32", flush=True)
        op = api.Operator(mesh, integral_equation=ie, eps_r_re=eps[0], eps_r_im=eps[1], far_precision=api.FP32,
                          near_precision=api.FP32)
        orc = pyoracle.Oracle(mesh, ie=ie, eps=eps)
        _cache[name] = (mesh, op, orc)
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_matvec_error(name):
    _, op, orc = pair(name)
    for seed in (1, 2):
        x = seeded_x(op.n, seed)
        yo = orc.apply(x)
        op.set_apply_precision(api.FP32)
        e32 = relerr(op.apply(x), yo)
        op.set_apply_precision(api.FP64)
        e64 = relerr(op.apply(x), yo)
        op.set_apply_precision(api.FP32)
        print(f"[fp32] {name} seed={seed} FP32 rel={e32:.3e} FP64 rel={e64:.3e}")
        assert e64 <= 1e-10
        assert e32 <= 1e-5
        assert e32 > 1e-12   # the FP32 tables are really in use


def test_fp32_needs_fp32_tables(sphere_small):
    op = api.Operator(sphere_small, integral_equation=api.CFIE)
    with pytest.raises(RuntimeError):
        op.set_apply_precision(api.FP32)


def test_fp32_far_only_and_near_only(sphere_ml):
    """Each half alone: the error of the mixed operator splits into its parts."""
    orc = pyoracle.Oracle(sphere_ml, ie=api.CFIE)
    x = seeded_x(orc.n, 7)
    yo = orc.apply(x)
    errs = {}
    for far, near in ((api.FP32, api.FP64), (api.FP64, api.FP32), (api.FP32, api.FP32)):
        op = api.Operator(sphere_ml, integral_equation=api.CFIE, far_precision=far, near_precision=near)
        op.set_apply_precision(api.FP32)
        errs[(far, near)] = relerr(op.apply(x), yo)
        op.close()
    print(f"[fp32 parts] far only {errs[(1, 0)]:.2e} near only {errs[(0, 1)]:.2e} both {errs[(1, 1)]:.2e}")
    assert all(1e-12 < e <= 1e-5 for e in errs.values())


@pytest.mark.parametrize("name", ["cfie_ml", "pmchwt_1lvl"])
def test_fp32_solve_refined(name):
    _, op, orc = pair(name)
    op.set_apply_precision(api.FP32)
    b = op.excitation(0.0, 0.0, 0.0)
    for tol in (1e-3, 1e-6):
        x, st = op.solve(b, tol_rel=tol, restart=30)
        # the reported residual is the FP64 operator's
        r = relerr(orc.apply(x), b)
        print(f"[fp32 solve] {name} tol={tol} it={st['iterations']} mv={st['matvecs']} "
              f"fp64 mv={st['fp64_matvecs']} res={st['rel_residual']:.3e} oracle res={r:.3e}")
        assert st["converged"] and st["rel_residual"] < tol
        assert abs(r - st["rel_residual"]) <= 1e-9 + 1e-6 * r
    # the solution matches the FP64 solve to the tolerance's order
    x64, _ = orc.gmres(b, tol_rel=1e-6, restart=30)
    assert relerr(x, x64) <= 1e-4
    # RCS of the FP32-mode solution within 0.1 dB of the FP64 solution's
    for th in (0.0, 45.0, 90.0, 180.0):
        a = op.rcs(x, th, 0.0)[0]
        o, _ = orc.rcs(x64, th, 0.0)
        assert np.max(np.abs(a - o)) <= 0.1
