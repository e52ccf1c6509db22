"""Accuracy mode of the translation operators (tf_mode = TF_COMPLEX_ARG) on the GPU.

The reference evaluates the translation series at the real |kR| even in a lossy medium
(Code/AlgoMLFMM.h:262-265), which drops the medium's decay between distant boxes: on a
6-level lossy PMCHWT tree its matvec is 25-33% off the dense Galerkin product and GMRES
stalls.  With the Hankel functions at the complex kR the same tree is within 2%.

* product = oracle in that mode, <= 1e-10 (the FP64 bar), tables and matvec;
* product matvec against the dense Galerkin product of the committed fixture
  (tests/golden/dense_lossy_r284.npz): <= 2% in accuracy mode, > 15% in parity mode;
* an FP64 GMRES solve of a plane-wave excitation reaches a true residual of 1e-3.
"""
from pathlib import Path

import numpy as np
import pytest

from conftest import relerr
from mlfma_pmchwt_amd import api, geometry
import pyoracle

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "dense_lossy_r284.npz"
_ops = {}


def case(tf_mode):
    if tf_mode not in _ops:
        d = np.load(FIX)
        mesh = geometry.geodesic_sphere(int(d["freq"]), float(d["radius"]))
        eps = tuple(d["eps"])
        print(f"[case] lossy sphere r={float(d['radius'])} geodesic {int(d['freq'])} eps={eps} tf_mode={tf_mode}",
              flush=True)
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=eps[0], eps_r_im=eps[1], tf_mode=tf_mode)
        rng = np.random.default_rng(int(d["seed"]))
        x = rng.standard_normal(op.n) + 1j * rng.standard_normal(op.n)
        _ops[tf_mode] = (mesh, op, x, d["y"], eps)
    return _ops[tf_mode]


def test_accuracy_mode_matches_oracle():
    mesh, op, x, _, eps = case(api.TF_COMPLEX_ARG)
    orc = pyoracle.Oracle(mesh, ie=2, eps=eps, tf_mode=2)
    assert op.max_level == orc.max_level == 5
    for lvl in range(2, orc.max_level + 1):
        assert relerr(op.tf(2, lvl), orc.tf(2, lvl)) < 1e-12, lvl
    e = relerr(op.apply(x), orc.apply(x))
    print(f"[accuracy] product vs oracle (complex kR): {e:.2e}")
    assert e < 1e-10


@pytest.mark.parametrize("tf_mode,lo,hi", [(api.TF_COMPLEX_ARG, 0.0, 0.02), (0, 0.15, 10.0)])
def test_against_dense_galerkin(tf_mode, lo, hi):
    _, op, x, y_dense, _ = case(tf_mode)
    e = relerr(op.apply(x), y_dense)
    print(f"[accuracy] tf_mode={tf_mode}: MLFMA vs dense Galerkin {e:.4f}")
    assert lo <= e <= hi


def test_accuracy_mode_fp64_solve():
    _, op, _, _, _ = case(api.TF_COMPLEX_ARG)
    b = op.excitation(0.0, 0.0, 0.0)
    xs, st = op.solve(b, restart=50, tol_rel=1e-3, itr_max=3000)
    r = relerr(op.apply(xs), b)
    print(f"[accuracy] FP64 GMRES(50): {st['iterations']} iterations, true residual {r:.2e}")
    assert st["converged"] and r <= 1.001e-3
