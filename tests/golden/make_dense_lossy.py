"""Write tests/golden/dense_lossy_r284.npz: the dense Galerkin product y = Z x of a lossy
PMCHWT sphere whose MLFMA tree has 6 levels.

Sphere r = 2.84 lambda (geodesic frequency 12: 4,320 RWG, 8,640 unknowns), eps_r = 2.5 - 0.1j,
x ~ CN(0, 1) from numpy.random.default_rng(12345).  Z comes from the oracle's dense assembly
(the same L / K operator routines as its near field; those are pinned by Duffy quadrature,
Galerkin symmetry and Mie series).  A test compares the oracle's and the product's MLFMA
matvecs, in the reference's translation series and in the complex-argument accuracy mode,
against it.  About three minutes on 8 cores:

    python tests/golden/make_dense_lossy.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import pyoracle  # noqa: E402
from mlfma_pmchwt_amd import geometry  # noqa: E402

OUT = Path(__file__).resolve().parent / "dense_lossy_r284.npz"
RADIUS, FREQ, EPS, SEED = 2.84, 12, (2.5, -0.1), 12345


def mesh():
    return geometry.geodesic_sphere(FREQ, RADIUS)


def main():
    o = pyoracle.Oracle(mesh(), ie=2, eps=EPS, mode=3)
    rng = np.random.default_rng(SEED)
    x = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    y = o.dense() @ x
    np.savez_compressed(OUT, y=y, n=o.n, max_level=o.max_level, radius=RADIUS, freq=FREQ,
                        eps=np.array(EPS), seed=SEED)
    print(f"wrote {OUT} ({o.n} unknowns, tree levels 0..{o.max_level})")


if __name__ == "__main__":
    main()
