"""Write tests/golden/c1_dielectric_sphere.npz: the oracle on the full-size C1 configuration.

Dielectric sphere r = 0.5 lambda, eps_r = 4, mu_r = 1, icosphere subdivision 4 (5,120 triangles,
7,680 RWG, 15,360 PMCHWT unknowns), leaf 0.3 lambda (3-level tree), one plane wave along -z
polarised along x (theta = 0, phi = 0, V), GMRES(30) to 1e-3, bistatic RCS at theta_s =
0..180 deg step 1, phi_s = 0 (181 angles).  Stored: y = A x for x ~ CN(0,1) from
default_rng(12345); the oracle's GMRES solution, its iteration count and its 181-angle RCS
(total, V, H dBsm).  The oracle build and solve take a few minutes on 8 cores, so the GPU tests
compare against this file instead of running the oracle at this size:

    python tests/golden/make_c1_fixture.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import pyoracle  # noqa: E402
from mlfma_pmchwt_amd import geometry  # noqa: E402

OUT = Path(__file__).resolve().parent / "c1_dielectric_sphere.npz"
SUBDIV, RADIUS, EPS, SEED, RESTART, TOL = 4, 0.5, (4.0, 0.0), 12345, 30, 1e-3


def mesh():
    return geometry.icosphere(SUBDIV, RADIUS)


def main():
    o = pyoracle.Oracle(mesh(), ie=2, eps=EPS)
    rng = np.random.default_rng(SEED)
    x = rng.standard_normal(o.n) + 1j * rng.standard_normal(o.n)
    y = o.apply(x)
    b = o.rhs(0.0, 0.0, 0.0)
    xs, st = o.gmres(b, restart=RESTART, tol_rel=TOL)
    th = np.arange(181, dtype=np.float64)
    rcs = np.array([o.rcs(xs, t, 0.0, pol=0.0)[0] for t in th])
    np.savez_compressed(OUT, y=y, x_solution=xs, rhs=b, rcs=rcs, theta=th, iterations=st["iterations"],
                        rel_residual=st["rel_residual"], n=o.n, max_level=o.max_level, seed=SEED)
    print(f"wrote {OUT}: {o.n} unknowns, levels 0..{o.max_level}, GMRES({RESTART}) {st['iterations']} iterations")


if __name__ == "__main__":
    main()
