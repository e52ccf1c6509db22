"""The five benchmark workloads as meshes, media and incidences.

C1  dielectric sphere r = 0.5 lambda, eps_r = 4, icosphere subdivision 4 (15,360 PMCHWT
    unknowns, 3-level tree), one plane wave along -z polarised along x, GMRES(30) to 1e-3,
    bistatic RCS at theta_s = 0..180 deg (181 angles), phi_s = 0.
C2  PEC sphere r = 4 lambda, CFIE alpha = 0.5, 6-level tree, single right-hand side
    (geodesic frequency 100: 300,000 RWG; a coarse mesh of the same sphere has the same tree).
C3  lossy dielectric sphere r = 2 lambda, eps_r = 2.5 - 0.1j, PMCHWT, 5-level tree, 64
    incidences (32 theta x V/H polarisation, phi = 0) (geodesic frequency 50: 150,000 unknowns).
C4  16 objects from default_rng(4): ellipsoids (semi-axes U[0.5, 3]), cubes (side U[0.5, 3]) and
    cylinders (radius U[0.3, 1], height U[1, 4]), eps_r U[1.5, 6], mesh edge lambda/10, 8
    incidences each (theta = i 180/7 deg, phi = 0, V), 2-6 levels.
C5  dielectric ellipsoid (3, 2, 1) lambda, eps_r U[2, 8] from default_rng(seed) (training seeds
    0..199, held-out seeds 1000..1015), PMCHWT, 6-level tree, 64 incidences as in C3
    (geodesic frequency 71: 151,230 RWG, 302,460 unknowns).

Lengths are in free-space wavelengths (the operators run at 299,792,458 Hz, lambda = 1 m).
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry

# ---------------------------------------------------------------- C1
C1_EPS = (4.0, 0.0)


def c1_mesh(subdivisions: int = 4) -> geometry.Mesh:
    return geometry.icosphere(subdivisions, 0.5)


C1_INCIDENCE = (0.0, 0.0, 0.0)                  # pol V, theta 0, phi 0: k = -z, E along -x
C1_BISTATIC_THETA = np.arange(181, dtype=np.float64)

# ---------------------------------------------------------------- C2
C2_RADIUS, C2_ALPHA = 4.0, 0.5


def c2_mesh(frequency: int = 100) -> geometry.Mesh:
    return geometry.geodesic_sphere(frequency, C2_RADIUS)


# ---------------------------------------------------------------- C3 / C5 incidences
def incidences_64():
    """32 theta values 0..180 deg (phi = 0) for each of the V and H polarisations."""
    thetas = np.arange(32) * 180.0 / 31.0
    return [(pol, float(th), 0.0) for pol in (0.0, 90.0) for th in thetas]


C3_RADIUS, C3_EPS = 2.0, (2.5, -0.1)


def c3_mesh(frequency: int = 50) -> geometry.Mesh:
    return geometry.geodesic_sphere(frequency, C3_RADIUS)


# ---------------------------------------------------------------- C4
def _edge_frequency(longest: float, edge: float) -> int:
    # a class-I geodesic sphere of radius a and frequency n has edges of about 1.1 a / n
    return max(2, int(math.ceil(1.1 * longest / edge)))


def c4_objects(seed: int = 4, count: int = 16, scale: float = 1.0, edge: float = 0.1):
    """The seeded object mix: a list of dicts (kind, size parameters, eps_r, mesh).  `scale`
    shrinks every object (and keeps the edge length), for small versions of the mix."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = ("ellipsoid", "cube", "cylinder")[int(rng.integers(3))]
        if kind == "ellipsoid":
            axes = rng.uniform(0.5, 3.0, size=3) * scale
            mesh = geometry.ellipsoid(tuple(axes), frequency=_edge_frequency(float(axes.max()), edge))
            size = {"semi_axes": [float(a) for a in axes]}
        elif kind == "cube":
            side = float(rng.uniform(0.5, 3.0)) * scale
            mesh = geometry.cube(side, max(2, int(math.ceil(side / edge))))
            size = {"side": side}
        else:
            r = float(rng.uniform(0.3, 1.0)) * scale
            h = float(rng.uniform(1.0, 4.0)) * scale
            mesh = geometry.cylinder(r, h, max(8, int(math.ceil(2 * math.pi * r / edge))),
                                     max(2, int(math.ceil(h / edge))), max(2, int(math.ceil(r / edge))))
            size = {"radius": r, "height": h}
        eps = float(rng.uniform(1.5, 6.0))
        out.append({"index": i, "kind": kind, **size, "eps_r": eps, "mesh": mesh})
    return out


def c4_incidences():
    return [(0.0, i * 180.0 / 7.0, 0.0) for i in range(8)]


# ---------------------------------------------------------------- C5
C5_AXES = (3.0, 2.0, 1.0)
C5_TRAIN_SEEDS = range(0, 200)
C5_HELD_OUT_SEEDS = range(1000, 1016)


def c5_eps(seed: int) -> float:
    return float(np.random.default_rng(seed).uniform(2.0, 8.0))


def c5_mesh(frequency: int = 71, scale: float = 1.0) -> geometry.Mesh:
    return geometry.ellipsoid(tuple(a * scale for a in C5_AXES), frequency=frequency)
