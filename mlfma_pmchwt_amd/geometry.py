"""Synthetic triangle meshes and the NAS (Nastran) text format.

The reference reads meshes only from NAS files (Code/NASReader.cpp:60-144):
``GRID* id x y`` followed by a continuation ``* id z`` line, and
``CTRIA3 fid pid v1 v2 v3``; whitespace-separated, not fixed-width.  This module
writes exactly that token format (17 significant digits, so a read-back is
bit-exact) and generates the closed surfaces the workloads are defined on.
Point and face ids are 1-based, as in NAS files.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "icosphere", "geodesic_sphere", "ellipsoid", "cube", "cylinder",
    "write_nas", "read_nas", "Mesh",
]


class Mesh:
    """Points (with ids) and triangles (vertex ids, outward winding)."""

    def __init__(self, xyz: np.ndarray, faces: np.ndarray,
                 point_ids: np.ndarray | None = None, face_ids: np.ndarray | None = None):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = self.xyz.shape[0]
        self.point_ids = (np.arange(1, n + 1, dtype=np.int32) if point_ids is None
                          else np.ascontiguousarray(point_ids, dtype=np.int32))
        f = np.asarray(faces)
        if point_ids is None:
            f = f + 1  # 0-based indices -> 1-based ids
        self.faces = np.ascontiguousarray(f, dtype=np.int32)      # vertex ids
        m = self.faces.shape[0]
        self.face_ids = (np.arange(1, m + 1, dtype=np.int32) if face_ids is None
                         else np.ascontiguousarray(face_ids, dtype=np.int32))

    @property
    def n_points(self) -> int:
        return self.xyz.shape[0]

    @property
    def n_faces(self) -> int:
        return self.faces.shape[0]

    @property
    def n_edges_interior(self) -> int:
        e = np.sort(np.concatenate([self.faces[:, [0, 1]], self.faces[:, [1, 2]],
                                    self.faces[:, [2, 0]]]), axis=1)
        _, counts = np.unique(e, axis=0, return_counts=True)
        return int(np.sum(counts == 2))


_ICO_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array([
    [-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0],
    [0, -1, _ICO_T], [0, 1, _ICO_T], [0, -1, -_ICO_T], [0, 1, -_ICO_T],
    [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1]], dtype=np.float64)
_ICO_F = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def icosphere(subdivisions: int, radius: float = 1.0,
              centre=(0.0, 0.0, 0.0)) -> Mesh:
    """Recursive midpoint subdivision of the icosahedron, projected on the sphere.

    ``subdivisions = s`` gives 20*4^s triangles and 30*4^s interior edges (RWG).
    """
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    for _ in range(subdivisions):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.vstack([v, mid])
        nf = f.shape[0]
        a = base + inv[:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return Mesh(v * radius + np.asarray(centre, dtype=np.float64), f)


def geodesic_sphere(frequency: int, radius: float = 1.0,
                    centre=(0.0, 0.0, 0.0)) -> Mesh:
    """Class-I geodesic sphere: every icosahedron face split into frequency^2
    triangles on a barycentric lattice, then projected on the sphere.

    Gives 20 n^2 triangles and 30 n^2 interior edges for any integer n >= 1
    (n = 100 -> 300,000 RWG functions).
    """
    n = int(frequency)
    v0 = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    # global vertex numbering: icosahedron corners, then edge interiors, then face interiors
    edges = {}
    verts = [v0[i] for i in range(12)]

    def edge_points(a, b):
        key = (min(a, b), max(a, b))
        if key not in edges:
            ids = []
            for t in range(1, n):
                p = v0[key[0]] * (n - t) / n + v0[key[1]] * t / n
                ids.append(len(verts))
                verts.append(p)
            edges[key] = ids
        ids = edges[key]
        return ids if a < b else ids[::-1]

    faces = []
    for (a, b, c) in _ICO_F:
        eab, ebc, eca = edge_points(a, b), edge_points(b, c), edge_points(c, a)
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                k = n - i - j
                if i == n:
                    idx = a
                elif j == n:
                    idx = b
                elif k == n:
                    idx = c
                elif k == 0:        # on edge a-b, j steps from a to b
                    idx = eab[j - 1]
                elif i == 0:        # on edge b-c, k steps from b to c
                    idx = ebc[k - 1]
                elif j == 0:        # on edge c-a, i steps from c to a
                    idx = eca[i - 1]
                else:
                    idx = len(verts)
                    verts.append((v0[a] * i + v0[b] * j + v0[c] * k) / n)
                grid[(i, j)] = idx
        for i in range(n):
            for j in range(n - i):
                faces.append((grid[(i + 1, j)], grid[(i, j + 1)], grid[(i, j)]))
                if i + j < n - 1:
                    faces.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    v = np.array(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(faces, dtype=np.int64)
    # orient outward (the lattice order above can flip with the corner order)
    p = v[f]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    flip = np.einsum("ij,ij->i", nrm, p.mean(axis=1)) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return Mesh(v * radius + np.asarray(centre, dtype=np.float64), f)


def ellipsoid(semi_axes, subdivisions: int = 3, frequency: int | None = None,
              centre=(0.0, 0.0, 0.0)) -> Mesh:
    """Unit sphere mesh scaled by the semi-axes (a, b, c)."""
    s = geodesic_sphere(frequency) if frequency is not None else icosphere(subdivisions)
    xyz = (s.xyz - 0.0) * np.asarray(semi_axes, dtype=np.float64)
    return Mesh(xyz + np.asarray(centre, dtype=np.float64), s.faces - 1)


def cube(side: float, n: int, centre=(0.0, 0.0, 0.0)) -> Mesh:
    """Closed cube surface, each face an n x n grid split into 2 n^2 triangles."""
    h = side / 2.0
    pts = {}
    verts = []

    def vid(p):
        key = tuple(np.round(np.asarray(p) / side * (2 * n)).astype(int))
        if key not in pts:
            pts[key] = len(verts)
            verts.append(p)
        return pts[key]

    faces = []
    lin = np.linspace(-h, h, n + 1)
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            u, w = [a for a in range(3) if a != axis]
            for i in range(n):
                for j in range(n):
                    q = []
                    for (di, dj) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = np.zeros(3)
                        p[axis] = sgn * h
                        p[u] = lin[i + di]
                        p[w] = lin[j + dj]
                        q.append(vid(p))
                    faces.append((q[0], q[1], q[2]))
                    faces.append((q[0], q[2], q[3]))
    v = np.array(verts, dtype=np.float64)
    f = np.array(faces, dtype=np.int64)
    p = v[f]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    flip = np.einsum("ij,ij->i", nrm, p.mean(axis=1)) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return Mesh(v + np.asarray(centre, dtype=np.float64), f)


def cylinder(radius: float, height: float, n_around: int, n_height: int, n_rings: int,
             centre=(0.0, 0.0, 0.0)) -> Mesh:
    """Closed cylinder (axis z): side grid plus two capped disks."""
    verts = []
    faces = []
    zs = np.linspace(-height / 2, height / 2, n_height + 1)
    ang = np.linspace(0.0, 2 * np.pi, n_around, endpoint=False)
    side = np.zeros((n_height + 1, n_around), dtype=np.int64)
    for i, z in enumerate(zs):
        for j, a in enumerate(ang):
            side[i, j] = len(verts)
            verts.append((radius * np.cos(a), radius * np.sin(a), z))
    for i in range(n_height):
        for j in range(n_around):
            jn = (j + 1) % n_around
            faces.append((side[i, j], side[i, jn], side[i + 1, jn]))
            faces.append((side[i, j], side[i + 1, jn], side[i + 1, j]))
    for cap, z, ring0 in ((0, zs[0], side[0]), (1, zs[-1], side[-1])):
        prev = ring0
        for r in range(n_rings - 1, 0, -1):
            rr = radius * r / n_rings
            cur = np.array([len(verts) + j for j in range(n_around)])
            for j, a in enumerate(ang):
                verts.append((rr * np.cos(a), rr * np.sin(a), z))
            for j in range(n_around):
                jn = (j + 1) % n_around
                faces.append((prev[j], prev[jn], cur[jn]))
                faces.append((prev[j], cur[jn], cur[j]))
            prev = cur
        c = len(verts)
        verts.append((0.0, 0.0, z))
        for j in range(n_around):
            faces.append((prev[j], prev[(j + 1) % n_around], c))
    v = np.array(verts, dtype=np.float64)
    f = np.array(faces, dtype=np.int64)
    p = v[f]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    flip = np.einsum("ij,ij->i", nrm, p.mean(axis=1) - np.array([0, 0, 0])) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return Mesh(v + np.asarray(centre, dtype=np.float64), f)


def write_nas(path, mesh: Mesh) -> None:
    """Write the reference's NAS token format (Code/NASReader.cpp:72-120)."""
    with open(path, "w") as fh:
        fh.write("$ written by mlfma_pmchwt_amd.geometry\nBEGIN BULK\n")
        for pid, (x, y, z) in zip(mesh.point_ids, mesh.xyz):
            fh.write(f"GRID*   {int(pid):d} {float(x):.17g} {float(y):.17g}\n*       {int(pid):d} {float(z):.17g}\n")
        for fid, (a, b, c) in zip(mesh.face_ids, mesh.faces):
            fh.write(f"CTRIA3  {int(fid):d} 1 {int(a):d} {int(b):d} {int(c):d}\n")
        fh.write("ENDDATA\n")


def read_nas(path) -> Mesh:
    """Parse the same token format (first token after the keyword is the id)."""
    ids, xyz, fids, fv = [], [], [], []
    pending = None
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\r\n")
            if line.startswith("GRID*"):
                t = line[5:].split()
                pending = (int(t[0]), float(t[1]), float(t[2]))
            elif pending is not None and line.startswith("*"):
                t = line[1:].split()
                if int(t[0]) != pending[0]:
                    raise ValueError("ID mismatch between GRID* and * lines")
                ids.append(pending[0])
                xyz.append((pending[1], pending[2], float(t[1])))
                pending = None
            elif line.startswith("CTRIA3"):
                t = line[6:].split()
                fids.append(int(t[0]))
                fv.append((int(t[2]), int(t[3]), int(t[4])))
    if pending is not None:
        raise ValueError("Unfinished GRID* entry at end of file")
    return Mesh(np.array(xyz), np.array(fv, dtype=np.int32), np.array(ids, dtype=np.int32),
                np.array(fids, dtype=np.int32))
