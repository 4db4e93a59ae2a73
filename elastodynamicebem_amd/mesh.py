"""MEDIT ``.mesh`` reader and the per-triangle data the single-layer path uses.

Follows util/fileRead/readSpaceMesh.m:13-151 of the reference: four header
lines, a ``Vertices`` section (x y z ref), a ``Triangles`` section (i j k ref,
1-based); the triangle reference column is forced to 4 (readSpaceMesh.m:80-81)
and is not kept here.  Normals, areas, centres and the longest edge are derived
by the same expressions as readSpaceMesh.m:93-139 (also computed, identically,
inside the native library from the raw coordinates).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SpaceMesh:
    coordinates: np.ndarray  # (Nv, 3) float64
    triangles: np.ndarray    # (N, 3) int32, 0-based
    normal: np.ndarray       # (N, 3)
    area: np.ndarray         # (N,)
    center: np.ndarray       # (N, 3)
    maxL: np.ndarray         # (N,)

    @property
    def number_nodes(self) -> int:
        return int(self.coordinates.shape[0])

    @property
    def numberTriangles(self) -> int:
        return int(self.triangles.shape[0])

    @property
    def lMax(self) -> float:
        """Longest edge of the mesh (calcParamMesh.m:1-30)."""
        return float(self.maxL.max())


def derive(coords: np.ndarray, tri: np.ndarray):
    """Normals, areas, centres and longest edges as in readSpaceMesh.m:93-139."""
    P1 = coords[tri[:, 0]]
    P2 = coords[tri[:, 1]]
    P3 = coords[tri[:, 2]]
    vg1 = P1 - P3
    vg2 = P2 - P3
    vg3 = P1 - P2
    n = np.empty_like(vg1)
    n[:, 0] = vg1[:, 1] * vg2[:, 2] - vg2[:, 1] * vg1[:, 2]
    n[:, 1] = -vg1[:, 0] * vg2[:, 2] + vg2[:, 0] * vg1[:, 2]
    n[:, 2] = vg1[:, 0] * vg2[:, 1] - vg2[:, 0] * vg1[:, 1]
    area = np.sqrt(n[:, 0] ** 2 + n[:, 1] ** 2 + n[:, 2] ** 2) / 2
    normal = n / (2 * area)[:, None]
    center = (P1 + P2 + P3) / 3
    lens = np.stack([np.sqrt((vg3 ** 2).sum(1)), np.sqrt((vg2 ** 2).sum(1)),
                     np.sqrt((vg1 ** 2).sum(1))], axis=1)
    return normal, area, center, lens.max(1)


def from_arrays(coords, tri) -> SpaceMesh:
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    if coords.ndim != 2 or coords.shape[1] != 3:
        raise ValueError("coordinates must be (Nv, 3)")
    if tri.ndim != 2 or tri.shape[1] != 3:
        raise ValueError("triangles must be (N, 3)")
    if tri.size and (tri.min() < 0 or tri.max() >= coords.shape[0]):
        raise ValueError("triangle vertex index out of range")
    normal, area, center, maxL = derive(coords, tri)
    return SpaceMesh(coords, tri, normal, area, center, maxL)


def read_mesh(path: str) -> SpaceMesh:
    """Read a MEDIT mesh (readSpaceMesh.m:13-88)."""
    with open(path, "r") as fh:
        toks = fh.read().split()
    try:
        iv = toks.index("Vertices")
        nv = int(toks[iv + 1])
        vals = np.array(toks[iv + 2: iv + 2 + 4 * nv], dtype=np.float64).reshape(nv, 4)
        it = toks.index("Triangles", iv + 2 + 4 * nv)
        nt = int(toks[it + 1])
        tv = np.array(toks[it + 2: it + 2 + 4 * nt], dtype=np.float64).reshape(nt, 4)
    except (ValueError, IndexError) as exc:
        raise ValueError(f"malformed MEDIT mesh file {path!r}") from exc
    return from_arrays(vals[:, :3], tv[:, :3].astype(np.int64) - 1)


def write_mesh(path: str, mesh: SpaceMesh) -> None:
    """Write a MEDIT mesh that read_mesh (and readSpaceMesh.m) reads back."""
    with open(path, "w") as fh:
        fh.write(" MeshVersionFormatted\n 2\n Dimension\n 3\n Vertices\n")
        fh.write(f" {mesh.number_nodes}\n")
        for x, y, z in mesh.coordinates:
            fh.write(f" {x!r} {y!r} {z!r} 0\n")
        fh.write(" Triangles\n")
        fh.write(f" {mesh.numberTriangles}\n")
        for a, b, c in mesh.triangles + 1:
            fh.write(f" {a} {b} {c} 0\n")
        fh.write(" End\n")


def screen_uniform(n: int, half: float = 0.5) -> SpaceMesh:
    """Flat square screen [-half, half]^2 at z=0 split into n x n squares, each
    cut into two right triangles with normals +z."""
    xs = np.linspace(-half, half, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    idx = lambda i, j: i * (n + 1) + j
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    return from_arrays(coords, np.array(tris))


def icosphere(level: int) -> SpaceMesh:
    """Unit icosphere by repeated 4-way subdivision (20 * 4**level triangles),
    outward normals."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    coords = np.array(verts)
    tri = np.array(faces)
    m = from_arrays(coords, tri)
    # orient outward
    flip = (m.normal * m.center).sum(1) < 0
    if flip.any():
        tri = tri.copy()
        tri[flip] = tri[flip][:, [0, 2, 1]]
        m = from_arrays(coords, tri)
    return m
