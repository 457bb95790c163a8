"""Synthetic tetrahedral meshes: seeded, jittered Kuhn boxes.

A box of ``nx x ny x nz`` cubes (spacing ``h``) is split into 6 tetrahedra
per cube along the main diagonal (Kuhn / Freudenthal split).  Every cube is
split the same way, so the mesh is conforming.  Interior nodes are moved by
a seeded uniform offset in ``[-jitter*h, +jitter*h]`` per coordinate
(``numpy.random.default_rng(seed)``); boundary nodes stay on the box.

Layouts follow the reference's ``Mesh`` class (src/Mesh.m:8-20): ``P`` is
``[3 x Nn]`` column-major (returned as a C-order ``(Nn, 3)`` array, the same
bytes), ``T`` is ``[4 x Ne]`` column-major of 1-based node ids (returned as
``(Ne, 4)`` int32), ``RegionTags`` is ``[1 x Ne]``.

Element order: cube-major (x fastest, then y, then z), 6 consecutive
tetrahedra per cube, one per permutation of the axis steps.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Axis-step permutations; tet t of a cube is (v000, v000+e_p0, v000+e_p0+e_p1, v111).
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


@dataclass
class KuhnMesh:
    P: np.ndarray          # (Nn, 3) float64 == MATLAB P [3 x Nn]
    T: np.ndarray          # (Ne, 4) int32, 1-based == MATLAB T [4 x Ne]
    RegionTags: np.ndarray # (Ne,) int32
    shape: tuple           # (nx, ny, nz)
    h: tuple               # (hx, hy, hz)

    @property
    def num_nodes(self) -> int:
        return int(self.P.shape[0])

    @property
    def num_elems(self) -> int:
        return int(self.T.shape[0])

    def centroids(self) -> np.ndarray:
        t0 = self.T - 1
        return self.P[t0].mean(axis=1)


def kuhn_box(nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0), jitter: float = 0.1,
             seed: int = 1, origin=(0.0, 0.0, 0.0)) -> KuhnMesh:
    nx, ny, nz = int(nx), int(ny), int(nz)
    lx, ly, lz = (float(v) for v in lengths)
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    nnx, nny, nnz_ = nx + 1, ny + 1, nz + 1
    nn = nnx * nny * nnz_

    ii = np.arange(nnx, dtype=np.float64)
    jj = np.arange(nny, dtype=np.float64)
    kk = np.arange(nnz_, dtype=np.float64)
    K3, J3, I3 = np.meshgrid(kk, jj, ii, indexing="ij")  # node id = i + nnx*(j + nny*k)
    P = np.empty((nn, 3), dtype=np.float64)
    P[:, 0] = origin[0] + I3.ravel() * hx
    P[:, 1] = origin[1] + J3.ravel() * hy
    P[:, 2] = origin[2] + K3.ravel() * hz

    if jitter:
        ib = I3.ravel().astype(np.int64)
        jb = J3.ravel().astype(np.int64)
        kb = K3.ravel().astype(np.int64)
        interior = (ib > 0) & (ib < nx) & (jb > 0) & (jb < ny) & (kb > 0) & (kb < nz)
        idx = np.nonzero(interior)[0]
        rng = np.random.default_rng(seed)
        off = rng.uniform(-1.0, 1.0, size=(idx.size, 3))
        off *= np.array([jitter * hx, jitter * hy, jitter * hz])
        P[idx] += off

    # cubes, x fastest
    cz, cy, cx = np.meshgrid(np.arange(nz, dtype=np.int64), np.arange(ny, dtype=np.int64),
                             np.arange(nx, dtype=np.int64), indexing="ij")
    cx = cx.ravel(); cy = cy.ravel(); cz = cz.ravel()
    base = cx + nnx * (cy + nny * cz)
    step = np.array([1, nnx, nnx * nny], dtype=np.int64)
    v111 = base + step.sum()
    ncube = base.size
    T = np.empty((ncube, 6, 4), dtype=np.int64)
    for t, (p0, p1, _) in enumerate(_PERMS):
        T[:, t, 0] = base
        T[:, t, 1] = base + step[p0]
        T[:, t, 2] = base + step[p0] + step[p1]
        T[:, t, 3] = v111
    T = (T.reshape(ncube * 6, 4) + 1).astype(np.int32)
    tags = np.ones(T.shape[0], dtype=np.int32)
    return KuhnMesh(P=P, T=T, RegionTags=tags, shape=(nx, ny, nz), h=(hx, hy, hz))


def tag_box_region(mesh: KuhnMesh, lo, hi, tag: int) -> None:
    """Set ``RegionTags = tag`` for tetrahedra whose centroid lies in [lo, hi]^3."""
    c = mesh.centroids()
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    m = np.all((c >= lo) & (c <= hi), axis=1)
    mesh.RegionTags[m] = tag


def kuhn_counts(nx: int, ny: int, nz: int):
    """Exact (nodes, tets, edges) of a consistent Kuhn split.

    Edges: axis edges + one face diagonal per face + one body diagonal per cube.
    """
    nodes = (nx + 1) * (ny + 1) * (nz + 1)
    tets = 6 * nx * ny * nz
    axis = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    faces = nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz
    edges = axis + faces + nx * ny * nz
    return nodes, tets, edges


def centroid_axis(mesh: KuhnMesh, axis: int) -> np.ndarray:
    """Centroid coordinate along one axis for every tetrahedron (memory-light)."""
    c = np.zeros(mesh.num_elems)
    for k in range(4):
        c += mesh.P[mesh.T[:, k] - 1, axis]
    return c * 0.25


def tag_slab(mesh: KuhnMesh, axis: int, lo: float, hi: float, tag: int) -> None:
    """Set ``RegionTags = tag`` where the centroid's ``axis`` coordinate is in [lo, hi]."""
    c = centroid_axis(mesh, axis)
    mesh.RegionTags[(c >= lo) & (c <= hi)] = tag


def tag_box(mesh: KuhnMesh, lo, hi, tag: int) -> None:
    """Memory-light variant of tag_box_region for large meshes."""
    m = np.ones(mesh.num_elems, dtype=bool)
    for ax in range(3):
        c = centroid_axis(mesh, ax)
        m &= (c >= lo[ax]) & (c <= hi[ax])
    mesh.RegionTags[m] = tag


@dataclass
class ECore:
    """Synthetic E-core problem: region tags 1 = air, 2 = iron (two legs and
    a yoke), 3 = winding around the first leg; ``js_dir`` (Ne, 3) is the unit
    current direction of the winding elements (zero elsewhere)."""
    mesh: KuhnMesh
    js_dir: np.ndarray
    lengths: tuple


def ecore_box(nx: int = 64, ny: int = 64, nz: int = 40, lengths=(0.16, 0.16, 0.10),
              seed: int = 5) -> ECore:
    """Two vertical iron legs joined by a yoke inside an air box, and a
    rectangular winding around the first leg carrying current that circulates
    counter-clockwise about +z.  Regions are fractions of the box, so any
    resolution gives the same geometry; 64 x 64 x 40 cubes is 983,040 tets."""
    m = kuhn_box(nx, ny, nz, lengths=lengths, seed=seed)
    L = np.asarray(lengths, dtype=np.float64)
    c = np.stack([centroid_axis(m, ax) for ax in range(3)], axis=1) / L

    def box(lo, hi):
        return np.all((c >= np.asarray(lo)) & (c <= np.asarray(hi)), axis=1)

    leg1 = box((0.25, 0.375, 0.2), (0.375, 0.625, 0.85))
    leg2 = box((0.625, 0.375, 0.2), (0.75, 0.625, 0.85))
    yoke = box((0.25, 0.375, 0.7), (0.75, 0.625, 0.85))
    iron = leg1 | leg2 | yoke
    coil = box((0.15, 0.275, 0.3), (0.475, 0.725, 0.6)) & ~iron
    m.RegionTags[:] = 1
    m.RegionTags[iron] = 2
    m.RegionTags[coil] = 3
    # tangential direction around the first leg's axis (rectangular loop)
    dx = (c[:, 0] - 0.3125) / 0.0625
    dy = (c[:, 1] - 0.5) / 0.125
    xside = np.abs(dx) >= np.abs(dy)
    js = np.zeros((m.num_elems, 3))
    js[:, 0] = np.where(xside, 0.0, -np.sign(dy))
    js[:, 1] = np.where(xside, np.sign(dx), 0.0)
    js[~coil] = 0.0
    return ECore(m, js, tuple(float(v) for v in lengths))
