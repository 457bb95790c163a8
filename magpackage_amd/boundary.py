"""BoundaryCondition.findOuterBoundaryDofs / applyDirichlet for Nedelec spaces.

Mesh set-up work on the host (numpy), done once per mesh like the
reference's MATLAB code: faces seen once are on the outer boundary
(src/BoundaryCondition.m:9-17), their three edges are boundary edges
(:36-45), and the DOFs of those edges are fixed (:47, :62-65).  Applying the
condition to a system is done on the device by the linear solver
(``mpk_pcg_solve``'s fixed mask) and by masking residuals.
"""
from __future__ import annotations

import numpy as np

_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def _pair_keys(a: np.ndarray, b: np.ndarray, base: int) -> np.ndarray:
    lo = np.minimum(a, b).astype(np.int64)
    hi = np.maximum(a, b).astype(np.int64)
    return lo * base + hi


def outer_boundary_edge_mask(T: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """Boolean mask over ``edges`` ((E, 2) sorted (min, max) 1-based node pairs,
    as Mesh.generateEdges returns them) of the edges on the outer boundary of
    the tetrahedra ``T`` ((Ne, 4) 1-based)."""
    T = np.asarray(T, dtype=np.int64)
    edges = np.asarray(edges, dtype=np.int64)
    if T.size == 0 or edges.size == 0:
        return np.zeros(len(edges), dtype=bool)
    base = int(max(T.max(), edges.max())) + 1
    if base ** 3 >= 2 ** 63:
        raise ValueError("too many nodes for 64-bit face keys")
    F = np.concatenate([np.sort(T[:, list(f)], axis=1) for f in _FACES], axis=0)
    fkey = (F[:, 0] * base + F[:, 1]) * base + F[:, 2]
    uniq, counts = np.unique(fkey, return_counts=True)
    bnd = uniq[counts == 1]
    a = bnd // (base * base)
    b = (bnd // base) % base
    c = bnd % base
    ekeys = np.unique(np.concatenate([_pair_keys(a, b, base), _pair_keys(a, c, base),
                                      _pair_keys(b, c, base)]))
    gkeys = edges[:, 0] * base + edges[:, 1]
    pos = np.searchsorted(ekeys, gkeys)
    pos = np.minimum(pos, len(ekeys) - 1)
    return ekeys[pos] == gkeys


def find_outer_boundary_dofs(mesh, dof_handler, space) -> np.ndarray:
    """BoundaryCondition.findOuterBoundaryDofs (src/BoundaryCondition.m:4-70),
    Nedelec branch: boolean [NumGlobalDofs] mask of the fixed DOFs."""
    key = space.toString()
    if key not in dof_handler.EntityMaps:
        raise KeyError(f"DofHandler EntityMap missing for {key}. Did you call distributeDofs?")
    if space.Type.lower() != "nedelec":
        raise ValueError("Unsupported space type.")
    edges = mesh.Edges.cpu().numpy()
    on_bnd = outer_boundary_edge_mask(mesh.T, edges)
    emap = dof_handler.EntityMaps[key].cpu().numpy()
    dofs = emap[np.nonzero(on_bnd)[0]]
    dofs = dofs[dofs > 0]
    mask = np.zeros(dof_handler.NumGlobalDofs, dtype=bool)
    mask[dofs - 1] = True
    return mask
