"""Shared builders for the tests: meshes, reference tables, independent math."""
from __future__ import annotations

import numpy as np

from magpackage_amd.fem_tables import (EDGE_NODES, get_quadrature_data, matlab_flat,
                                       nedelec_tet_p1)
from magpackage_amd.material import MU0
from magpackage_amd.meshgen import kuhn_box, tag_box_region


def stiffness_tables():
    q_pts, q_w = get_quadrature_data("tet", 2)
    _, curl = nedelec_tet_p1(q_pts)
    return q_w, matlab_flat(curl), q_pts


def mass_tables(order=2):
    q_pts, q_w = get_quadrature_data("tet", order)
    val, _ = nedelec_tet_p1(q_pts)
    return q_w, matlab_flat(val), q_pts


def core_box_mesh(n=20, seed=1):
    """Unit Kuhn box of n^3 cubes, iron core [0.3,0.7]^3 (tag 2) in air (tag 1)."""
    m = kuhn_box(n, n, n, seed=seed)
    tag_box_region(m, (0.3, 0.3, 0.3), (0.7, 0.7, 0.7), 2)
    return m


def core_box_nu(mesh):
    nu = np.where(mesh.RegionTags == 2, 1.0 / (1000.0 * MU0), 1.0 / MU0)
    return nu.astype(np.float64)


def random_tets(n, seed=0, scale=1.0):
    """n independent well-shaped tetrahedra, P (4n,3), T (n,4) 1-based, both orientations."""
    rng = np.random.default_rng(seed)
    pts = np.empty((n, 4, 3))
    i = 0
    while i < n:
        J = rng.normal(size=(3, 3)) * scale
        if abs(np.linalg.det(J)) < 0.2 * scale ** 3:
            continue
        p0 = rng.uniform(-1, 1, 3)
        pts[i, 0] = p0
        pts[i, 1:] = p0 + J.T
        i += 1
    P = pts.reshape(-1, 3)
    T = np.arange(1, 4 * n + 1, dtype=np.int32).reshape(n, 4)
    return P, T


def grads(p):
    """Physical gradients of the 4 barycentric functions of tet p (4,3) via numpy inverse."""
    J = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1)  # columns
    Jit = np.linalg.inv(J).T
    g_ref = np.array([[-1.0, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    return g_ref @ Jit.T, np.linalg.det(J)


def whitney_curls(p):
    """curl W_ab = 2 grad l_a x grad l_b (independent of the J c / det formula)."""
    g, det = grads(p)
    return np.array([2.0 * np.cross(g[a], g[b]) for a, b in EDGE_NODES]), det


def gradient_matrix(edges, nn):
    """Discrete gradient G (E x Nn): edge (n1<n2) -> phi_n2 - phi_n1."""
    import scipy.sparse as sp
    E = edges.shape[0]
    rows = np.repeat(np.arange(E), 2)
    cols = (edges - 1).ravel()
    vals = np.tile([-1.0, 1.0], E)
    return sp.csr_matrix((vals, (rows, cols)), shape=(E, nn))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def csc_equal_bits(A, B):
    """Exact equality of two scipy CSC matrices (pattern and values, bitwise)."""
    A = A.tocsc(); B = B.tocsc()
    if A.shape != B.shape or A.nnz != B.nnz:
        return False
    A.sort_indices(); B.sort_indices()
    return (np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            and np.array_equal(A.data.view(np.uint64), B.data.view(np.uint64)))


def element_B(P, T, Dofs, Signs, sol):
    """Per-element B = sum_i a_i s_i curl W_i (vectorised Whitney curls,
    2 grad l_a x grad l_b); P (Nn, 3), T (Ne, 4), Dofs / Signs (Ne, 6), ids
    1-based (the memory of the MATLAB [3 x Nn], [4 x Ne], [6 x Ne] arrays)."""
    X = P[(T - 1).astype(np.int64)]  # (Ne, 4, 3)
    Jm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    Jit = np.transpose(np.linalg.inv(Jm), (0, 2, 1))
    g_ref = np.array([[-1.0, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    g = np.einsum("ij,ekj->eik", g_ref, Jit)  # (Ne, 4, 3) physical gradients
    d = Dofs.astype(np.int64)
    a = np.where(d > 0, sol[np.maximum(d, 1) - 1], 0.0) * Signs
    B = np.zeros((T.shape[0], 3))
    for i, (p, q) in enumerate(EDGE_NODES):
        B += a[:, i:i + 1] * 2.0 * np.cross(g[:, p], g[:, q])
    return B


def csc_from_plan(plan, vals):
    """scipy CSC matrix of structural-pattern values from the plan."""
    import scipy.sparse as sp
    colptr, rowind = plan.pattern()
    v = vals.detach().cpu().numpy() if hasattr(vals, "detach") else np.asarray(vals)
    return sp.csc_matrix((v, rowind.cpu().numpy(), colptr.cpu().numpy()),
                         shape=(plan.n, plan.n))
