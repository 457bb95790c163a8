"""Reference-element tables: quadrature rule and P1 shape functions.

Host-side mirror of the reference's table builders, same layouts and the
same floating-point operation order, so the tables fed to the GPU kernels
are bit-identical to what ``Assembler.m`` passes to its MEX kernels.

- ``get_quadrature_data``  src/utils/get_quadrature_data.m:12-29
  (1-point rule, or the 4-point Keast rule with the *truncated* constants
  a = 0.58541020, b = 0.13819660, w = 1/24).
- ``lagrange_tet_p1``      src/shape_functions/lagrange_tet_p1.m:1-38
- ``nedelec_tet_p1``       src/shape_functions/nedelec_tet_p1.m:18-51
  (Whitney W_ab = l_a grad l_b - l_b grad l_a; reference curl 2 grad l_a x grad l_b).

Arrays are returned with MATLAB's memory order (column-major), i.e. the
``[6 x 3 x Nq]`` tables are flat float64 vectors indexed ``i + 6*c + 18*q``
exactly as ``assemble_*_kernel_mex.cpp`` reads them.
"""
from __future__ import annotations

import numpy as np

# Local edge definitions [1 2; 1 3; 1 4; 2 3; 2 4; 3 4] (src/Mesh.m:118,
# src/shape_functions/nedelec_tet_p1.m:13), 0-based here.
EDGE_NODES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))

_A_TRUNC = 0.58541020
_B_TRUNC = 0.13819660


def get_quadrature_data(elem_type: str = "tet", order: int = 2):
    """Return ``(q_pts [3 x Nq], q_w [Nq])`` (get_quadrature_data.m:12-29)."""
    if elem_type.lower() != "tet":
        raise ValueError(f"Unsupported element type: {elem_type}")
    if order <= 1:
        q_pts = np.array([[0.25], [0.25], [0.25]], dtype=np.float64)
        q_w = np.array([1.0 / 6.0], dtype=np.float64)
    else:
        a, b = _A_TRUNC, _B_TRUNC
        q_pts = np.array([[a, b, b, b],
                          [b, a, b, b],
                          [b, b, a, b]], dtype=np.float64)
        q_w = np.full(4, 1.0 / 24.0, dtype=np.float64)
    return q_pts, q_w


def lagrange_tet_p1(q_pts: np.ndarray):
    """Return ``(val [4 x Nq], grad [4 x 3 x Nq])`` (lagrange_tet_p1.m:9-38).

    ``val(1,:) = 1 - xi - eta - zeta`` is evaluated left to right as in MATLAB.
    """
    xi, eta, zeta = q_pts[0], q_pts[1], q_pts[2]
    n = q_pts.shape[1]
    val = np.zeros((4, n), dtype=np.float64)
    val[0] = ((1.0 - xi) - eta) - zeta
    val[1] = xi
    val[2] = eta
    val[3] = zeta
    grad = np.zeros((4, 3, n), dtype=np.float64)
    grad[0, :, :] = -1.0
    grad[1, 0, :] = 1.0
    grad[2, 1, :] = 1.0
    grad[3, 2, :] = 1.0
    return val, grad


def nedelec_tet_p1(q_pts: np.ndarray):
    """Return ``(val, curl)`` as ``[6 x 3 x Nq]`` arrays (nedelec_tet_p1.m:18-51).

    ``val(e,c,p) = L_i(p) * G_j(c) - L_j(p) * G_i(c)``;
    ``curl(e,:,p) = 2 * cross(G_i, G_j)``.
    """
    lv, lg = lagrange_tet_p1(q_pts)
    n = q_pts.shape[1]
    val = np.zeros((6, 3, n), dtype=np.float64)
    curl = np.zeros((6, 3, n), dtype=np.float64)
    for e, (i, j) in enumerate(EDGE_NODES):
        gi = lg[i, :, 0]
        gj = lg[j, :, 0]
        term1 = gj[None, :] * lv[i][:, None]   # [n x 3]
        term2 = gi[None, :] * lv[j][:, None]
        val[e] = (term1 - term2).T
        c = 2.0 * np.cross(gi, gj)
        curl[e] = np.repeat(c[:, None], n, axis=1)
    return val, curl


def matlab_flat(a: np.ndarray) -> np.ndarray:
    """Column-major flattening (MATLAB memory order) as a contiguous float64 vector."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def stiffness_tables(order: int = 2):
    """Tables ``Assembler.assembleStiffness`` passes (Assembler.m:30-33): order = max(order, 2)."""
    q_pts, q_w = get_quadrature_data("tet", max(order, 2))
    _, curl = nedelec_tet_p1(q_pts)
    return q_w.copy(), matlab_flat(curl)


def mass_tables(order: int = 2, weighted: bool = True):
    """Tables for the mass kernels.

    ``assembleMassWeighted`` uses the un-maxed ``DefaultQuadratureOrder``
    (Assembler.m:64); ``assembleMass`` uses ``max(order, 2)`` (Assembler.m:45).
    """
    o = order if weighted else max(order, 2)
    q_pts, q_w = get_quadrature_data("tet", o)
    val, _ = nedelec_tet_p1(q_pts)
    return q_w.copy(), matlab_flat(val), q_pts


def barycentric_moments(q_pts: np.ndarray, q_w: np.ndarray) -> np.ndarray:
    """``Q[x, y] = sum_q w_q l_x(q) l_y(q)`` for the rule's own points.

    The closed-form mass path uses these moments of the (truncated) rule, so
    that it integrates exactly what the 4-point loop of the reference does.
    """
    lv, _ = lagrange_tet_p1(q_pts)
    q = np.zeros((4, 4), dtype=np.float64)
    for k in range(q_pts.shape[1]):
        q += q_w[k] * np.outer(lv[:, k], lv[:, k])
    return q
