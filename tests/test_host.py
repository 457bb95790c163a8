"""Host-side logic (no GPU): tables, materials, packing, meshes, library exports."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest
from scipy.interpolate import PchipInterpolator, PPoly

from magpackage_amd.fem_tables import (EDGE_NODES, barycentric_moments, get_quadrature_data,
                                       lagrange_tet_p1, matlab_flat, nedelec_tet_p1)
from magpackage_amd.material import (MU0, create_linear, create_nonlinear, marrocco,
                                     marrocco_exact, pack_material_data, pchip, ppval)
from magpackage_amd.meshgen import kuhn_box, kuhn_counts, tag_box, tag_box_region

ROOT = Path(__file__).resolve().parent.parent


def test_quadrature_tables():
    p1, w1 = get_quadrature_data("tet", 1)
    assert p1.shape == (3, 1) and w1[0] == 1.0 / 6.0
    p, w = get_quadrature_data("tet", 2)
    assert p.shape == (3, 4) and np.all(w == 1.0 / 24.0)
    assert p[0, 0] == 0.58541020 and p[1, 0] == 0.13819660  # truncated constants
    # the rule integrates linear functions exactly on the reference tet (volume 1/6)
    lv, _ = lagrange_tet_p1(p)
    assert np.allclose(lv @ w, 1.0 / 24.0, rtol=0, atol=1e-9)


def test_nedelec_tables_closed_form():
    p, _ = get_quadrature_data("tet", 2)
    val, curl = nedelec_tet_p1(p)
    ref_c = np.array([[0, -2, 2], [2, 0, -2], [-2, 2, 0], [0, 0, 2], [0, -2, 0], [2, 0, 0]])
    for q in range(4):
        assert np.array_equal(curl[:, :, q], ref_c)
    lv, g = lagrange_tet_p1(p)
    for e, (a, b) in enumerate(EDGE_NODES):
        W = lv[a][None, :] * g[b, :, 0][:, None] - lv[b][None, :] * g[a, :, 0][:, None]
        assert np.array_equal(val[e], W)
    flat = matlab_flat(curl)
    assert flat[0 + 6 * 1 + 18 * 2] == curl[0, 1, 2]  # i + 6c + 18q indexing


def test_barycentric_moments():
    p, w = get_quadrature_data("tet", 2)
    Q = barycentric_moments(p, w)
    exact = (np.ones((4, 4)) + np.eye(4)) / 120.0
    assert np.abs(Q - exact).max() / exact.max() < 1e-7  # truncation of the constants
    assert np.abs(Q - Q.T).max() == 0.0


def test_pchip_matches_scipy():
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.uniform(0.1, 1.0, 25))
    y = np.cumsum(rng.uniform(0.0, 1.0, 25)) ** 1.5
    pp = pchip(x, y)
    s = np.linspace(x[0], x[-1], 1000)
    ref = PchipInterpolator(x, y)(s)
    assert np.max(np.abs(ppval(pp, s) - ref)) <= 1e-12 * np.max(np.abs(ref))
    # non-monotone data (sign changes of the slopes)
    y2 = np.sin(x)
    assert np.max(np.abs(ppval(pchip(x, y2), s) - PchipInterpolator(x, y2)(s))) < 1e-12
    # interpolates the knots and is C1 (scipy PPoly derivative)
    P = PPoly(pp.coefs.T, pp.breaks)
    assert np.allclose(P(x), y, rtol=1e-14, atol=0)
    d = P.derivative()
    assert np.allclose(d(x[1:-1] - 1e-12), d(x[1:-1] + 1e-12), rtol=1e-6)


def test_create_nonlinear_table():
    B = np.array([0.0, 0.5, 1.0, 1.5, 1.8, 2.0, 1.0])  # duplicate 1.0: unique 'last'
    H = np.array([0.0, 100.0, 250.0, 1000.0, 8000.0, 30000.0, 260.0])
    m = create_nonlinear(B, H)
    assert m.Type == "Nonlinear"
    Bu = np.array([0.0, 0.5, 1.0, 1.5, 1.8, 2.0])
    Hu = np.array([0.0, 100.0, 260.0, 1000.0, 8000.0, 30000.0])
    nu = Hu / np.where(Bu > 0, Bu, 1.0)
    nu[0] = nu[1]
    assert m.MaxBSq == pytest.approx(4.0 + 0.1)
    bsq = m.SplineNu.breaks
    assert np.allclose(bsq[:-1], Bu ** 2) and bsq[-1] == pytest.approx(4.1)
    assert np.allclose(ppval(m.SplineNu, Bu ** 2), nu, rtol=1e-13)
    slope = (1.0 / MU0 - nu[-1]) / (2.0 * 4.0)
    assert ppval(m.SplineNu, 4.1) == pytest.approx(nu[-1] + slope * 0.1, rel=1e-13)


def test_marrocco_tabulation():
    m = marrocco()
    s = np.linspace(0, 4.84, 5001)
    ex, dex = marrocco_exact(s)
    got = ppval(m.SplineNu, s)
    assert np.max(np.abs(got - ex) / ex) < 5e-4
    # knots are exact
    assert np.allclose(ppval(m.SplineNu, m.SplineNu.breaks), marrocco_exact(m.SplineNu.breaks)[0],
                       rtol=1e-14)
    # the curve reaches the vacuum value near B^2 = ln((1-k3)/k1)/k2 ~ 1.52
    s_vac = np.log((1 - 2e-4) / 5e-4) / 5.0
    assert 1.5 < s_vac < 1.53


def test_pack_material_layout():
    lib = {1: create_linear(1.0), 3: marrocco(pieces=10), 2: create_linear(1000.0)}
    pk = pack_material_data(lib)
    assert pk["linNu"].size == 4  # maxTag + 1
    assert pk["linNu"][0] == pytest.approx(1 / MU0) and pk["linNu"][1] == pytest.approx(1 / (1000 * MU0))
    assert pk["isNon"][2] == 1 and pk["counts"][2] == 10 and pk["starts"][2] == 0
    assert pk["breaks"].size == 11 and pk["coefs"].size == 40
    # coefficients per piece [c3 c2 c1 c0] in row-major order (c_mat')(:)
    mm = lib[3].SplineNu
    assert np.array_equal(pk["coefs"][4:8], mm.coefs[1])
    # list form: entry i is tag i+1
    pk2 = pack_material_data([create_linear(2.0), marrocco(pieces=4)])
    assert pk2["linNu"].size == 3 and pk2["isNon"][1] == 1 and pk2["coe_starts"][1] == 0


def test_kuhn_box_properties():
    m = kuhn_box(4, 3, 2, seed=3)
    nn, ne, E = kuhn_counts(4, 3, 2)
    assert m.num_nodes == nn and m.num_elems == ne
    # deterministic for a seed, different for another
    assert np.array_equal(m.P, kuhn_box(4, 3, 2, seed=3).P)
    assert not np.array_equal(m.P, kuhn_box(4, 3, 2, seed=4).P)
    # boundary nodes stay on the box, jitter bounded by 0.1 h
    grid = kuhn_box(4, 3, 2, jitter=0.0).P
    d = np.abs(m.P - grid)
    h = np.array(m.h)
    assert np.all(d <= 0.1 * h + 1e-15)
    onb = (np.isclose(grid, 0) | np.isclose(grid, 1)).any(axis=1)
    assert np.all(d[onb] == 0)
    # volumes sum to the box, no inverted / degenerate tets
    p = m.P[m.T - 1]
    vol = np.abs(np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]),
                           p[:, 3] - p[:, 0])) / 6.0
    assert vol.min() > 0 and abs(vol.sum() - 1.0) < 1e-12
    # two region taggers agree
    a = kuhn_box(6, 6, 6, seed=1); b = kuhn_box(6, 6, 6, seed=1)
    tag_box_region(a, (0.3,) * 3, (0.7,) * 3, 2)
    tag_box(b, (0.3,) * 3, (0.7,) * 3, 2)
    assert np.array_equal(a.RegionTags, b.RegionTags) and (a.RegionTags == 2).sum() > 0


def test_library_exports_header_symbols():
    """libmpk_assembly.so loads and exports every function include/*.h declares."""
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(mpk_\w+)\s*\(", h.read_text(),
                                flags=re.M))
    assert len(names) >= 30
    lib = C.CDLL(str(ROOT / "magpackage_amd" / "libmpk_assembly.so"))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    import magpackage_amd._lib as L
    assert set(L.EXPORTED) == names
    assert L.lib.mpk_abi_version() == 1


def test_mex_arity_error_without_gpu():
    """Wrong input counts raise MagPackage:Args before any device work."""
    from magpackage_amd import mex
    from magpackage_amd._lib import MpkError
    with pytest.raises(MpkError) as e:
        mex.assemble_curl_curl_kernel_mex(1, 2, 3)
    assert "MagPackage:Args" in str(e.value) and "Expected 7 inputs." in str(e.value)
    with pytest.raises(MpkError):
        mex.assemble_jacobian_kernel_mex(*range(16))


def test_outer_boundary_edges_match_oracle():
    """Host boundary finder (product path) against the oracle restatement of
    BoundaryCondition.findOuterBoundaryDofs, on a box and on a mesh with a
    hole (a removed block of cubes leaves inner boundary faces)."""
    from magpackage_amd.boundary import outer_boundary_edge_mask
    from magpackage_amd.meshgen import kuhn_box
    from oracle import oracle as O
    m = kuhn_box(5, 4, 3, seed=41)
    edges, _, _ = O.generate_edges(m.T)
    assert np.array_equal(outer_boundary_edge_mask(m.T, edges), O.boundary_edges(m.T, edges))
    c = m.P[m.T - 1].mean(axis=1)
    keep = ~((np.abs(c[:, 0] - 0.5) < 0.2) & (np.abs(c[:, 1] - 0.5) < 0.25))
    T2 = m.T[keep]
    edges2, _, _ = O.generate_edges(T2)
    got = outer_boundary_edge_mask(T2, edges2)
    assert np.array_equal(got, O.boundary_edges(T2, edges2))
    assert got.sum() > O.boundary_edges(m.T, edges).sum()


def test_check_krylov():
    """The solve checker: a residual above the tolerance or non-finite raises,
    unless only finiteness is asked for (inexact Newton steps)."""
    import ctypes as C
    from magpackage_amd._lib import KrylovError, check_krylov
    assert check_krylov(C.c_int(12), C.c_double(1e-13), 1e-12) == (12, 1e-13)
    with pytest.raises(KrylovError, match="after 40 iterations"):
        check_krylov(40, 2e-12, 1e-12, "x")
    with pytest.raises(KrylovError):
        check_krylov(3, float("nan"), 1e-12)
    assert check_krylov(40, 2e-3, 1e-12, finite_only=True) == (40, 2e-3)
    with pytest.raises(KrylovError):
        check_krylov(3, float("inf"), 1e-12, finite_only=True)
