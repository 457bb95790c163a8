"""Pins the CPU oracle before it is trusted as the checker of the GPU engine.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails: the reference's
own known-answer test, closed forms computed independently with numpy,
mathematical invariants of the discrete operators, and independent library
routines (numpy.unique, scipy.sparse, scipy.interpolate.PPoly).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.interpolate import PPoly

from oracle import oracle as O
from magpackage_amd.fem_tables import EDGE_NODES, barycentric_moments
from magpackage_amd.material import (MU0, create_linear, create_nonlinear, marrocco,
                                     pack_material_data)
from magpackage_amd.meshgen import kuhn_box, kuhn_counts
from tests.helpers import (core_box_mesh, core_box_nu, gradient_matrix, grads, mass_tables,
                           random_tets, rel_err, stiffness_tables, whitney_curls)

REF_CURLS = np.array([[0, -2, 2], [2, 0, -2], [-2, 2, 0], [0, 0, 2], [0, -2, 0], [2, 0, 0]],
                     dtype=np.float64)


def _block(V, e):
    """6x6 block K[r][c] of element e from the COO V (slot 36e + 6c + r)."""
    return V[36 * e: 36 * e + 36].reshape(6, 6).T


# ---------------------------------------------------------------- KAT ------
def test_kat_reference_single_tet():
    """tests/test_mex_assembly.m:31-73: unit tet, Dofs 1..6, Signs [1,-1,1,-1,1,-1],
    Nu = 500.  Known answer K_ij = (500/6) (c_i . c_j) s_i s_j, max |dK| < 1e-13."""
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    T = np.array([[1, 2, 3, 4]], dtype=np.float64)
    D = np.arange(1, 7, dtype=np.float64)[None]
    S = np.array([[1, -1, 1, -1, 1, -1]], dtype=np.float64)
    qw, curl, _ = stiffness_tables()
    I, J, V = O.curl_curl(P, T, D, S, np.array([500.0]), qw, curl)
    s = S[0]
    K_exp = (500.0 / 6.0) * (REF_CURLS @ REF_CURLS.T) * np.outer(s, s)
    assert np.max(np.abs(_block(V, 0) - K_exp)) < 1e-13
    # triplet layout: slot 6c + r -> (dof r, dof c)
    assert np.array_equal(I.reshape(6, 6), np.tile(np.arange(1, 7), (6, 1)))
    assert np.array_equal(J.reshape(6, 6), np.repeat(np.arange(1, 7), 6).reshape(6, 6))
    assert _block(V, 0)[0, 0] == pytest.approx(4000.0 / 6.0, rel=1e-15)


# ---------------------------------------------------------- closed forms ---
def test_curl_curl_closed_form_random_tets():
    """K_e = nu |T| C_r . C_c s_r s_c with C = 2 grad l_a x grad l_b (numpy inverse)."""
    P, T = random_tets(200, seed=3)
    ne = T.shape[0]
    rng = np.random.default_rng(4)
    S = rng.choice([-1.0, 1.0], size=(ne, 6))
    D = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    nu = rng.uniform(0.5, 2.0, ne) * 1e3
    qw, curl, _ = stiffness_tables()
    _, _, V = O.curl_curl(P, T.astype(np.float64), D, S, nu, qw, curl)
    for e in range(ne):
        p = P[T[e] - 1]
        C, det = whitney_curls(p)
        K = nu[e] * abs(det) / 6.0 * (C @ C.T) * np.outer(S[e], S[e])
        assert rel_err(_block(V, e), K) < 1e-12


def _mass_closed(p, moments):
    """sum_q w_q W_r . W_c |det| with W_ab = l_a grad l_b - l_b grad l_a and the
    barycentric moments Q_xy = int l_x l_y / |T_ref-volume| of a rule."""
    g, det = grads(p)
    G = g @ g.T
    M = np.zeros((6, 6))
    for r, (a, b) in enumerate(EDGE_NODES):
        for c, (k, l) in enumerate(EDGE_NODES):
            M[r, c] = (moments[a, k] * G[b, l] - moments[a, l] * G[b, k]
                       - moments[b, k] * G[a, l] + moments[b, l] * G[a, k])
    return M * abs(det)


def test_mass_closed_form_exact_and_rule():
    P, T = random_tets(100, seed=5)
    ne = T.shape[0]
    rng = np.random.default_rng(6)
    S = rng.choice([-1.0, 1.0], size=(ne, 6))
    D = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    sig = rng.uniform(1.0, 3.0, ne)
    qw, val, q_pts = mass_tables(2)
    _, _, V = O.mass(P, T.astype(np.float64), D, S, sig, qw, val)
    # exact integral: int l_a l_b = |T| (1 + delta_ab) / 20, reference volume 1/6
    exact = (np.ones((4, 4)) + np.eye(4)) / 120.0
    # moments of the truncated 4-point rule (get_quadrature_data.m:20-27), numpy
    lam = np.vstack([1.0 - q_pts.sum(axis=0), q_pts])
    rule = (lam * qw) @ lam.T
    worst_exact = worst_rule = 0.0
    for e in range(ne):
        p = P[T[e] - 1]
        ss = np.outer(S[e], S[e]) * sig[e]
        got = _block(V, e)
        worst_exact = max(worst_exact, rel_err(got, _mass_closed(p, exact) * ss))
        worst_rule = max(worst_rule, rel_err(got, _mass_closed(p, rule) * ss))
    assert worst_rule < 1e-12
    # the truncated constants move the reference off the exact integral by ~5e-9
    assert 1e-10 < worst_exact < 1e-8


def test_mass_sigma_rules():
    """sigma = 1 when Coeffs is empty; |sigma| <= 1e-12 gives a zero block; linear in sigma."""
    P, T = random_tets(20, seed=7)
    ne = T.shape[0]
    D = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    S = np.ones((ne, 6))
    qw, val, _ = mass_tables(2)
    Tf = T.astype(np.float64)
    _, _, V1 = O.mass(P, Tf, D, S, None, qw, val)
    _, _, Vo = O.mass(P, Tf, D, S, np.ones(ne), qw, val)
    assert np.array_equal(V1, Vo)
    _, _, Vw = O.mass(P, Tf, D, S, np.ones(ne + 1), qw, val)  # wrong length -> sigma = 1
    assert np.array_equal(V1, Vw)
    sig = np.full(ne, 2.5)
    sig[3] = 1e-12
    sig[4] = -5e-13
    _, _, V2 = O.mass(P, Tf, D, S, sig, qw, val)
    assert np.all(V2[36 * 3: 36 * 5] == 0.0)
    keep = np.ones(36 * ne, dtype=bool)
    keep[36 * 3: 36 * 5] = False
    assert rel_err(V2[keep], 2.5 * V1[keep]) < 1e-15


# --------------------------------------------------- global invariants -----
@pytest.fixture(scope="module")
def small_system():
    m = core_box_mesh(6, seed=11)
    mi = O.mesh_inputs(m)
    nu = core_box_nu(m)
    qw, curl, _ = stiffness_tables()
    I, J, V = O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl)
    n = mi["ndof"]
    K = O.sparse_scipy(I, J, V, n, n)
    return m, mi, nu, K


def test_stiffness_invariants(small_system):
    m, mi, nu, K = small_system
    n = K.shape[0]
    assert n == kuhn_counts(6, 6, 6)[2]
    # symmetric
    assert abs(K - K.T).max() <= 1e-12 * abs(K).max()
    # K G = 0 for the discrete gradient (curl grad = 0)
    G = gradient_matrix(mi["edges"], m.num_nodes)
    assert abs(K @ G).max() <= 1e-9 * abs(K).max()
    # positive semi-definite
    w = np.linalg.eigvalsh(K.toarray())
    assert w.min() > -1e-9 * w.max()
    # energy identity 0.5 a'Ka = sum_e 0.5 nu_e |B_e|^2 |T_e| (tests/test_energy.m:35-47)
    rng = np.random.default_rng(12)
    a = rng.normal(size=n)
    W = 0.0
    for e in range(m.num_elems):
        p = m.P[m.T[e] - 1]
        C, det = whitney_curls(p)
        B = (a[mi["cell"][e] - 1] * mi["sign"][e]) @ C
        W += 0.5 * nu[e] * (B @ B) * abs(det) / 6.0
    assert abs(0.5 * a @ (K @ a) - W) <= 1e-12 * abs(W)


def test_mass_spd():
    m = core_box_mesh(5, seed=13)
    mi = O.mesh_inputs(m)
    qw, val, _ = mass_tables(2)
    I, J, V = O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], None, qw, val)
    n = mi["ndof"]
    M = O.sparse_scipy(I, J, V, n, n).toarray()
    assert np.allclose(M, M.T, rtol=0, atol=1e-15 * abs(M).max())
    assert np.linalg.eigvalsh(M).min() > 0


# ---------------------------------------------------------- Jacobian --------
def _mat_lib_core_box():
    return {1: create_linear(1.0), 2: create_linear(1000.0)}


def test_jacobian_linear_equals_stiffness(small_system):
    """Linear materials: J == K bitwise and R == K a (MexMaterialUtils.hpp:30-34)."""
    m, mi, nu, K = small_system
    packed = pack_material_data(_mat_lib_core_box())
    qw, curl, _ = stiffness_tables()
    n = mi["ndof"]
    a = np.random.default_rng(14).normal(size=n)
    I, J, V, Ri, Rv = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"],
                                 m.RegionTags.astype(np.float64), a, qw, curl, packed)
    Ik, Jk, Vk = O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl)
    assert np.array_equal(V, Vk) and np.array_equal(I, Ik) and np.array_equal(J, Jk)
    c, r, v = O.sparse(Ri, np.ones_like(Ri), Rv, n, 1, drop_zeros=True)
    R = np.zeros(n)
    R[r] = v
    assert rel_err(R, K @ a) < 1e-12
    # residual-only mode leaves J empty and R unchanged
    I0, J0, V0, Ri0, Rv0 = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"],
                                      m.RegionTags.astype(np.float64), a, qw, curl, packed,
                                      calcJ=False)
    assert I0.size == 0 and np.array_equal(Rv0, Rv)


def test_jacobian_is_residual_derivative():
    """Nonlinear Marrocco iron: J = dR/da by central differences of the residual mode."""
    m = core_box_mesh(4, seed=15)
    mi = O.mesh_inputs(m)
    lib = {1: create_linear(1.0), 2: marrocco()}
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    n = mi["ndof"]
    rng = np.random.default_rng(16)
    a = rng.normal(size=n) * 0.02  # |B| ~ O(1) T on this mesh
    tags = m.RegionTags.astype(np.float64)

    def resid(x):
        _, _, _, Ri, Rv = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, x, qw,
                                     curl, packed, calcJ=False)
        R = np.zeros(n)
        np.add.at(R, (Ri - 1).astype(np.int64), Rv)
        return R

    I, J, V, _, _ = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, a, qw, curl,
                               packed)
    Jm = O.sparse_scipy(I, J, V, n, n)
    for k in range(3):
        d = rng.normal(size=n)
        h = 1e-7 * np.linalg.norm(a) / np.linalg.norm(d)
        fd = (resid(a + h * d) - resid(a - h * d)) / (2 * h)
        assert rel_err(Jm @ d, fd) < 2e-5


# ------------------------------------------------- B-H material eval --------
def test_eval_nu_against_ppoly_and_clamps():
    B = np.linspace(0, 2.0, 30)
    H = B / (1000 * MU0) * (1 + B ** 2)  # saturating-ish curve
    nl = create_nonlinear(B, H)
    mar = marrocco()
    lib = {1: create_linear(2.0), 2: nl, 3: mar}
    pk = pack_material_data(lib)
    rng = np.random.default_rng(17)
    # tag 2: interior points vs scipy PPoly on the same pp tables
    s = rng.uniform(0, nl.MaxBSq, 2000)
    nu, dnu = O.eval_nu(s, np.full(s.size, 2), pk)
    pp = PPoly(nl.SplineNu.coefs.T, nl.SplineNu.breaks)
    ref_nu, ref_d = pp(s), pp.derivative()(s)
    clamp = (ref_nu > nl.nu0) | (ref_nu < 1e-7)
    assert rel_err(nu[~clamp], ref_nu[~clamp]) < 1e-13
    assert rel_err(dnu[~clamp], ref_d[~clamp]) < 1e-12
    # breaks themselves (binary-search boundaries)
    sb = nl.SplineNu.breaks.copy()
    nub, _ = O.eval_nu(sb, np.full(sb.size, 2), pk)
    assert rel_err(nub, np.minimum(pp(sb), nl.nu0)) < 1e-13
    # extrapolation beyond MaxBSq: linear with the end slope (MexMaterialUtils.hpp:70-76)
    se = nl.MaxBSq + np.array([0.01, 0.5])
    nue, due = O.eval_nu(se, np.full(2, 2), pk)
    end_v, end_d = pp(nl.MaxBSq), pp.derivative()(nl.MaxBSq)
    exp_v = end_v + end_d * (se - nl.MaxBSq)
    exp_v_cl = np.where((exp_v > nl.nu0) | (exp_v < 1e-7), nl.nu0, exp_v)
    assert rel_err(nue, exp_v_cl) < 1e-13
    # linear material: (linNu, 0)
    nul, dnul = O.eval_nu(np.array([0.0, 5.0]), np.array([1, 1]), pk)
    assert np.all(nul == 1.0 / (2.0 * MU0)) and np.all(dnul == 0)
    # Marrocco above the vacuum value clamps to (nu0, 0): nu(2.0) > nu0
    nuc, duc = O.eval_nu(np.array([2.0, 3.0]), np.array([3, 3]), pk)
    assert np.all(nuc == 1.0 / MU0) and np.all(duc == 0)
    # a table that goes below 1e-7 clamps to (nu0, 0) too
    neg = create_nonlinear(np.array([0.0, 1.0, 2.0, 3.0]), np.array([0.0, 1e-9, 2e-9, 3e-9]))
    pk2 = pack_material_data({1: neg})
    nun, dun = O.eval_nu(np.array([1.0]), np.array([1]), pk2)
    assert nun[0] == neg.nu0 and dun[0] == 0


# ---------------------------------------------------------- sparse() --------
def test_sparse_against_scipy():
    rng = np.random.default_rng(18)
    m, n, nt = 40, 30, 3000
    I = rng.integers(0, m + 1, nt).astype(np.float64)  # 0 = filtered out
    J = rng.integers(0, n + 1, nt).astype(np.float64)
    V = rng.integers(-3, 4, nt).astype(np.float64)     # integer values: order-free sums
    c, r, v = O.sparse(I, J, V, m, n, drop_zeros=True)
    keep = (I > 0) & (J > 0)
    ref = sp.coo_matrix((V[keep], (I[keep] - 1, J[keep] - 1)), shape=(m, n)).tocsc()
    ref.sum_duplicates()
    ref.eliminate_zeros()
    ref.sort_indices()
    assert np.array_equal(c, ref.indptr) and np.array_equal(r, ref.indices)
    assert np.array_equal(v, ref.data)
    # without dropping: the structural pattern keeps the zero sums
    c2, r2, v2 = O.sparse(I, J, V, m, n, drop_zeros=False)
    assert r2.size >= r.size and np.count_nonzero(v2 == 0) == r2.size - r.size
    # duplicates are summed in triplet order (floating point)
    Vf = rng.normal(size=nt)
    _, _, vf = O.sparse(I, J, Vf, m, n, drop_zeros=True)
    reff = sp.coo_matrix((Vf[keep], (I[keep] - 1, J[keep] - 1)), shape=(m, n)).tocsc()
    reff.sum_duplicates()
    reff.sort_indices()
    assert rel_err(vf, reff.data) < 1e-14
    with pytest.raises(IndexError):
        O.sparse(np.array([m + 1.0]), np.array([1.0]), np.array([1.0]), m, n)


def test_sparse_par_matches_serial_bitwise():
    """The threaded sparse() (used to time the host path) returns what the
    serial restatement does, bit for bit: buckets past the insertion-sort
    size, exact cancellations, filtered triplets, both drop modes."""
    rng = np.random.default_rng(41)
    for m, n, nt in ((40, 30, 6000), (5000, 4000, 20000), (3, 3, 0)):
        I = rng.integers(0, m + 1, nt).astype(np.float64)
        J = rng.integers(0, n + 1, nt).astype(np.float64)
        V = rng.normal(size=nt)
        if nt:
            # exact cancellations: a value and its negative on the same entry
            k = rng.choice(nt, size=nt // 10, replace=False)
            I = np.concatenate([I, I[k]]); J = np.concatenate([J, J[k]]); V = np.concatenate([V, -V[k]])
        for dz in (True, False):
            c1, r1, v1 = O.sparse(I, J, V, m, n, drop_zeros=dz)
            c2, r2, v2 = O.sparse_par(I, J, V, m, n, drop_zeros=dz)
            assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
            assert np.array_equal(v1.view(np.int64), v2.view(np.int64))
    with pytest.raises(IndexError):
        O.sparse_par(np.array([1.0]), np.array([n + 1.0]), np.array([1.0]), m, n)


def test_csc_add_jw_against_scipy():
    rng = np.random.default_rng(42)
    m = n = 60
    w = 2 * np.pi * 50.0
    mats = []
    for nt in (900, 150):
        I = rng.integers(1, m + 1, nt).astype(np.float64)
        J = rng.integers(1, n + 1, nt).astype(np.float64)
        mats.append(O.sparse(I, J, rng.normal(size=nt), m, n))
    c, r, v = O.csc_add_jw(mats[0], mats[1], w, n)
    K = sp.csc_matrix((mats[0][2], mats[0][1], mats[0][0]), shape=(m, n))
    M = sp.csc_matrix((mats[1][2], mats[1][1], mats[1][0]), shape=(m, n))
    ref = (K + 1j * w * M).tocsc()
    ref.sort_indices()
    assert np.array_equal(c, ref.indptr) and np.array_equal(r, ref.indices)
    assert np.array_equal(v, ref.data)


# ------------------------------------------------------ topology / DOFs -----
def test_edges_against_numpy_unique():
    mesh = kuhn_box(5, 4, 3, seed=19)
    edges, t2e, sign = O.generate_edges(mesh.T)
    raw = np.stack([mesh.T[:, [a for a, _ in EDGE_NODES]].ravel(),
                    mesh.T[:, [b for _, b in EDGE_NODES]].ravel()], axis=1)
    srt = np.sort(raw, axis=1)
    u, inv = np.unique(srt, axis=0, return_inverse=True)
    assert np.array_equal(edges, u)
    assert np.array_equal(t2e.ravel(), inv.ravel() + 1)
    assert np.array_equal(sign.ravel(), np.where(raw[:, 0] == srt[:, 0], 1, -1))
    assert edges.shape[0] == kuhn_counts(5, 4, 3)[2]


def test_boundary_edges_brute_force():
    mesh = kuhn_box(3, 3, 2, seed=20)
    edges, _, _ = O.generate_edges(mesh.T)
    got = O.boundary_edges(mesh.T, edges)
    # brute force: an edge is on the boundary iff all its nodes lie on the box surface
    # and it belongs to a face seen once; for a Kuhn box the surface test suffices
    P = mesh.P
    onb = lambda x: np.isclose(x[:, 0], 0) | np.isclose(x[:, 0], 1) | np.isclose(x[:, 1], 0) \
        | np.isclose(x[:, 1], 1) | np.isclose(x[:, 2], 0) | np.isclose(x[:, 2], 1)
    a, b = P[edges[:, 0] - 1], P[edges[:, 1] - 1]
    same_face = np.zeros(edges.shape[0], dtype=bool)
    for ax in range(3):
        for v in (0.0, 1.0):
            same_face |= np.isclose(a[:, ax], v) & np.isclose(b[:, ax], v)
    assert np.array_equal(got, same_face & onb(a) & onb(b))


def test_distribute_dofs_restricted():
    mesh = core_box_mesh(4, seed=21)
    edges, t2e, _ = O.generate_edges(mesh.T)
    act = mesh.RegionTags == 2
    emap, cell, nd = O.distribute_dofs(t2e, edges.shape[0], act, offset=7)
    used = np.unique(t2e[act])
    assert nd == used.size
    assert np.array_equal(emap[used - 1], 7 + np.arange(1, nd + 1))
    assert np.count_nonzero(emap) == nd
    assert np.array_equal(cell, emap[t2e - 1])


# ---------------------------------------------------- source and B ---------
def test_source_closed_form_random_tets():
    """F_i = s_i Js . int W_i = s_i Js . (|T|/4)(grad l_b - grad l_a): the
    4-point rule integrates the linear Whitney functions exactly (its
    truncated constants still sum a + 3b = 1)."""
    P, T = random_tets(150, seed=21)
    ne = T.shape[0]
    rng = np.random.default_rng(22)
    S = rng.choice([-1.0, 1.0], size=(ne, 6))
    D = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    Js = rng.normal(size=(ne, 3)) * 3e6
    Js[::7] = 0.0  # zero-source elements give zeros
    qw, val, _ = mass_tables(2)
    I, V = O.source(P, T.astype(np.float64), D, S, Js, qw, val)
    assert np.array_equal(I, D.ravel())
    worst = 0.0
    for e in range(ne):
        g, det = grads(P[T[e] - 1])
        vol = abs(det) / 6.0
        ref = np.array([S[e, i] * Js[e] @ (g[b] - g[a]) * vol / 4.0
                        for i, (a, b) in enumerate(EDGE_NODES)])
        if not Js[e].any():
            assert np.all(V[6 * e:6 * e + 6] == 0.0)
            continue
        worst = max(worst, rel_err(V[6 * e:6 * e + 6], ref))
    assert worst < 1e-12


def test_element_b_against_whitney_curls():
    """computeElementB restatement vs curl W_ab = 2 grad l_a x grad l_b."""
    P, T = random_tets(80, seed=23)
    ne = T.shape[0]
    rng = np.random.default_rng(24)
    S = rng.choice([-1.0, 1.0], size=(ne, 6))
    D = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    D[5, 2] = 0.0  # an inactive DOF contributes nothing
    sol = rng.normal(size=6 * ne)
    _, curl, _ = stiffness_tables()
    B = O.element_b(P, T.astype(np.float64), D, S, sol, curl)
    for e in range(ne):
        c, _ = whitney_curls(P[T[e] - 1])
        a = np.where(D[e] > 0, sol[np.maximum(D[e].astype(int), 1) - 1], 0.0) * S[e]
        assert rel_err(B[e], a @ c) < 1e-12


def _grad_ref_table():
    from magpackage_amd.fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
    q_pts, qw = get_quadrature_data("tet", 2)
    _, lg = lagrange_tet_p1(q_pts)
    return qw, matlab_flat(lg)


def test_scalar_laplacian_closed_form_random_tets():
    """K_ij = sigma |T| grad l_i . grad l_j (P1 gradients are constant), zero
    blocks for |sigma| <= 1e-12, node DOFs passed through (0 = inactive)."""
    P, T = random_tets(120, seed=31)
    ne = T.shape[0]
    rng = np.random.default_rng(32)
    D = rng.integers(0, 50, size=(ne, 4)).astype(np.float64)
    sig = rng.uniform(1e6, 5e7, ne)
    sig[::5] = 1e-13
    qw, G = _grad_ref_table()
    I, J, V = O.scalar_laplacian(P, T.astype(np.float64), D, sig, qw, G)
    worst = 0.0
    for e in range(ne):
        g, det = grads(P[T[e] - 1])
        K = sig[e] * abs(det) / 6.0 * (g @ g.T)
        blk = slice(16 * e, 16 * e + 16)
        assert np.array_equal(I[blk], np.tile(D[e], 4))
        assert np.array_equal(J[blk], np.repeat(D[e], 4))
        if sig[e] <= 1e-12:
            assert not V[blk].any()
            continue
        worst = max(worst, rel_err(V[blk], K.T.ravel()))
    assert worst < 1e-12


def test_coupling_closed_form_random_tets():
    """C_ij = s_i sigma (grad l_j) . int W_i = s_i sigma (|T|/4)(grad l_b - grad l_a) .
    grad l_j: the 4-point rule integrates the linear Whitney functions exactly."""
    P, T = random_tets(120, seed=33)
    ne = T.shape[0]
    rng = np.random.default_rng(34)
    S = rng.choice([-1.0, 1.0], size=(ne, 6))
    DE = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    DN = rng.integers(1, 40, size=(ne, 4)).astype(np.float64)
    sig = rng.uniform(1e6, 5e7, ne)
    sig[::6] = 0.0
    qw, G = _grad_ref_table()
    _, val, _ = mass_tables(2)
    I, J, V = O.coupling(P, T.astype(np.float64), DE, S, DN, sig, qw, val, G)
    worst = 0.0
    for e in range(ne):
        g, det = grads(P[T[e] - 1])
        vol = abs(det) / 6.0
        C = np.array([[S[e, i] * sig[e] * vol / 4.0 * (g[b] - g[a]) @ g[j] for j in range(4)]
                      for i, (a, b) in enumerate(EDGE_NODES)])
        blk = slice(24 * e, 24 * e + 24)
        assert np.array_equal(I[blk], np.tile(DE[e], 4))
        assert np.array_equal(J[blk], np.repeat(DN[e], 6))
        if sig[e] == 0.0:
            assert not V[blk].any()
            continue
        worst = max(worst, rel_err(V[blk], C.T.ravel()))
    assert worst < 1e-12


def test_winding_closed_form_random_tets():
    """F_i = s_i Scale J_dir . (|T|/4)(grad l_b - grad l_a) on the coil's
    elements, (0, 0) pairs elsewhere; DirField per element or UniformDir."""
    P, T = random_tets(90, seed=35)
    ne = T.shape[0]
    rng = np.random.default_rng(36)
    S = rng.choice([-1.0, 1.0], size=(ne, 6))
    D = np.arange(1, 6 * ne + 1, dtype=np.float64).reshape(ne, 6)
    tags = rng.choice([1.0, 5.0, 7.0], size=ne)
    dirf = rng.normal(size=(ne, 3))
    uni = np.array([0.3, -0.5, 0.8])
    qw, val, _ = mass_tables(2)
    for field in (dirf, None):
        I, V = O.winding(P, T.astype(np.float64), D, S, tags, 5, field, uni, 250.0, qw, val)
        worst = 0.0
        for e in range(ne):
            blk = slice(6 * e, 6 * e + 6)
            if tags[e] != 5.0:
                assert not I[blk].any() and not V[blk].any()
                continue
            g, det = grads(P[T[e] - 1])
            jd = 250.0 * (dirf[e] if field is not None else uni)
            ref = np.array([S[e, i] * jd @ (g[b] - g[a]) * abs(det) / 24.0
                            for i, (a, b) in enumerate(EDGE_NODES)])
            assert np.array_equal(I[blk], D[e])
            worst = max(worst, rel_err(V[blk], ref))
        assert worst < 1e-12


def test_hbfem_linear_closed_form():
    """For a linear material the harmonic-balance residual of harmonic h is
    K(nu) A_h (the DFT of the periodic field at the samples recovers each
    harmonic, Nt > 2 max h), and the Jacobian blocks are K(nu) scaled by 1
    (DC) or 2 (AC)."""
    from magpackage_amd.hbfem import AFT
    mesh = kuhn_box(3, 3, 2, seed=41)
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    aft = AFT([0, 1, 3], 16)
    H = aft.NumHarmonics
    rng = np.random.default_rng(42)
    A = rng.normal(size=(n, H)) + 1j * rng.normal(size=(n, H))
    A[:, 0] = A[:, 0].real  # the DC coefficient is real
    lib = {1: create_linear(1.0), 2: create_linear(250.0)}
    tags = np.where(rng.uniform(size=mesh.num_elems) < 0.5, 1.0, 2.0)
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    sc = aft.jacobian_scalings()
    I, J, V, Rr, Rc, Rvr, Rvi = O.hbfem(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, A.real,
                                        A.imag, aft.D_matrix.real, aft.D_matrix.imag,
                                        aft.P_matrix.real, aft.P_matrix.imag, qw, curl, packed,
                                        True, sc)
    nu = np.where(tags == 2.0, packed["linNu"][1], packed["linNu"][0])
    K = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl), n, n)
    R = np.zeros((n, H), dtype=complex)
    np.add.at(R, (Rr.astype(int) - 1, Rc.astype(int) - 1), Rvr + 1j * Rvi)
    ref = K @ A
    assert np.abs(R - ref).max() / np.abs(ref).max() < 1e-12
    Jm = sp.csc_matrix((V, (I.astype(int) - 1, J.astype(int) - 1)), shape=(n * H, n * H))
    Jref = sp.block_diag([K * f for f in sc]).tocsc()
    assert abs(Jm - Jref).max() / abs(Jref).max() < 1e-12
