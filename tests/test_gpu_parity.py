"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bar: bit-exact for integer work (edges, DOFs, patterns) and for the
FP64_EXACT precision mode; <= 1e-12 relative (max|dA| / max|A|) for FP64;
<= 1e-6 for FP32_LOCAL.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from magpackage_amd.material import create_linear, marrocco, pack_material_data, MU0
from magpackage_amd.meshgen import kuhn_box, kuhn_counts
from tests.helpers import (core_box_mesh, core_box_nu, csc_equal_bits, csc_from_plan, element_B,
                           gradient_matrix, mass_tables, random_tets, rel_err, stiffness_tables)

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32_local": 1e-6}


def _t(a, dev, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev) if dtype is None else \
        torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)


def _setup(mesh, gpu, active_tags=None):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    M = Mesh.from_kuhn(mesh) if hasattr(mesh, "shape") else mesh
    dh = DofHandler(M)
    sp_ = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(sp_, active_tags)
    return M, dh, sp_, Assembler(M, dh)


# ------------------------------------------------------------ topology -----
def test_generate_edges_and_dofs_bit_exact(gpu):
    mesh = core_box_mesh(12, seed=2)
    M, dh, space, asm = _setup(mesh, gpu)
    edges, t2e, sign = O.generate_edges(mesh.T)
    assert np.array_equal(M.Edges.cpu().numpy(), edges)
    assert np.array_equal(M.T2E.cpu().numpy(), t2e)
    assert np.array_equal(M.T2E_Sign.cpu().numpy(), sign)
    emap, cell, nd = O.distribute_dofs(t2e, edges.shape[0])
    assert dh.NumGlobalDofs == nd == kuhn_counts(12, 12, 12)[2]
    assert np.array_equal(dh.DofMaps["Nedelec_P1"].cpu().numpy(), cell)


def test_distribute_dofs_restricted_bit_exact(gpu):
    mesh = core_box_mesh(8, seed=3)
    M, dh, space, asm = _setup(mesh, gpu, active_tags=[2])
    edges, t2e, _ = O.generate_edges(mesh.T)
    emap, cell, nd = O.distribute_dofs(t2e, edges.shape[0], mesh.RegionTags == 2)
    assert dh.NumGlobalDofs == nd
    assert np.array_equal(dh.DofMaps["Nedelec_P1"].cpu().numpy(), cell)
    assert np.array_equal(dh.EntityMaps["Nedelec_P1"].cpu().numpy(), emap)


# -------------------------------------------------- MEX-compatible COO -----
def test_coo_kernels_bit_exact_kat(gpu):
    """tests/test_mex_assembly.m:31-46 single tet, through mpk_curl_curl_coo."""
    from magpackage_amd import mex
    P = np.array([[0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)  # [3 x 4]
    T = np.array([[1], [2], [3], [4]], dtype=np.float64)
    Dofs = np.arange(1, 7, dtype=np.float64)[:, None]
    Signs = np.array([1, -1, 1, -1, 1, -1], dtype=np.float64)[:, None]
    qw, curl, _ = stiffness_tables()
    I, J, V = mex.assemble_curl_curl_kernel_mex(P, T, Dofs, Signs, 500.0, qw, curl)
    Io, Jo, Vo = O.curl_curl(P.T, T.T, Dofs.T, Signs.T, np.array([500.0]), qw, curl)
    assert np.array_equal(I, Io) and np.array_equal(J, Jo)
    assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))


@pytest.mark.parametrize("ne_cubes", [(1, 1, 1), (7, 5, 3), (20, 20, 20)])
def test_coo_kernels_bit_exact_mesh(gpu, ne_cubes):
    from magpackage_amd import mex
    mesh = core_box_mesh(1) if ne_cubes == (1, 1, 1) else kuhn_box(*ne_cubes, seed=4)
    if ne_cubes == (20, 20, 20):
        mesh = core_box_mesh(20, seed=1)
    mi = O.mesh_inputs(mesh)
    rng = np.random.default_rng(5)
    ne = mesh.num_elems
    nu = core_box_nu(mesh) if ne_cubes == (20, 20, 20) else rng.uniform(1, 2, ne)
    qw, curl, _ = stiffness_tables()
    args = (mi["P"].T, mi["T"].T, mi["Dofs"].T, mi["Signs"].T)
    I, J, V = mex.assemble_curl_curl_kernel_mex(*args, nu, qw, curl)
    Io, Jo, Vo = O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl)
    assert np.array_equal(I, Io) and np.array_equal(J, Jo)
    assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))
    # mass, with a sigma vector that has zero regions, and with [] (sigma = 1)
    qm, val, _ = mass_tables(2)
    sig = np.where(rng.uniform(size=ne) < 0.5, 0.0, rng.uniform(1e6, 5e7, ne))
    for coeff in (sig, np.zeros(0)):
        I, J, V = mex.assemble_mass_kernel_mex(*args, coeff, qm, val)
        Io, Jo, Vo = O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"],
                            coeff if coeff.size else None, qm, val)
        assert np.array_equal(I, Io) and np.array_equal(J, Jo)
        assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))
    # Jacobian with Marrocco iron + linear air, both calcJ modes
    lib = {1: create_linear(1.0), 2: marrocco()}
    packed = pack_material_data(lib)
    tags = np.where(rng.uniform(size=ne) < 0.5, 1.0, 2.0)
    a = rng.normal(size=mi["ndof"]) * (0.05 / ne_cubes[0])
    for calcJ in (True, False):
        out = mex.assemble_jacobian_kernel_mex(*args, tags, a, qw, curl, *[
            packed[k] for k in ("linNu", "isNon", "maxB", "breaks", "coefs", "starts", "counts",
                                "coe_starts")], float(calcJ))
        ref = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, a, qw, curl, packed,
                         calcJ=calcJ)
        for g, r in zip(out, ref):
            assert g.size == r.size
            assert np.array_equal(np.asarray(g).view(np.uint64), r.view(np.uint64))


def _nodal_tables():
    from magpackage_amd.fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
    q_pts, qw = get_quadrature_data("tet", 2)
    _, lg = lagrange_tet_p1(q_pts)
    _, val, _ = mass_tables(2)
    return qw, matlab_flat(lg), val


def test_nodal_and_winding_coo_bit_exact(gpu):
    """assemble_scalar_laplacian / coupling / winding_kernel_mex drop-ins against
    the oracle, bit for bit (a zero-sigma region, inactive node DOFs, both
    DirField forms, elements outside the coil)."""
    from magpackage_amd import mex
    mesh = core_box_mesh(6, seed=9)
    mi = O.mesh_inputs(mesh)
    ne = mesh.num_elems
    rng = np.random.default_rng(40)
    qw, G, val = _nodal_tables()
    T = mi["T"]
    # node DOFs: a permutation of the nodes, node 1 inactive (0)
    perm = rng.permutation(mesh.num_nodes) + 1.0
    perm[0] = 0.0
    DN = perm[(T - 1).astype(np.int64)]
    sig = np.where(mesh.RegionTags == 2, 0.0, rng.uniform(1e6, 5e7, ne))
    I, J, V = mex.assemble_scalar_laplacian_kernel_mex(mi["P"].T, T.T, DN.T, sig, qw, G)
    Io, Jo, Vo = O.scalar_laplacian(mi["P"], T, DN, sig, qw, G)
    assert np.array_equal(I, Io) and np.array_equal(J, Jo)
    assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))
    I, J, V = mex.assemble_coupling_kernel_mex(mi["P"].T, T.T, mi["Dofs"].T, mi["Signs"].T, DN.T,
                                               sig, qw, val, G)
    Io, Jo, Vo = O.coupling(mi["P"], T, mi["Dofs"], mi["Signs"], DN, sig, qw, val, G)
    assert np.array_equal(I, Io) and np.array_equal(J, Jo)
    assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))
    tags = mesh.RegionTags.astype(np.float64)
    dirf = rng.normal(size=(ne, 3))
    for field in (dirf, None):
        I, V = mex.assemble_winding_kernel_mex(mi["P"].T, T.T, mi["Dofs"].T, mi["Signs"].T, tags,
                                               2, field.T if field is not None else np.zeros(0),
                                               np.array([0.0, 0.0, 1.0]), 3.5e6, qw, val)
        Io, Vo = O.winding(mi["P"], T, mi["Dofs"], mi["Signs"], tags, 2, field,
                           np.array([0.0, 0.0, 1.0]), 3.5e6, qw, val)
        assert np.array_equal(I, Io)
        assert np.array_equal(V.view(np.uint64), Vo.view(np.uint64))
    with pytest.raises(Exception, match="Expected 6 inputs"):
        mex.assemble_scalar_laplacian_kernel_mex(mi["P"].T, T.T, DN.T, sig, qw)


def _hbfem_args(mi, tags, A, aft, packed, qw, curl, request_j, imag=True):
    sc = aft.jacobian_scalings()
    mat = [packed[k] for k in ("linNu", "isNon", "maxB", "breaks", "coefs", "starts", "counts")]
    return (mi["P"].T, mi["T"].T, mi["Dofs"].T, mi["Signs"].T, tags, A.real,
            A.imag if imag else np.zeros(0), aft.D_matrix.real, aft.D_matrix.imag,
            aft.P_matrix.real, aft.P_matrix.imag, qw, curl, *mat, float(request_j), sc)


@pytest.mark.parametrize("request_j", [True, False])
def test_hbfem_coo_bit_exact(gpu, request_j):
    """assemble_hbfem_kernel_mex drop-in against the oracle, bit for bit:
    harmonics [0, 1, 3, 5] on 24 samples, a Marrocco core driven into
    saturation, complex phasors and the real-phasor (empty SolH_Imag) form; the
    coefficient offsets derived as packMaterialData builds them."""
    from magpackage_amd import mex
    from magpackage_amd.hbfem import AFT
    mesh = core_box_mesh(6, seed=12)
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    aft = AFT([0, 1, 3, 5], 24)
    rng = np.random.default_rng(44)
    A = (rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))) * 0.02
    lib = {1: create_linear(1.0), 2: marrocco()}
    packed = pack_material_data(lib)
    tags = mesh.RegionTags.astype(np.float64)
    qw, curl, _ = stiffness_tables()
    for imag in (True, False):
        Ai = A if imag else A.real + 0j
        got = mex.assemble_hbfem_kernel_mex(*_hbfem_args(mi, tags, Ai, aft, packed, qw, curl,
                                                         request_j, imag))
        ref = O.hbfem(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, Ai.real,
                      Ai.imag if imag else None, aft.D_matrix.real, aft.D_matrix.imag,
                      aft.P_matrix.real, aft.P_matrix.imag, qw, curl, packed, request_j,
                      aft.jacobian_scalings())
        for g, r in zip(got, ref):
            assert g.size == r.size
            assert np.array_equal(np.asarray(g).view(np.uint64), r.view(np.uint64))
    with pytest.raises(Exception, match="Need 22 inputs"):
        mex.assemble_hbfem_kernel_mex(*_hbfem_args(mi, tags, A, aft, packed, qw, curl, True)[:21])


def test_assemble_hbfem_linear_and_nonlinear(gpu):
    """Assembler.assembleHBFEM: for linear materials R = K(nu) A per harmonic and
    the Jacobian is blockdiag(K * [1, 2, 2]); with a saturating core, R equals the
    oracle's triplets summed per (DOF, harmonic)."""
    import scipy.sparse as sps
    from magpackage_amd.hbfem import AFT
    mesh = core_box_mesh(6, seed=13)
    M_, dh, space, asm = _setup(mesh, gpu)
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    aft = AFT([0, 1, 3], 16)
    rng = np.random.default_rng(45)
    A = (rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))) * 0.02
    A[:, 0] = A[:, 0].real
    qw, curl, _ = stiffness_tables()
    lib = {1: create_linear(1.0), 2: create_linear(400.0)}
    J_tri, R = asm.assembleHBFEM(space, A, aft, lib, True)
    packed = pack_material_data(lib)
    nu = np.where(mesh.RegionTags == 2, packed["linNu"][1], packed["linNu"][0])
    K = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl), n, n)
    ref = K @ A
    assert np.abs(R.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-12
    Jm = sps.csc_matrix((J_tri["V"].cpu().numpy(), (J_tri["I"].cpu().numpy().astype(int) - 1,
                                                    J_tri["J"].cpu().numpy().astype(int) - 1)),
                        shape=(3 * n, 3 * n))
    Jref = sps.block_diag([K, 2 * K, 2 * K]).tocsc()
    assert abs(Jm - Jref).max() / abs(Jref).max() < 1e-12
    lib = {1: create_linear(1.0), 2: marrocco()}
    packed = pack_material_data(lib)
    _, R = asm.assembleHBFEM(space, A * 5, aft, lib, False)
    out = O.hbfem(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], mesh.RegionTags.astype(np.float64),
                  (A * 5).real, (A * 5).imag, aft.D_matrix.real, aft.D_matrix.imag,
                  aft.P_matrix.real, aft.P_matrix.imag, qw, curl, packed, False,
                  aft.jacobian_scalings())
    Rr = np.zeros((n, 3), dtype=complex)
    np.add.at(Rr, (out[3].astype(int) - 1, out[4].astype(int) - 1), out[5] + 1j * out[6])
    assert np.abs(R.cpu().numpy() - Rr).max() / np.abs(Rr).max() < 1e-13


def test_coo_index_errors(gpu):
    from magpackage_amd import mex
    from magpackage_amd._lib import MpkError
    qw, curl, _ = stiffness_tables()
    P = np.zeros((3, 4))
    T = np.array([[1], [2], [3], [9]], dtype=np.float64)  # node 9 does not exist
    with pytest.raises(MpkError):
        mex.assemble_curl_curl_kernel_mex(P, T, np.ones((6, 1)), np.ones((6, 1)), 1.0, qw, curl)
    with pytest.raises(MpkError) as ei:
        mex.assemble_curl_curl_kernel_mex(P, T, np.ones((6, 1)))
    assert "MagPackage:Args" in str(ei.value)


# --------------------------------------------------------------- sparse() --
def test_sparse_coo_to_csc_bit_exact(gpu):
    from magpackage_amd import mex
    rng = np.random.default_rng(6)
    m, n, nt = 500, 400, 200000
    I = rng.integers(0, m + 1, nt).astype(np.float64)
    J = rng.integers(0, n + 1, nt).astype(np.float64)
    V = rng.normal(size=nt)
    V[rng.uniform(size=nt) < 0.05] = 0.0
    for drop in (True, False):
        gc, gr, gv = mex.sparse_gpu(I, J, V, m, n, drop_zeros=drop)
        rc, rr, rv = O.sparse(I, J, V, m, n, drop_zeros=drop)
        assert np.array_equal(gc, rc)
        assert np.array_equal(gr.astype(np.int64), rr)
        assert np.array_equal(gv.view(np.uint64), rv.view(np.uint64))


# ------------------------------------------------------------ CSC engine ---
@pytest.fixture(scope="module")
def core20(gpu):
    mesh = core_box_mesh(20, seed=1)
    M, dh, space, asm = _setup(mesh, gpu)
    mi = O.mesh_inputs(mesh)
    return mesh, M, dh, space, asm, mi


@pytest.fixture(params=["ring", "element"])
def driver(request, monkeypatch):
    """Closed-form modes: rotation-order (default) and element-order drivers."""
    monkeypatch.setenv("MPK_CSC_DRIVER", request.param)
    return request.param


def _one_exact_driver(prec, driver):
    if prec == "fp64_exact" and driver != "ring":
        pytest.skip("the exact mode has a single (element-order) driver")


def test_plan_pattern_matches_sparse(core20):
    mesh, M, dh, space, asm, mi = core20
    plan = asm.plan(space)
    qw, curl, _ = stiffness_tables()
    I, J, V = O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], np.ones(mesh.num_elems),
                          qw, curl)
    c, r, _ = O.sparse(I, J, V, mi["ndof"], mi["ndof"], drop_zeros=False)
    colptr, rowind = plan.pattern()
    assert plan.nnz == r.size == 938060  # nnz(K) of the 20^3 Kuhn box
    assert plan.ring_order  # a Kuhn mesh is manifold: every column is a ring or chain
    assert np.array_equal(colptr.cpu().numpy(), c)
    assert np.array_equal(rowind.cpu().numpy(), r)


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_local"])
def test_csc_stiffness(core20, prec, driver):
    _one_exact_driver(prec, driver)
    mesh, M, dh, space, asm, mi = core20
    nu = core_box_nu(mesh)
    qw, curl, _ = stiffness_tables()
    ref = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl),
                         mi["ndof"], mi["ndof"])
    K = asm.assembleStiffness(space, nu, precision=prec).to_scipy()
    if prec == "fp64_exact":
        assert csc_equal_bits(K, ref)
    else:
        K2 = asm.assembleStiffness(space, nu, precision=prec, drop_zeros=False).to_scipy()
        assert np.array_equal(K2.indptr, ref.indptr) or ref.nnz < K2.nnz
        d = (K2 - ref)
        assert abs(d).max() / abs(ref).max() < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_local"])
def test_csc_mass(core20, prec, driver):
    _one_exact_driver(prec, driver)
    mesh, M, dh, space, asm, mi = core20
    qm, val, _ = mass_tables(2)
    rng = np.random.default_rng(7)
    sig = np.where(mesh.RegionTags == 2, 3.5e7, 0.0)
    for coeff in (None, sig):
        ref = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], coeff, qm, val),
                             mi["ndof"], mi["ndof"])
        if coeff is None:
            Mg = asm.assembleMass(space, precision=prec).to_scipy()
        else:
            Mg = asm.assembleMassWeighted(space, coeff, precision=prec).to_scipy()
        if prec == "fp64_exact":
            assert csc_equal_bits(Mg, ref)
        else:
            assert abs(Mg - ref).max() / abs(ref).max() < TOL[prec]


def test_source_vector_bit_exact(core20):
    """assemble_source_kernel_mex (COO) and assembleSource (per-DOF sums in
    triplet order) against the oracle + sparse(I, 1, V, N, 1), bitwise."""
    from magpackage_amd.mex import assemble_source_kernel_mex
    mesh, M, dh, space, asm, mi = core20
    qm, val, _ = mass_tables(2)
    rng = np.random.default_rng(31)
    Js = rng.normal(size=(mesh.num_elems, 3)) * 3e6
    Js[mesh.RegionTags == 2] = 0.0
    I, V = O.source(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], Js, qm, val)
    n = mi["ndof"]
    # MATLAB-shaped arguments: P [3 x Nn], T [4 x Ne], Dofs / Signs [6 x Ne], Js [3 x Ne]
    Ig, Vg = assemble_source_kernel_mex(mi["P"].T, mi["T"].T, mi["Dofs"].T, mi["Signs"].T,
                                        Js.T, qm, val)
    assert np.array_equal(Ig, I) and np.array_equal(Vg.view(np.uint64), V.view(np.uint64))
    c, r, v = O.sparse(I, np.ones_like(I), V, n, 1, drop_zeros=False)
    Fref = np.zeros(n)
    Fref[r] = v
    F = asm.assembleSource(space, Js).cpu().numpy()
    assert np.array_equal(F.view(np.uint64), Fref.view(np.uint64))
    # a region map {tag: (jx, jy, jz)} as Assembler.expandSourceMap takes it
    F2 = asm.assembleSource(space, {1: (0.0, 0.0, 3e6)}).cpu().numpy()
    Jm = np.zeros((mesh.num_elems, 3))
    Jm[mesh.RegionTags == 1, 2] = 3e6
    I2, V2 = O.source(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], Jm, qm, val)
    F2ref = np.zeros(n)
    c, r, v = O.sparse(I2, np.ones_like(I2), V2, n, 1, drop_zeros=False)
    F2ref[r] = v
    assert np.array_equal(F2.view(np.uint64), F2ref.view(np.uint64))


def test_element_b_bit_exact(core20):
    mesh, M, dh, space, asm, mi = core20
    _, curl, _ = stiffness_tables()
    a = np.random.default_rng(32).normal(size=mi["ndof"])
    Bref = O.element_b(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], a, curl)
    B = asm.computeElementB(a, space).cpu().numpy()
    assert np.array_equal(B.view(np.uint64), Bref.view(np.uint64))


@pytest.mark.parametrize("prec", ["fp64", "fp32_local"])
def test_csc_preloaded_coordinates(core20, prec, driver):
    """P == NULL reads the coordinates loaded by mpk_plan_set_coords: values
    are bit-identical to the call given P; FP64_EXACT and a plan that never
    saw coordinates reject P == NULL."""
    import ctypes as C
    import torch
    from magpackage_amd import _lib as L
    from magpackage_amd.assembler import Plan
    mesh, M, dh, space, asm, mi = core20
    nu = torch.from_numpy(core_box_nu(mesh)).cuda()
    q = asm.stiffness_quadrature()
    pr = L.PRECISIONS[prec]
    fresh = Plan(M, dh.DofMaps[space.toString()], dh.NumGlobalDofs)
    v1 = torch.empty(fresh.nnz, dtype=torch.float64, device="cuda")
    with pytest.raises(L.MpkError, match="no coordinates"):
        L.check(L.lib.mpk_csc_curl_curl(fresh.handle, None, L.ptr(nu), C.byref(q), pr,
                                        L.ptr(v1), L.stream_ptr()))
    L.check(L.lib.mpk_csc_curl_curl(fresh.handle, L.ptr(M.P_dev), L.ptr(nu), C.byref(q), pr,
                                    L.ptr(v1), L.stream_ptr()))
    L.check(L.lib.mpk_plan_set_coords(fresh.handle, L.ptr(M.P_dev), L.stream_ptr()))
    v2 = torch.full_like(v1, float("nan"))
    L.check(L.lib.mpk_csc_curl_curl(fresh.handle, None, L.ptr(nu), C.byref(q), pr, L.ptr(v2),
                                    L.stream_ptr()))
    assert torch.equal(v1.view(torch.int64), v2.view(torch.int64))
    with pytest.raises(L.MpkError, match="needs the coordinate array"):
        L.check(L.lib.mpk_csc_curl_curl(fresh.handle, None, L.ptr(nu), C.byref(q),
                                        L.PREC_FP64_EXACT, L.ptr(v2), L.stream_ptr()))


def test_mass_weighted_empty_when_sigma_zero(core20):
    """With sigma = 0 everywhere, sparse() of the weighted mass is empty."""
    mesh, M, dh, space, asm, mi = core20
    Mg = asm.assembleMassWeighted(space, np.zeros(mesh.num_elems), precision="fp64_exact")
    assert Mg.nnz == 0


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_local"])
def test_csc_frequency_operator(core20, prec, driver):
    _one_exact_driver(prec, driver)
    mesh, M, dh, space, asm, mi = core20
    nu = core_box_nu(mesh)
    sig = np.where(mesh.RegionTags == 2, 3.5e7, 0.0)
    omega = 2 * np.pi * 50.0
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    n = mi["ndof"]
    K = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl), n, n)
    Ms = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sig, qm, val), n, n)
    ref = (K + 1j * omega * Ms).tocsc()
    A = asm.assembleFrequencyOperator(space, nu, sig, omega, precision=prec).to_scipy()
    if prec == "fp64_exact":
        A.sort_indices(); ref.sort_indices()
        assert np.array_equal(A.indptr, ref.indptr) and np.array_equal(A.indices, ref.indices)
        assert np.array_equal(A.data.real, ref.data.real)
        assert np.array_equal(A.data.imag, ref.data.imag)
    else:
        d = A - ref
        assert abs(d.real).max() / abs(ref.real).max() < TOL[prec]
        assert abs(d.imag).max() / abs(ref.imag).max() < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_local"])
def test_csc_jacobian(core20, prec, driver):
    _one_exact_driver(prec, driver)
    mesh, M, dh, space, asm, mi = core20
    lib = {1: create_linear(1.0), 2: marrocco()}
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    n = mi["ndof"]
    rng = np.random.default_rng(8)
    a = rng.normal(size=n) * 0.004
    tags = mesh.RegionTags.astype(np.float64)
    I, J, V, Ri, Rv = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, a, qw, curl,
                                 packed)
    Jref = O.sparse_scipy(I, J, V, n, n)
    c, r, v = O.sparse(Ri, np.ones_like(Ri), Rv, n, 1, drop_zeros=False)
    Rref = np.zeros(n)
    Rref[r] = v
    Jg, Rg = asm.assembleJacobian(space, a, lib, calcJ=True, precision=prec)
    Rg = Rg.cpu().numpy()
    if prec == "fp64_exact":
        assert csc_equal_bits(Jg.to_scipy(), Jref)
        assert np.array_equal(Rg.view(np.uint64), Rref.view(np.uint64))
    else:
        assert abs(Jg.to_scipy() - Jref).max() / abs(Jref).max() < TOL[prec]
        assert rel_err(Rg, Rref) < TOL[prec]
    _, R0 = asm.assembleJacobian(space, a, lib, calcJ=False, precision=prec)
    if prec == "fp64_exact":
        assert np.array_equal(R0.cpu().numpy(), Rg)
    else:
        # the residual-only pass is another kernel instantiation (no LDS tile):
        # same arithmetic, FMA contraction may round differently
        assert rel_err(R0.cpu().numpy(), Rg) < 1e-13


@pytest.mark.parametrize("prec", ["fp64", "fp32_local"])
def test_csc_newton_set(core20, prec, driver):
    """Fused K(nu(B)) + M_sigma + J + R against the oracle's separate kernels:
    J and R from the Jacobian kernel, M from the mass kernel, K from the
    curl-curl kernel with nu_e = eval_nu(|B_e|^2) at the element's B."""
    mesh, M, dh, space, asm, mi = core20
    lib = {1: create_linear(1.0), 2: marrocco()}
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    n = mi["ndof"]
    a = np.random.default_rng(18).normal(size=n) * 0.004
    tags = mesh.RegionTags.astype(np.float64)
    sig = np.where(mesh.RegionTags == 2, 2e6, 0.0)
    out = asm.assembleNewtonSet(space, a, lib, sigmaMap=sig, precision=prec)
    plan = asm.plan(space)
    I, J, V, Ri, Rv = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, a, qw, curl,
                                 packed)
    Jref = O.sparse_scipy(I, J, V, n, n)
    Rref = np.zeros(n)
    np.add.at(Rref, Ri.astype(np.int64) - 1, Rv)
    B = element_B(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], a)
    nu_e, _ = O.eval_nu((B * B).sum(axis=1), mesh.RegionTags, packed)
    Kref = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu_e, qw,
                                       curl), n, n)
    Mref = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sig, qm, val),
                          n, n)
    # the ring driver splits the 64-column tiles into those an M or D term
    # reaches (an element with sigma != 0 or dnu != 0 among the tile's
    # incidences) and the rest; this workload has both kinds
    _, dnu_e = O.eval_nu((B * B).sum(axis=1), mesh.RegionTags, packed)
    hot = (sig != 0) | (dnu_e != 0)
    dofs = mi["Dofs"].astype(np.int64)[hot] - 1  # Ne x 6, 0-based, -1 inactive
    hot_tiles = np.unique(dofs[dofs >= 0] // 64)
    ntiles = (n + 63) // 64
    assert 0 < hot_tiles.size < ntiles
    tol = TOL[prec]
    for name, ref in (("K", Kref), ("M", Mref), ("J", Jref)):
        A = csc_from_plan(plan, out[name])
        assert abs(A - ref).max() / abs(ref).max() < tol, name
    assert rel_err(out["R"].cpu().numpy(), Rref) < tol
    # tiles no M or D term reaches: M exactly 0 and J exactly K there
    colptr = plan.pattern()[0].cpu().numpy().astype(np.int64)
    cold = np.ones(n, dtype=bool)
    for t in hot_tiles:
        cold[64 * t: 64 * (t + 1)] = False
    sel = np.repeat(cold, np.diff(colptr))
    Mv, Jv, Kv = (out[k].cpu().numpy() for k in ("M", "J", "K"))
    assert sel.any() and not Mv[sel].any()
    assert np.array_equal(Jv[sel], Kv[sel])
    # subsets of outputs give the same values
    only = asm.assembleNewtonSet(space, a, lib, sigmaMap=sig, precision=prec, outputs=("J",))
    assert torch_equal_bits(only["J"], out["J"])
    with pytest.raises(Exception, match="FP64 and FP32_LOCAL"):
        asm.assembleNewtonSet(space, a, lib, precision="fp64_exact")


def _newton_refs(mesh, mi, lib, a, sig):
    """Oracle K(nu(B)), M_sigma, J, R for the fused Newton set."""
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    n = mi["ndof"]
    tags = mesh.RegionTags.astype(np.float64)
    I, J, V, Ri, Rv = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, a, qw, curl,
                                 packed)
    Rref = np.zeros(n)
    np.add.at(Rref, Ri.astype(np.int64) - 1, Rv)
    B = element_B(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], a)
    nu_e, _ = O.eval_nu((B * B).sum(axis=1), mesh.RegionTags, packed)
    return {"K": O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu_e, qw,
                                             curl), n, n),
            "M": O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sig, qm, val),
                                n, n),
            "J": O.sparse_scipy(I, J, V, n, n), "R": Rref}


@pytest.mark.parametrize("overlap", ["0", "1"])
@pytest.mark.parametrize("case", ["all_linear", "all_sigma", "tiny"])
def test_newton_set_tile_classes(gpu, case, overlap, monkeypatch):
    """The ring driver runs tiles that no M or D term reaches with a one-plane
    op and lists the others for a persistent pass, after it or beside it on a
    second stream (overlap).  Edge cases: no listed tile (linear,
    non-conducting), every tile listed (sigma everywhere), and a mesh of 2
    tiles (fewer than the 8 XCDs the persistent grid is split over).  Against
    the oracle and against the element-order driver, which runs every tile
    with the three-plane op."""
    monkeypatch.setenv("MPK_NEWTON_OVERLAP", overlap)
    from magpackage_amd.meshgen import tag_box_region
    if case == "tiny":
        mesh = kuhn_box(2, 2, 2, seed=7)
        tag_box_region(mesh, (0.0, 0.0, 0.0), (0.5, 1.0, 1.0), 2)
    else:
        mesh = core_box_mesh(10, seed=5)
    M_, dh, space, asm = _setup(mesh, gpu)
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    if case == "all_linear":
        lib = {1: create_linear(1.0), 2: create_linear(1000.0)}
        sig = np.zeros(mesh.num_elems)
    elif case == "all_sigma":
        lib = {1: create_linear(1.0), 2: marrocco()}
        sig = np.full(mesh.num_elems, 3e6)
    else:
        lib = {1: create_linear(1.0), 2: marrocco()}
        sig = np.where(mesh.RegionTags == 2, 2e6, 0.0)
    a = np.random.default_rng(21).normal(size=n) * 0.004
    plan = asm.plan(space)
    if case == "tiny":
        assert (n + 63) // 64 < 8 and plan.ring_order
    monkeypatch.setenv("MPK_CSC_DRIVER", "ring")
    ring = asm.assembleNewtonSet(space, a, lib, sigmaMap=sig, precision="fp64")
    monkeypatch.setenv("MPK_CSC_DRIVER", "element")
    elem = asm.assembleNewtonSet(space, a, lib, sigmaMap=sig, precision="fp64")
    refs = _newton_refs(mesh, mi, lib, a, sig)
    for name in ("K", "M", "J"):
        A = csc_from_plan(plan, ring[name])
        scale = max(abs(refs[name]).max(), 1e-300)
        assert abs(A - refs[name]).max() / scale < TOL["fp64"], name
        d = (ring[name] - elem[name]).abs().max().item()
        assert d <= 1e-13 * max(elem[name].abs().max().item(), 1e-300), name
    assert rel_err(ring["R"].cpu().numpy(), refs["R"]) < TOL["fp64"]
    if case == "all_linear":
        assert not ring["M"].cpu().numpy().any()
        assert np.array_equal(ring["J"].cpu().numpy(), ring["K"].cpu().numpy())


@pytest.mark.parametrize("overlap", ["0", "1"])
@pytest.mark.parametrize("prec", ["fp64", "fp32_local"])
def test_newton_set_classes_change_between_calls(gpu, prec, overlap, monkeypatch):
    """One plan, four calls whose listed tiles move: sigma on the left half,
    then the right half, then nowhere, then the left half again (linear
    materials, so sigma alone sets the classes).  The tile classes are
    per-call epochs in flags that are not cleared between calls; every output
    of every call against the oracle, so a class or an M value left over from
    an earlier call would show."""
    monkeypatch.setenv("MPK_NEWTON_OVERLAP", overlap)
    monkeypatch.setenv("MPK_CSC_DRIVER", "ring")
    mesh = core_box_mesh(10, seed=5)
    M_, dh, space, asm = _setup(mesh, gpu)
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    lib = {1: create_linear(1.0), 2: create_linear(1000.0)}
    from magpackage_amd.meshgen import centroid_axis
    cx = centroid_axis(mesh, 0)
    left = np.where(cx < 0.5, 2e6, 0.0)
    right = np.where(cx >= 0.5, 3e6, 0.0)
    none = np.zeros(mesh.num_elems)
    plan = asm.plan(space)
    assert plan.ring_order
    tol = TOL[prec]
    for k, sig in enumerate((left, right, none, left)):
        a = np.random.default_rng(30 + k).normal(size=n) * 0.004
        out = asm.assembleNewtonSet(space, a, lib, sigmaMap=sig, precision=prec)
        refs = _newton_refs(mesh, mi, lib, a, sig)
        for name in ("K", "J"):
            A = csc_from_plan(plan, out[name])
            assert abs(A - refs[name]).max() / abs(refs[name]).max() < tol, (k, name)
        A = csc_from_plan(plan, out["M"])
        if refs["M"].nnz:
            assert abs(A - refs["M"]).max() / abs(refs["M"]).max() < tol, (k, "M")
        else:
            assert not out["M"].cpu().numpy().any(), k
        assert rel_err(out["R"].cpu().numpy(), refs["R"]) < tol, (k, "R")


def torch_equal_bits(a, b):
    import torch
    return torch.equal(a.view(torch.int64), b.view(torch.int64))


# --------------------------------------- size-independent properties -------
def test_large_mesh_properties(gpu):
    """A 60^3 Kuhn box (1.3M tets): K G = 0, symmetry, exact-vs-fast agreement."""
    import torch
    mesh = kuhn_box(60, 60, 60, seed=9)
    M, dh, space, asm = _setup(mesh, gpu)
    nu = np.random.default_rng(10).uniform(1e5, 8e5, mesh.num_elems)
    K = asm.assembleStiffness(space, nu, precision="fp64", drop_zeros=False)
    Ks = K.to_scipy()
    assert K.nnz == asm.plan(space).nnz
    G = gradient_matrix(M.Edges.cpu().numpy(), mesh.num_nodes)
    phi = np.random.default_rng(11).normal(size=mesh.num_nodes)
    assert np.abs(Ks @ (G @ phi)).max() < 1e-9 * abs(Ks).max() * np.abs(G @ phi).max()
    assert abs(Ks - Ks.T).max() < 1e-12 * abs(Ks).max()
    Ke = asm.assembleStiffness(space, nu, precision="fp64_exact", drop_zeros=False).to_scipy()
    assert abs(Ks - Ke).max() < 1e-12 * abs(Ke).max()
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_local"])
def test_restricted_dofs_csc(gpu, prec, driver):
    """DOFs on the iron region only (DofHandler.distributeDofs with target tags):
    air elements carry inactive (0) DOFs, the region's boundary edges have
    chains instead of rings.  K and M_sigma through the plan against the
    oracle's MEX triplets with Assembler's valid = I>0 & J>0 filter."""
    _one_exact_driver(prec, driver)
    mesh = core_box_mesh(8, seed=14)
    M_, dh, space, asm = _setup(mesh, gpu, active_tags=[2])
    mi = O.mesh_inputs(mesh, active=mesh.RegionTags == 2)
    n = mi["ndof"]
    assert dh.NumGlobalDofs == n and (mi["Dofs"] == 0).any()
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    nu = core_box_nu(mesh)
    sig = np.where(mesh.RegionTags == 2, 2e6, 0.0)
    Kref = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl),
                          n, n)
    Mref = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sig, qm, val), n, n)
    K = asm.assembleStiffness(space, nu, precision=prec).to_scipy()
    Mw = asm.assembleMassWeighted(space, sig, precision=prec).to_scipy()
    if prec == "fp64_exact":
        assert csc_equal_bits(K, Kref) and csc_equal_bits(Mw, Mref)
    else:
        assert abs(K - Kref).max() / abs(Kref).max() < TOL[prec]
        assert abs(Mw - Mref).max() / abs(Mref).max() < TOL[prec]


def test_empty_and_single_element(gpu):
    """No elements: every MEX drop-in returns empty outputs; sparse() of no
    triplets is an all-zero CSC.  One tetrahedron: the CSC path (a single tile
    of 6 one-incidence chains) against the oracle."""
    from magpackage_amd import mex
    from magpackage_amd.hbfem import AFT
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    P = np.zeros((3, 4))
    T0 = np.zeros((4, 0))
    D0 = np.zeros((6, 0))
    z = np.zeros(0)
    for out in (mex.assemble_curl_curl_kernel_mex(P, T0, D0, D0, z, qw, curl),
                mex.assemble_mass_kernel_mex(P, T0, D0, D0, z, qm, val),
                mex.assemble_source_kernel_mex(P, T0, D0, D0, np.zeros((3, 0)), qm, val),
                mex.assemble_winding_kernel_mex(P, T0, D0, D0, z, 1, z, np.ones(3), 1.0, qm, val)):
        assert all(np.size(o) == 0 for o in out)
    colptr, rowind, vals = mex.sparse_gpu(z, z, z, 5, 4)
    assert np.array_equal(colptr, np.zeros(5)) and rowind.size == 0 and vals.size == 0
    # one tetrahedron
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    Pt = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.0, 0.1], [0.1, 0.3, 1.2]])
    Tt = np.array([[1, 2, 3, 4]])
    mesh = Mesh(Pt, Tt, np.array([1]))
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)

    class _One:
        P, T, RegionTags = Pt, Tt, np.array([1])
    mi = O.mesh_inputs(_One)
    nu = np.array([1234.5])
    Kref = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl),
                          6, 6)
    for prec in ("fp64_exact", "fp64"):
        K = asm.assembleStiffness(space, nu, precision=prec).to_scipy()
        if prec == "fp64_exact":
            assert csc_equal_bits(K, Kref)
        else:
            assert abs(K - Kref).max() / abs(Kref).max() < 1e-12
