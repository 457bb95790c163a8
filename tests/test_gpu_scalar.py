"""DC conduction (magpackage_amd.scalar, ScalarSolver) on the GPU:
- the reference's own test (tests/test_scalar_conduction.m): a copper bar, 0 V
  at one end and 0.1 V at the other. P1 elements reproduce the linear potential
  exactly, so V is 0.1 z / L at every node and the current V^T S V / 0.1 is
  sigma W^2 0.1 / L, whatever the (jittered) split;
- a two-conductivity box against a direct scipy solve of the oracle's
  assembled Laplacian, with the reference's Dirichlet handling."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _space(km, tags=None):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    sp = FunctionSpace("Lagrange", 1)
    dh.distributeDofs(sp, tags)
    return sp, Assembler(mesh, dh)


def test_copper_bar_resistance(gpu):
    import torch
    from magpackage_amd.meshgen import kuhn_box
    from magpackage_amd.scalar import ScalarSolver
    Lz, W, sigma = 1.0, 0.1, 5.8e7
    km = kuhn_box(3, 3, 20, lengths=(W, W, Lz))
    sp, asm = _space(km)
    z = km.P[:, 2]
    bottom = np.nonzero(np.abs(z) < 1e-6)[0]
    top = np.nonzero(np.abs(z - Lz) < 1e-6)[0]
    bc = np.concatenate([bottom, top]) + 1
    vals = np.concatenate([np.zeros(bottom.size), 0.1 * np.ones(top.size)])
    V, info = ScalarSolver(asm).solve(sp, sigma, bc, vals)
    v = V.cpu().numpy()
    S = asm.assembleScalarLaplacian(sp, sigma)
    Sv = torch.empty_like(V)
    from magpackage_amd import _lib as L
    L.check(L.lib.mpk_spmv_sym(L.ptr(S.colptr.to(torch.int32)), L.ptr(S.rowind), L.ptr(S.values),
                               S.shape[0], L.ptr(V), L.ptr(Sv), L.stream_ptr()))
    I_fem = float(V @ Sv) / 0.1
    I_theo = 0.1 / (Lz / (sigma * W * W))
    print(f"{info['Iterations']} CG iterations, rel {info['RelResidual']:.1e}; "
          f"I_fem {I_fem:.6f} A, I_theo {I_theo:.6f} A", flush=True)
    assert np.abs(v - 0.1 * z / Lz).max() <= 1e-10
    assert abs(I_fem - I_theo) <= 1e-9 * I_theo


def test_two_conductivities_against_direct_solve(gpu):
    from magpackage_amd.fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
    from magpackage_amd.scalar import ScalarSolver
    from tests.helpers import core_box_mesh
    km = core_box_mesh(8, seed=11)
    sp, asm = _space(km)
    sigma = {1: 1.0e6, 2: 3.0e7}
    x = km.P[:, 0]
    left = np.nonzero(x < 1e-9)[0]
    right = np.nonzero(x > 1 - 1e-9)[0]
    bc = np.concatenate([left, right]) + 1
    vals = np.concatenate([np.full(left.size, -0.2), np.full(right.size, 0.3)])
    V, info = ScalarSolver(asm).solve(sp, sigma, bc, vals)
    n = km.P.shape[0]
    q, w = get_quadrature_data("tet", 1)
    g = matlab_flat(lagrange_tet_p1(q)[1])
    sig_e = np.where(km.RegionTags == 2, 3.0e7, 1.0e6)
    S = O.sparse_scipy(*O.scalar_laplacian(km.P, km.T.astype(np.float64),
                                           km.T.astype(np.float64), sig_e, w, g), n, n).tocsc()
    fixed = np.zeros(n, dtype=bool)
    fixed[bc - 1] = True
    vf = np.zeros(n)
    vf[bc - 1] = vals
    F = -(S @ vf)
    free = ~fixed
    ref = vf.copy()
    ref[free] = spla.spsolve(S[free][:, free].tocsc(), F[free])
    v = V.cpu().numpy()
    err = np.abs(v - ref).max() / np.abs(ref).max()
    print(f"{info['Iterations']} CG iterations; max rel diff {err:.2e}", flush=True)
    assert err <= 1e-9
