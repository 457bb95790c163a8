"""GPU linear algebra of the transient driver against scipy on the same matrices."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from tests.helpers import core_box_mesh, core_box_nu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system(gpu):
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    km = core_box_mesh(8, seed=30)
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    K = asm.assembleStiffness(space, core_box_nu(km), drop_zeros=False)
    M = asm.assembleMass(space, drop_zeros=False)
    edges = mesh.Edges.cpu().numpy()
    fixed = O.boundary_edges(km.T, edges)
    return torch, asm, space, K, M, fixed


def test_spmv_and_combine(system):
    torch, asm, space, K, M, fixed = system
    from magpackage_amd import _lib as L
    plan = asm.plan(space)
    dev = K.values.device
    A = torch.empty_like(K.values)
    L.check(L.lib.mpk_csc_combine(plan.nnz, 1.0, L.ptr(K.values), 1e5, L.ptr(M.values), 0.0, None,
                                  0.0, None, L.ptr(A), L.stream_ptr()))
    As = K.to_scipy() + 1e5 * M.to_scipy()
    x = np.random.default_rng(0).normal(size=plan.n)
    y = torch.empty(plan.n, dtype=torch.float64, device=dev)
    colptr32 = K.colptr.to(torch.int32)
    L.check(L.lib.mpk_spmv_sym(L.ptr(colptr32), L.ptr(K.rowind), L.ptr(A), plan.n,
                               L.ptr(torch.from_numpy(x).to(dev)), L.ptr(y), L.stream_ptr()))
    ref = As @ x
    assert np.max(np.abs(y.cpu().numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_pcg_dirichlet_against_scipy(system):
    torch, asm, space, K, M, fixed = system
    from magpackage_amd import _lib as L
    plan = asm.plan(space)
    dev = K.values.device
    reg = 1e-5 * float(np.mean(np.abs(K.to_scipy().diagonal())))
    A = torch.empty_like(K.values)
    L.check(L.lib.mpk_csc_combine(plan.nnz, 1.0, L.ptr(K.values), reg, L.ptr(M.values), 0.0,
                                  None, 0.0, None, L.ptr(A), L.stream_ptr()))
    n = plan.n
    b = np.random.default_rng(1).normal(size=n)
    # applyDirichlet: D_free A D_free + D_fixed, b(fixed) = 0
    As = (K.to_scipy() + reg * M.to_scipy()).tocsr()
    Df = sp.diags((~fixed).astype(float))
    Abc = (Df @ As @ Df + sp.diags(fixed.astype(float))).tocsc()
    bb = b.copy()
    bb[fixed] = 0
    ref = spla.spsolve(Abc, bb)
    ws = C.c_int64()
    L.check(L.lib.mpk_pcg_work_size(n, C.byref(ws)))
    work = torch.empty(ws.value, dtype=torch.float64, device=dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    fx = torch.from_numpy(fixed.astype(np.uint8)).to(dev)
    it, rr = C.c_int(), C.c_double()
    colptr32 = K.colptr.to(torch.int32)
    L.check(L.lib.mpk_pcg_solve(L.ptr(colptr32), L.ptr(K.rowind), L.ptr(A), n, L.ptr(fx),
                                L.ptr(torch.from_numpy(b).to(dev)), L.ptr(x), L.ptr(work), 1e-11,
                                20000, 25, C.byref(it), C.byref(rr), L.stream_ptr()))
    assert rr.value <= 1e-11
    xs = x.cpu().numpy()
    assert np.all(xs[fixed] == 0)
    assert np.max(np.abs(xs - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_pcg_aux_space_against_scipy(system):
    """Jacobi + gradient-space correction: same solution as a direct solve, and
    far fewer iterations than Jacobi alone on a weakly regularised curl-curl
    system whose right-hand side has a gradient part."""
    torch, asm, space, K, M, fixed = system
    from magpackage_amd import _lib as L
    from magpackage_amd.transient import build_aux_space
    plan = asm.plan(space)
    dev = K.values.device
    n = plan.n
    reg = 1e-5 * float(np.mean(np.abs(K.to_scipy().diagonal())))
    A = torch.empty_like(K.values)
    L.check(L.lib.mpk_csc_combine(plan.nnz, 1.0, L.ptr(K.values), reg, L.ptr(M.values), 0.0,
                                  None, 0.0, None, L.ptr(A), L.stream_ptr()))
    Cv = M.values * reg
    colptr32 = K.colptr.to(torch.int32)
    aux, keep = build_aux_space(asm, space.toString(), fixed, Cv, colptr32, K.rowind)
    b = np.random.default_rng(2).normal(size=n)
    As = (K.to_scipy() + reg * M.to_scipy()).tocsr()
    Df = sp.diags((~fixed).astype(float))
    Abc = (Df @ As @ Df + sp.diags(fixed.astype(float))).tocsc()
    bb = b.copy()
    bb[fixed] = 0
    ref = spla.spsolve(Abc, bb)
    fx = torch.from_numpy(fixed.astype(np.uint8)).to(dev)
    bd = torch.from_numpy(b).to(dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    it, rr = C.c_int(), C.c_double()
    L.check(L.lib.mpk_pcg_aux_solve(L.ptr(colptr32), L.ptr(K.rowind), L.ptr(A), n, L.ptr(fx),
                                    C.byref(aux), L.ptr(bd), L.ptr(x), L.ptr(keep["work"]), 1e-11,
                                    20000, 5, C.byref(it), C.byref(rr), L.stream_ptr()))
    assert rr.value <= 1e-11
    xs = x.cpu().numpy()
    assert np.all(xs[fixed] == 0)
    assert np.abs(xs - ref).max() <= 1e-7 * np.abs(ref).max()
    # Jacobi alone on the same system, same tolerance
    ws = C.c_int64()
    L.check(L.lib.mpk_pcg_work_size(n, C.byref(ws)))
    work = torch.empty(ws.value, dtype=torch.float64, device=dev)
    it_j, rr_j = C.c_int(), C.c_double()
    L.check(L.lib.mpk_pcg_solve(L.ptr(colptr32), L.ptr(K.rowind), L.ptr(A), n, L.ptr(fx),
                                L.ptr(bd), L.ptr(x), L.ptr(work), 1e-11, 20000, 5,
                                C.byref(it_j), C.byref(rr_j), L.stream_ptr()))
    print(f"aux-space PCG {it.value} iterations, Jacobi PCG {it_j.value}", flush=True)
    assert it.value * 3 < it_j.value
