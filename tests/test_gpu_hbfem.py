"""Harmonic-balance solvers (magpackage_amd.hbfem.HBFEMSolver,
HBFEMCoupledSolver) on the GPU:
- HBFEMSolver with linear materials: the converged harmonics solve
  (K + K_add) X_h = F_h, against a sparse LU solve of that system;
- HBFEMCoupledSolver against tests/hbfem_ref.py, the CPU restatement of the
  reference's solver with its monolithic field-circuit Jacobian solved by
  sparse LU, on a small E-core box driven into saturation (3rd and 5th
  harmonic currents), with the same Newton and line-search path."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle as O
from tests.hbfem_ref import hbfem_coupled_cpu
from tests.helpers import mass_tables, stiffness_tables

pytestmark = pytest.mark.gpu

# tutorials/run_ecore_hbfem.m:35-40
B_DATA = [0, 0.1, 0.2, 0.3001, 0.4, 0.5001, 0.6001, 0.7, 0.8, 0.9001, 1, 1.1001, 1.2001, 1.3002,
          1.4, 1.4999, 1.5999, 1.6991, 1.7987, 1.8978, 1.98, 1.9925, 2.118164]
H_DATA = [0, 7.86, 13.99, 19.51, 24.58, 29.42, 34.07, 38.64, 43.12, 47.59, 52.1, 56.77, 61.67,
          67.19, 74.03, 83.86, 1.00e2, 1.38e2, 2.95e2, 1.01e3, 2500, 12500, 112500]


def _box(gpu, eps_rel):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.meshgen import ecore_box
    em = ecore_box(6, 6, 4)
    km = em.mesh
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    K = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"],
                                    np.full(km.T.shape[0], 1 / (4e-7 * np.pi)), qw, curl), n, n)
    M = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], None, qm, val), n, n)
    # MatrixK_Add as run_ecore_hbfem.m:73-79 builds it (ratio eps_rel of the 1-norms)
    eps = spla.norm(K, 1) * eps_rel / spla.norm(M, 1)
    fixed = O.boundary_edges(km.T, mi["edges"])
    emap = np.zeros(mi["edges"].shape[0], dtype=np.int64)
    emap[mi["t2e"].ravel() - 1] = mi["cell"].ravel()
    fx = np.zeros(n, dtype=bool)
    fd = emap[fixed]
    fx[fd[fd > 0] - 1] = True
    return em, km, asm, space, (eps * M).tocsc(), fx, mi


def test_hbfem_linear_converges_to_the_linear_solve(gpu):
    """Linear materials, harmonics [1, 3]: the block Jacobian is 2K per AC
    harmonic (the DC-reluctivity form), so Newton contracts by 1/2 per step;
    at Tolerance 1e-9 the harmonics equal (K + K_add)^-1 F_h."""
    from magpackage_amd.hbfem import AFT, HBFEMSolver
    from magpackage_amd.material import create_linear
    em, km, asm, space, Kadd, fx, mi = _box(gpu, 1e-4)
    lib = {1: create_linear(1.0), 2: create_linear(300.0), 3: create_linear(1.0)}
    aft = AFT([1, 3], None, 50)
    src1 = np.where((km.RegionTags == 3)[:, None], em.js_dir * 2e5, 0.0)
    src3 = np.where((km.RegionTags == 3)[:, None], em.js_dir * -4e4, 0.0)
    s = HBFEMSolver(asm, aft)
    s.MatrixK_Add = Kadd
    s.Tolerance = 1e-9
    X, info = s.solve(space, lib, [src1, src3], fx)
    n = mi["ndof"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    nu = np.where(km.RegionTags == 2, 1 / (300 * 4e-7 * np.pi), 1 / (4e-7 * np.pi))
    K = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl), n, n)
    A = (K + Kadd).tocsr()
    f = (~fx).astype(float)
    import scipy.sparse as sp
    Ab = (sp.diags(f) @ A @ sp.diags(f) + sp.diags(fx.astype(float))).tocsc()
    Xg = X.cpu().numpy()
    print(f"linear HBFEM: {info['Iterations']} Newton, {info['LinearIterations']} CG iterations",
          flush=True)
    assert info["Converged"]
    for h, src in enumerate((src1, src3)):
        I, V = O.source(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], src, qm, val)
        F = np.zeros(n)
        np.add.at(F, I.astype(int) - 1, V)
        F[fx] = 0.0
        ref = spla.spsolve(Ab, F)
        assert np.abs(Xg[:, h] - ref).max() <= 1e-7 * np.abs(ref).max()


def test_hbfem_coupled_matches_restatement(gpu):
    """Saturating E-core box, harmonics [1, 3, 5], 1000 V fundamental through
    10 ohm, 100 turns: A harmonics, current phasors and the Newton path equal
    the monolithic CPU solve."""
    from magpackage_amd.circuit import Winding
    from magpackage_amd.hbfem import AFT, HBFEMCoupledSolver
    from magpackage_amd.material import create_linear, create_nonlinear
    em, km, asm, space, Kadd, fx, mi = _box(gpu, 1e-2)
    lib = {1: create_linear(1.0), 2: create_nonlinear(B_DATA, H_DATA), 3: create_linear(1.0)}
    aft = AFT([1, 3, 5], None, 50)
    V = [-1000j, 0, 0]
    A_ref, I_ref, it_ref, hist_ref = hbfem_coupled_cpu(
        km, lib, aft, dict(region=3, turns=100, area=4.8e-4, dir=em.js_dir), 10.0, V, fx, Kadd,
        tol=1e-6)
    w = Winding("Primary", 3, 100, 10.0, 4.8e-4)
    w.setDirectionField(em.js_dir.T)
    s = HBFEMCoupledSolver(asm, aft, w, 10.0)
    s.MatrixK_Add = Kadd
    s.Tolerance = 1e-6
    A, I, info = s.solve(space, lib, V, fx)
    print(f"coupled HBFEM: {info['Iterations']} Newton (ref {it_ref}), I {np.round(I, 6)} "
          f"(ref {np.round(I_ref, 6)}), {info['LinearIterations']} CG iterations", flush=True)
    assert info["Converged"] and info["Iterations"] == it_ref
    assert np.allclose(info["Residuals"], hist_ref, rtol=1e-6, atol=0)
    assert np.abs(I - I_ref).max() <= 1e-6 * np.abs(I_ref).max()
    # the saturating core draws odd harmonics (exactly zero for a linear core)
    assert np.abs(I_ref[1:]).min() > 1e-5 * np.abs(I_ref[0])
    Ag = A.cpu().numpy()
    assert np.abs(Ag - A_ref).max() <= 1e-6 * np.abs(A_ref).max()


def test_ecore_hbfem_tutorial_matches_restatement(gpu):
    """tutorials/run_ecore_hbfem.m on the reference's E-core mesh (harmonics
    1, 3, 5, 7; 300 turns through 100 ohm; -76j V; MatrixK_Add = 1e-7 |K|_1 /
    |M|_1 x M; Tolerance 1e-3): the Newton path and the current phasors equal
    those of the CPU restatement (fixture tests/golden/ecore_hbfem_run.npz,
    made by tests/golden/make_hbfem_fixture.py)."""
    from pathlib import Path
    from types import SimpleNamespace
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.circuit import Winding
    from magpackage_amd.coil import CoilGeometryUtils as CG
    from magpackage_amd.hbfem import AFT, HBFEMCoupledSolver
    from magpackage_amd.material import create_linear, create_nonlinear
    golden = Path(__file__).resolve().parent / "golden"
    ref = np.load(golden / "ecore_hbfem_run.npz")
    d = np.load(golden / "ecore_mesh.npz")
    tags = d["tags"].copy()
    tags[np.isin(tags, [5, 6, 8, 9])] = 50
    tags[np.isin(tags, [3, 4, 7, 10])] = 60
    mesh = Mesh(d["P"], d["T"], tags)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    coil = SimpleNamespace(P=d["P"], T=d["T"], RegionTags=tags, NumElements=len(tags))
    c, _, area, ax = CG.autoDetectCircular(coil, 50)
    w = Winding("Primary", 50, 300, 100, area)
    w.setDirectionField(-CG.computeCircularDirection(coil, 50, c, ax))
    lib = {1: create_linear(1.0), 2: create_nonlinear(B_DATA, H_DATA), 50: create_linear(1.0),
           60: create_linear(1.0)}
    K = asm.assembleStiffness(space, np.full(mesh.NumElements, 1 / (4e-7 * np.pi))).to_scipy()
    M = asm.assembleMass(space).to_scipy()
    eps = spla.norm(K, 1) * 1e-7 / spla.norm(M, 1)
    s = HBFEMCoupledSolver(asm, AFT([1, 3, 5, 7], None, 50), w, 100.0)
    s.Tolerance = 1e-3
    kadd = asm.assembleMass(space, drop_zeros=False)
    kadd.values = kadd.values * eps
    s.MatrixK_Add = kadd
    A, I, info = s.solve(space, lib, [-76j, 0, 0, 0], find_outer_boundary_dofs(mesh, dh, space))
    print(f"E-core HBFEM: {info['Iterations']} Newton (ref {int(ref['iterations'])}), |I| "
          f"{np.abs(I)} (ref {np.abs(ref['current'])}), {info['LinearIterations']} CG iterations",
          flush=True)
    assert info["Converged"] and info["Iterations"] == int(ref["iterations"])
    assert np.allclose(info["Residuals"], ref["residuals"], rtol=1e-5, atol=0)
    assert np.abs(I - ref["current"]).max() <= 1e-5 * np.abs(ref["current"]).max()
    assert np.allclose(np.linalg.norm(A.cpu().numpy(), axis=0), ref["a_norms"], rtol=1e-5)
