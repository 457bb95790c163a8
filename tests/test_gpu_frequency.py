"""Time-harmonic path on the GPU: the complex symmetric COCG solve against a
direct scipy solve of the same matrices, the FrequencySolver (linear and
nonlinear), and the TEAM Problem 7 benchmark (the reference's own mesh and
tutorial set-up, tutorials/run_team7.m) against the published measurements."""
import ctypes as C
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from tests.helpers import core_box_mesh, core_box_nu

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def csystem(gpu):
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    km = core_box_mesh(8, seed=31)
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    K = asm.assembleStiffness(space, core_box_nu(km), drop_zeros=False)
    M = asm.assembleMass(space, drop_zeros=False)
    sigma = np.where(km.RegionTags == 2, 2e6, 0.0)
    Ms = asm.assembleMassWeighted(space, sigma, drop_zeros=False)
    fixed = O.boundary_edges(km.T, mesh.Edges.cpu().numpy())
    return torch, asm, space, K, M, Ms, fixed


def _cocg(torch, asm, space, R, I, fixed, b, aux_c=None, tol=1e-12, fill=None):
    from magpackage_amd import _lib as L
    from magpackage_amd.transient import build_aux_space
    plan = asm.plan(space)
    n = plan.n
    dev = R.device
    colptr, rowind = plan.pattern()
    colptr32 = colptr.to(torch.int32)
    fx = torch.from_numpy(fixed.astype(np.uint8)).to(dev)
    aux_p, keep, nv = None, {}, 0
    if aux_c is not None:
        aux, keep = build_aux_space(asm, space.toString(), fixed, aux_c, colptr32, rowind)
        aux_p, nv = C.byref(aux), aux.num_nodes
    ws = C.c_int64()
    L.check(L.lib.mpk_cocg_aux_work_size(n, nv, C.byref(ws)))
    work = torch.empty(ws.value, dtype=torch.float64, device=dev)
    if fill is not None:
        work.fill_(fill)
    bd = torch.from_numpy(b.astype(np.complex128)).to(dev)
    x = torch.empty(n, dtype=torch.complex128, device=dev)
    it, rr = C.c_int(), C.c_double()
    L.check(L.lib.mpk_cocg_aux_solve(L.ptr(colptr32), L.ptr(rowind), L.ptr(R), L.ptr(I), n,
                                     L.ptr(fx), aux_p, L.ptr(bd), L.ptr(x), L.ptr(work), tol,
                                     40000, 10, C.byref(it), C.byref(rr), L.stream_ptr()))
    return x.cpu().numpy(), it.value, rr.value


def _direct(R, I, fixed, b):
    A = (R + 1j * I).tocsr()
    Df = sp.diags((~fixed).astype(float))
    Abc = (Df @ A @ Df + sp.diags(fixed.astype(float))).tocsc()
    bb = b.astype(np.complex128).copy()
    bb[fixed] = 0
    return spla.spsolve(Abc, bb)


@pytest.mark.parametrize("omega", [2 * np.pi * 50, 2 * np.pi * 5000])
def test_cocg_against_direct_solve(csystem, omega):
    """A = K + eps M + 1j omega M_sigma with Dirichlet edges: the COCG solution
    (Jacobi and auxiliary-space preconditioners) equals a sparse LU solve;
    the gradient-space correction cuts the iteration count; runs repeat
    bitwise."""
    torch, asm, space, K, M, Ms, fixed = csystem
    eps = 1e-6 * float(np.abs(K.to_scipy().diagonal()).max())
    R = K.values + eps * M.values
    I = omega * Ms.values
    rng = np.random.default_rng(5)
    n = asm.plan(space).n
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    ref = _direct(K.to_scipy() + eps * M.to_scipy(), omega * Ms.to_scipy(), fixed, b)
    x_aux, it_aux, rr_aux = _cocg(torch, asm, space, R, I, fixed, b,
                                  aux_c=eps * M.values + I)
    x_jac, it_jac, rr_jac = _cocg(torch, asm, space, R, I, fixed, b)
    print(f"omega {omega:.0f}: COCG aux {it_aux} it ({rr_aux:.1e}), Jacobi {it_jac} it "
          f"({rr_jac:.1e})", flush=True)
    for x, rr in ((x_aux, rr_aux), (x_jac, rr_jac)):
        assert rr <= 1e-12
        assert np.all(x[fixed] == 0)
        assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()
    assert it_aux * 2 < it_jac
    x2, it2, _ = _cocg(torch, asm, space, R, I, fixed, b, aux_c=eps * M.values + I)
    assert it2 == it_aux and np.array_equal(x2.view(np.float64), x_aux.view(np.float64))


def test_cocg_workspace_contents_do_not_matter(csystem):
    """Regression: the workspace is uninitialised memory; a NaN-filled one
    gives the same iterations and the same bits as a zero-filled one."""
    torch, asm, space, K, M, Ms, fixed = csystem
    eps = 1e-6 * float(np.abs(K.to_scipy().diagonal()).max())
    R = K.values + eps * M.values
    I = 314.0 * Ms.values
    n = asm.plan(space).n
    b = np.random.default_rng(8).normal(size=n) + 0.5j
    x0, it0, r0 = _cocg(torch, asm, space, R, I, fixed, b, aux_c=eps * M.values + I, fill=0.0)
    x1, it1, r1 = _cocg(torch, asm, space, R, I, fixed, b, aux_c=eps * M.values + I,
                        fill=float("nan"))
    assert it0 > 10 and r0 <= 1e-12
    assert it1 == it0 and np.array_equal(x1.view(np.float64), x0.view(np.float64))


def test_cocg_real_system_matches_pcg(csystem):
    """With I = 0 and a real right-hand side COCG is preconditioned CG: same
    iterates as mpk_pcg_aux_solve up to rounding."""
    torch, asm, space, K, M, Ms, fixed = csystem
    from magpackage_amd import _lib as L
    from magpackage_amd.transient import build_aux_space
    eps = 1e-5 * float(np.abs(K.to_scipy().diagonal()).mean())
    R = K.values + eps * M.values
    I = torch.zeros_like(R)
    n = asm.plan(space).n
    b = np.random.default_rng(6).normal(size=n)
    x_c, it_c, _ = _cocg(torch, asm, space, R, I, fixed, b, aux_c=eps * M.values, tol=1e-10)
    plan = asm.plan(space)
    colptr, rowind = plan.pattern()
    colptr32 = colptr.to(torch.int32)
    aux, keep = build_aux_space(asm, space.toString(), fixed, eps * M.values, colptr32, rowind)
    dev = R.device
    x = torch.empty(n, dtype=torch.float64, device=dev)
    it, rr = C.c_int(), C.c_double()
    L.check(L.lib.mpk_pcg_aux_solve(L.ptr(colptr32), L.ptr(rowind), L.ptr(R), n,
                                    L.ptr(torch.from_numpy(fixed.astype(np.uint8)).to(dev)),
                                    C.byref(aux), L.ptr(torch.from_numpy(b).to(dev)), L.ptr(x),
                                    L.ptr(keep["work"]), 1e-10, 40000, 10, C.byref(it),
                                    C.byref(rr), L.stream_ptr()))
    assert abs(it.value - it_c) <= 10     # same convergence check interval
    assert np.abs(x_c.imag).max() == 0.0
    assert np.abs(x_c.real - x.cpu().numpy()).max() <= 1e-9 * np.abs(x_c.real).max()


def _oracle_ops(km, nu, sg, src):
    """K, unweighted M, M_sigma and F of a Kuhn mesh from the oracle's MEX
    restatement, through the oracle's sparse()."""
    from tests.helpers import mass_tables, stiffness_tables
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    args = (mi["P"], mi["T"], mi["Dofs"], mi["Signs"])
    K = O.sparse_scipy(*O.curl_curl(*args, nu, qw, curl), n, n)
    Mr = O.sparse_scipy(*O.mass(*args, None, qm, val), n, n)
    Ms = O.sparse_scipy(*O.mass(*args, sg, qm, val), n, n)
    Js = np.zeros((km.T.shape[0], 3))
    for tag, v in src.items():
        Js[km.RegionTags == tag] = v
    I, V = O.source(*args, Js, qm, val)
    c, r, v = O.sparse(I, np.ones_like(I), V, n, 1, drop_zeros=False)
    F = np.zeros(n)
    F[r] = v
    return K, Mr, Ms, F


def _frequency_setup(km, sigma_tag2, mat2):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.material import create_linear
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    matLib = {1: create_linear(1.0), 2: mat2}
    sigma = {1: 0.0, 2: sigma_tag2}
    return mesh, dh, space, asm, matLib, sigma


def test_frequency_solver_linear_against_direct(gpu):
    """FrequencySolver on a conducting core: A equals a direct solve of
    K + 1j w M_sigma + eps M_reg (FrequencySolver.m:127-139) assembled by the
    oracle's MEX restatement; without a scalar space the conductivity is
    ignored, as in the reference."""
    from magpackage_amd.assembler import FunctionSpace
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.frequency import FrequencySolver
    from magpackage_amd.material import create_linear
    km = core_box_mesh(6, seed=32)
    mesh, dh, space, asm, matLib, sigma = _frequency_setup(km, 1e6, create_linear(200.0))
    fixed = find_outer_boundary_dofs(mesh, dh, space)
    src = {1: [0.0, 0.0, 0.0], 2: [0.0, 0.0, 1e6]}
    f = 200.0
    fs = FrequencySolver(asm, f)
    A, V = fs.solve(space, FunctionSpace("Lagrange", 1), matLib, sigma, src, fixed)
    # oracle assembly of the same operator
    mu0 = 4e-7 * np.pi
    nu = np.where(km.RegionTags == 2, 1 / (200.0 * mu0), 1 / mu0)
    sg = np.where(km.RegionTags == 2, 1e6, 0.0)
    Kr, Mr, Msr, Fr = _oracle_ops(km, nu, sg, src)
    eps = 1e-6 * np.abs(Kr.diagonal()).max()
    ref = _direct(Kr + eps * Mr, 2 * np.pi * f * Msr, fixed, Fr)
    a = A.cpu().numpy()
    print(f"linear solve: {fs.LinearIterations} COCG iterations", flush=True)
    assert fs.Iterations == 1
    assert np.abs(a - ref).max() <= 1e-7 * np.abs(ref).max()
    assert V is not None and V.numel() == np.unique(km.T[km.RegionTags == 2]).size
    # no scalar space: M_sigma is left out (FrequencySolver.m:63-77)
    A0, V0 = FrequencySolver(asm, f).solve(space, None, matLib, sigma, src, fixed)
    ref0 = _direct(Kr + eps * Mr, 0 * Msr, fixed, Fr)
    assert V0 is None
    assert np.abs(A0.cpu().numpy() - ref0).max() <= 1e-7 * np.abs(ref0).max()


def test_frequency_solver_nonlinear_fixed_point(gpu):
    """Nonlinear core: the relaxed outer loop converges (relative change below
    Tolerance) and its result is the fixed point: re-solving with nu(|B|^2)
    of the result reproduces it."""
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.frequency import FrequencySolver
    from magpackage_amd.material import marrocco
    km = core_box_mesh(6, seed=33)
    mesh, dh, space, asm, matLib, sigma = _frequency_setup(km, 0.0, marrocco())
    fixed = find_outer_boundary_dofs(mesh, dh, space)
    src = {1: [0.0, 0.0, 0.0], 2: [0.0, 0.0, 1e5]}
    fs = FrequencySolver(asm, 50.0)
    fs.Tolerance = 1e-7
    A, _ = fs.solve(space, None, matLib, sigma, src, fixed)
    print(f"nonlinear: {fs.Iterations} outer iterations", flush=True)
    assert 1 < fs.Iterations < fs.MaxIterations
    nu = fs._effective_nu(matLib, A, space.toString())
    assert np.ptp(nu[km.RegionTags == 2]) > 0   # the field reaches the nonlinear range
    Kr, Mr, _, Fr = _oracle_ops(km, nu, np.zeros(km.T.shape[0]), src)
    eps = 1e-6 * np.abs(Kr.diagonal()).max()
    again = _direct(Kr + eps * Mr, 0 * Mr, fixed, Fr)
    a = A.cpu().numpy()
    assert np.abs(a - again).max() <= 1e-5 * np.abs(again).max()


# TEAM Problem 7 measured Bz (Gauss) on lines A1-B1 and A2-B2 at 50 Hz, as
# tutorials/run_team7.m:133-143 tabulates them (x in mm; published benchmark).
TEAM7_X = np.array([0, 18, 36, 54, 72, 90, 108, 126, 144, 162, 180, 198, 216, 234, 252, 270, 288],
                   dtype=np.float64)
TEAM7_BZ = {
    1: (np.array([-4.9, -17.88, -22.13, -20.19, -15.67, 0.36, 43.64, 78.11, 71.55, 60.44, 53.91,
                  52.62, 53.81, 56.91, 59.24, 52.78, 27.61]),
        np.array([-1.16, 2.84, 4.15, 4, 3.07, 2.31, 1.89, 4.97, 12.61, 14.15, 13.04, 12.4, 12.05,
                  12.27, 12.66, 9.96, 2.36])),
    2: (np.array([-1.83, -8.5, -13.6, -15.21, -14.48, -5.62, 28.77, 60.34, 61.84, 56.64, 53.4,
                  52.36, 53.93, 56.82, 59.48, 52.08, 26.56]),
        np.array([-1.63, -0.6, -0.43, 0.11, 1.26, 3.4, 6.53, 10.25, 11.83, 11.83, 11.01, 10.58,
                  10.8, 10.54, 10.62, 9.03, 1.79])),
}
TEAM7_LINES = {1: (0.072, 0.034), 2: (0.144, 0.034)}   # (y, z) in metres


def test_team7_benchmark(gpu):
    """tutorials/run_team7.m end to end on the reference's mesh: coil detected
    from the mesh (2742 A-turns over the detected cross-section), aluminium
    plate sigma = 3.526e7 S/m at 50 Hz, outer-boundary Dirichlet edges.  The
    solve's residual is checked on the CPU against the assembled matrices,
    and Bz on lines A1-B1 and A2-B2 is compared with the measurements
    (Re Bz and -Im Bz, Gauss)."""
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.coil import CoilGeometryUtils as CG
    from magpackage_amd.frequency import FrequencySolver
    from magpackage_amd.material import create_linear
    from magpackage_amd.post import PostProcessor
    d = np.load(GOLDEN / "team7_mesh.npz")
    mesh = Mesh(d["P"] * 1e-3, d["T"], d["tags"])
    tags = mesh.RegionTags.copy()
    coil = SimpleNamespace(P=mesh.P, T=mesh.T, NumElements=mesh.NumElements,
                           RegionTags=np.where(np.isin(tags, [3, 4, 5, 6]), 9999, tags))
    center, _, _, Lx, Ly, area, ax = CG.autoDetectRoundedRect(coil, 9999)
    J_field = CG.computeRoundedRectDirection(coil, 9999, center, Lx, Ly, ax) * (2742 / area)
    matLib = {t: create_linear(1.0) for t in range(1, int(tags.max()) + 1)}
    sigma = {t: (3.526e7 if t == 2 else 0.0) for t in range(1, int(tags.max()) + 1)}
    dh = DofHandler(mesh)
    space_A = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space_A)
    asm = Assembler(mesh, dh)
    fs = FrequencySolver(asm, 50.0)
    fs.LinearTolerance = 1e-9
    fixed = find_outer_boundary_dofs(mesh, dh, space_A)
    torch.cuda.synchronize()
    A, V = fs.solve(space_A, FunctionSpace("Lagrange", 1), matLib, sigma, J_field, fixed)
    print(f"TEAM7: {asm.plan(space_A).n} DOFs, COCG {fs.LinearIterations} iterations, "
          f"rel. residual {fs.LinearResiduals}", flush=True)
    assert fs.LinearResiduals[-1] <= 1e-9
    # residual of the returned solution on the CPU, from the assembled matrices
    K = asm.assembleStiffness(space_A, np.full(mesh.NumElements, 1 / (4e-7 * np.pi)),
                              drop_zeros=False).to_scipy()
    Mr = asm.assembleMass(space_A, drop_zeros=False).to_scipy()
    Ms = asm.assembleMassWeighted(space_A, sigma, drop_zeros=False).to_scipy()
    F = asm.assembleSource(space_A, J_field).cpu().numpy()
    eps = 1e-6 * np.abs(K.diagonal()).max()
    S = (K + eps * Mr + 1j * 2 * np.pi * 50.0 * Ms).tocsr()
    a = A.cpu().numpy()
    free = ~fixed
    res = (S @ np.where(free, a, 0))[free] - F[free]
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(F[free])
    post = PostProcessor(asm)
    B = post.computeElementB(A)[:, :, 0].cpu().numpy()      # (3, Ne) complex
    errs = []
    for line, (y, z) in TEAM7_LINES.items():
        pts = np.stack([TEAM7_X * 1e-3, np.full_like(TEAM7_X, y), np.full_like(TEAM7_X, z)], 1)
        idx = post.locate(pts)
        assert (idx >= 0).all()
        bz = B[2, idx] * 1e4
        re_m, im_m = TEAM7_BZ[line]
        err_re = np.abs(bz.real - re_m)
        err_im = np.abs(-bz.imag - im_m)
        scale = np.abs(re_m).max()
        print(f"line A{line}-B{line}: Re Bz " + " ".join(f"{v:6.1f}" for v in bz.real), flush=True)
        print(f"   measured   " + " ".join(f"{v:6.1f}" for v in re_m), flush=True)
        print(f"  -Im Bz      " + " ".join(f"{v:6.1f}" for v in -bz.imag), flush=True)
        print(f"   measured   " + " ".join(f"{v:6.1f}" for v in im_m), flush=True)
        rms_re = np.sqrt(np.mean(err_re ** 2)) / scale
        rms_im = np.sqrt(np.mean(err_im ** 2)) / scale
        errs.append((rms_re, rms_im))
        print(f"  rms error / max|Bz|: Re {rms_re:.3f}, Im {rms_im:.3f}", flush=True)
    # 57,892 tets, B constant per element: 6-7% (Re) and 1.5-2% (Im) rms on both lines
    for rms_re, rms_im in errs:
        assert rms_re <= 0.10 and rms_im <= 0.05


def test_probe_solution_reproduces_edge_space_fields(gpu):
    """probeSolution interpolates with the edge basis: fields the lowest-order
    edge space holds exactly (A = c and A = w x x) come back exactly at any
    point, from DOFs set to their edge integrals (global orientation low ->
    high node id)."""
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.post import PostProcessor
    km = core_box_mesh(5, seed=34)
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    post = PostProcessor(Assembler(mesh, dh))
    E = mesh.Edges.cpu().numpy().astype(np.int64) - 1
    emap = dh.EntityMaps[space.toString()].cpu().numpy().astype(np.int64)
    lo, hi = km.P[E[:, 0]], km.P[E[:, 1]]
    c = np.array([0.3, -1.2, 0.7])
    w = np.array([0.5, 0.2, -0.9])
    rng = np.random.default_rng(3)
    pts = rng.uniform(0.05, 0.95, size=(6, 3))
    for field, edge_int in ((lambda x: c, lambda: (hi - lo) @ c),
                            (lambda x: np.cross(w, x),
                             lambda: np.einsum("ij,ij->i", np.cross(w, 0.5 * (lo + hi)), hi - lo))):
        a = np.zeros(dh.NumGlobalDofs)
        a[emap - 1] = edge_int()
        for p in pts:
            val, e = post.probeSolution(a, p)
            assert e >= 1
            assert np.abs(val - field(p)).max() <= 1e-12 * max(1.0, np.abs(field(p)).max())


def test_frequency_coupled_solver(gpu):
    """FrequencyCoupledSolver on a small E-core box: a conducting second leg
    (sigma 2e6; the coil around the first leg touches no conductor), a
    100-turn winding on 10 V / 50 Hz through 2 ohm + 1 mH.
    (a) The current and A equal the elimination of the A* system assembled by
        the oracle and solved by sparse LU.
    (b) With ProjectWinding (the winding vector's divergence-free part, in
        both), the current equals the reference's monolithic A-V system
        (FrequencyCoupledSolver.m:107-121: coupling and scalar-Laplacian
        blocks of the oracle's MEX restatements, first V DOF grounded) within
        the O(eps) difference of the two gauges.  With the raw winding vector
        (gradient part 30% of its norm here) the A-V system's flux linkage carries an
        eps-held gradient field, and the two currents differ."""
    from types import SimpleNamespace
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.circuit import Winding
    from magpackage_amd.fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
    from magpackage_amd.frequency import FrequencyCoupledSolver
    from magpackage_amd.material import create_linear
    from magpackage_amd.meshgen import centroid_axis, ecore_box
    from tests.helpers import mass_tables, stiffness_tables
    em = ecore_box(6, 6, 4)
    km = em.mesh
    leg2 = (km.RegionTags == 2) & (centroid_axis(km, 0) > 0.08) & (centroid_axis(km, 2) < 0.07)
    km.RegionTags[leg2] = 4
    lib = {1: create_linear(1.0), 2: create_linear(500.0), 3: create_linear(1.0),
           4: create_linear(500.0)}
    sigma = {1: 0.0, 2: 0.0, 3: 0.0, 4: 2e6}
    f, turns, area = 50.0, 100.0, 4.8e-4
    circ = SimpleNamespace(R=2.0, L=1e-3, V_source=10.0)
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    fixed = find_outer_boundary_dofs(mesh, dh, space)
    w = Winding("Primary", 3, turns, 2.0, area)
    w.setDirectionField(em.js_dir.T)
    fs = FrequencyCoupledSolver(asm, f)
    fs.LinearTolerance = 1e-12
    A, V, I, info = fs.solve(space, FunctionSpace("Lagrange", 1), lib, sigma, circ, w, fixed)
    # (a) the A* system on the CPU
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    args = (mi["P"], mi["T"], mi["Dofs"], mi["Signs"])
    mu0 = 4e-7 * np.pi
    nu = np.where(np.isin(km.RegionTags, [2, 4]), 1 / (500 * mu0), 1 / mu0)
    sg = np.where(km.RegionTags == 4, 2e6, 0.0)
    K = O.sparse_scipy(*O.curl_curl(*args, nu, qw, curl), n, n)
    Ms = O.sparse_scipy(*O.mass(*args, sg, qm, val), n, n)
    Mr = O.sparse_scipy(*O.mass(*args, None, qm, val), n, n)
    Iw, Vw = O.winding(*args, km.RegionTags.astype(float), 3, em.js_dir, np.zeros(3),
                       turns / area, qm, val)
    Wv = np.zeros(n)
    np.add.at(Wv, Iw[Iw > 0].astype(int) - 1, Vw[Iw > 0])
    om = 2 * np.pi * f
    eps = 1e-6 * np.abs(K.data).max()
    Z = circ.R + 1j * om * circ.L
    Wf = np.where(fixed, 0.0, Wv)
    v = _direct(K + eps * Mr, om * Ms, fixed, Wf)
    I_a = circ.V_source / (Z + 1j * om * (Wf @ v))
    print(f"coupled frequency: I {I:.8e} (A* elimination {I_a:.8e}), "
          f"{fs.LinearIterations} COCG iterations", flush=True)
    assert abs(I - I_a) <= 1e-8 * abs(I_a)
    assert np.abs(A.cpu().numpy() - v * I_a).max() <= 1e-7 * np.abs(v * I_a).max()
    # (b) the reference's monolithic A-V system, with the projected winding
    from magpackage_amd.frequency import divergence_free_part
    import torch
    fs2 = FrequencyCoupledSolver(asm, f)
    fs2.LinearTolerance = 1e-12
    fs2.ProjectWinding = True
    _, _, I_p, _ = fs2.solve(space, FunctionSpace("Lagrange", 1), lib, sigma, circ, w, fixed)
    Wv_raw = Wv.copy()
    Wv = divergence_free_part(asm, space, torch.from_numpy(Wv).to(gpu), fixed).cpu().numpy()
    cond_nodes = np.unique(km.T[km.RegionTags == 4])          # 1-based, node order
    vdof = np.zeros(km.P.shape[0] + 1, dtype=np.int64)
    vdof[cond_nodes] = n + np.arange(1, cond_nodes.size + 1)
    DN = vdof[km.T].astype(float)                          # 0: node outside the conductor
    q_pts, qw2 = get_quadrature_data("tet", 2)
    grad2 = matlab_flat(lagrange_tet_p1(q_pts)[1])
    q1, qw1 = get_quadrature_data("tet", 1)
    grad1 = matlab_flat(lagrange_tet_p1(q1)[1])
    ntot = n + cond_nodes.size + 1
    import scipy.sparse as sps
    Ic, Jc, Vc = O.coupling(*args, DN, sg, qw2, val, grad2)
    ok = (Ic > 0) & (Jc > 0)
    Cav = sps.csc_matrix((Vc[ok], (Ic[ok].astype(int) - 1, Jc[ok].astype(int) - 1)),
                         shape=(ntot, ntot))
    Il, Jl, Vl = O.scalar_laplacian(mi["P"], mi["T"], DN, sg, qw1, grad1)
    ok = (Il > 0) & (Jl > 0)
    Lvv = sps.csc_matrix((Vl[ok], (Il[ok].astype(int) - 1, Jl[ok].astype(int) - 1)),
                         shape=(ntot, ntot))
    pad = lambda m: sps.block_diag([m, sps.csc_matrix((ntot - n, ntot - n))]).tocsc()  # noqa
    S = pad(K + 1j * om * Ms + eps * Mr) + 1j * om * Cav + 1j * om * Cav.T + 1j * om * Lvv
    wi = np.nonzero(Wv)[0]
    Wm = sps.csc_matrix((-Wv[wi], (wi, np.full(wi.size, ntot - 1))), shape=(ntot, ntot))
    Wt = sps.csc_matrix((-1j * om * Wv[wi], (np.full(wi.size, ntot - 1), wi)), shape=(ntot, ntot))
    Zm = sps.csc_matrix(([-Z], ([ntot - 1], [ntot - 1])), shape=(ntot, ntot))
    Sys = (S + Wm + Wt + Zm).tocsc()
    rhs = np.zeros(ntot, dtype=complex)
    rhs[-1] = -circ.V_source
    fx = np.zeros(ntot, dtype=bool)
    fx[:n] = fixed
    fx[n] = True                                           # grounded first V DOF
    fr = (~fx).astype(float)
    Sb = (sps.diags(fr) @ Sys @ sps.diags(fr) + sps.diags(fx.astype(float))).tocsc()
    rhs[fx] = 0
    x = spla.spsolve(Sb, rhs)
    print(f"  projected winding: I {I_p:.8e}, reference A-V system {x[-1]:.8e}; raw winding "
          f"|G^T W| / |W| = {np.linalg.norm(Wv_raw - Wv) / np.linalg.norm(Wv_raw):.2f}",
          flush=True)
    assert abs(I_p - x[-1]) <= 1e-6 * abs(x[-1])
