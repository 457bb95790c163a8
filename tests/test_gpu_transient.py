"""GPU transient Newton driver against its CPU restatement (oracle kernels +
scipy direct solves): ten strongly saturating backward-Euler steps with the
GPU arm's linear systems solved to 1e-12 (a direct solve's accuracy), the
same Newton iteration counts, and B within 1e-6 of the restatement's."""
import numpy as np
import pytest

from oracle import oracle as O
from magpackage_amd.boundary import outer_boundary_edge_mask
from magpackage_amd.material import create_linear, marrocco
from magpackage_amd.meshgen import ecore_box
from tests.helpers import stiffness_tables
from tests.transient_ref import transient_cpu

pytestmark = pytest.mark.gpu


def test_transient_matches_cpu_restatement(gpu):
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.transient import TransientSolver
    ec = ecore_box(10, 10, 8, seed=3)
    m = ec.mesh
    lib = {1: create_linear(1.0), 2: marrocco(), 3: create_linear(1.0)}
    sigma = np.where(m.RegionTags == 2, 2e6, 0.0)
    amp = 3e7
    steps = [1e-3] * 10
    src = lambda t: ec.js_dir * amp * np.sin(2 * np.pi * 50 * t)
    edges, _, _ = O.generate_edges(m.T)
    fixed_cpu = outer_boundary_edge_mask(m.T, edges)
    ref, ref_iters = transient_cpu(m, lib, sigma, src, steps, fixed_cpu)

    M = Mesh.from_kuhn(m)
    dh = DofHandler(M)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(M, dh)
    fixed = find_outer_boundary_dofs(M, dh, space)
    assert np.array_equal(fixed, fixed_cpu)
    solver = TransientSolver(asm)
    # every Newton step's PCG to 1e-12, as the restatement's direct solves
    solver.NewtonLinearTolerance = 1e-12
    js = torch.from_numpy(ec.js_dir).cuda()
    sols, infos = solver.solve(space, lib, {2: 2e6}, lambda t: js * (amp * np.sin(2 * np.pi * 50 * t)),
                               steps, fixed)
    mi = O.mesh_inputs(m)
    _, curl, _ = stiffness_tables()
    for k, (x, xr, info) in enumerate(zip(sols, ref, infos)):
        print(f"step {k + 1}: gpu {info.newton_iters} Newton / {info.pcg_iters} PCG, "
              f"cpu {ref_iters[k]} Newton, rel res {info.rel_residual:.2e}", flush=True)
        assert info.converged
        assert info.newton_iters == ref_iters[k]
        Bg = asm.computeElementB(x, space).cpu().numpy()
        Br = O.element_b(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], xr, curl)
        err = np.abs(Bg - Br).max() / np.abs(Br).max()
        print(f"    B error {err:.1e}", flush=True)
        assert err <= 1e-6, (k, err)
        assert np.all(x.cpu().numpy()[fixed] == 0.0)


def test_divergence_free_source(gpu):
    """G^T F_df = 0 over the free nodes and F - F_df lies in M G (range of gradients)."""
    import scipy.sparse as sp
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.transient import divergence_free_source
    ec = ecore_box(10, 10, 8, seed=3)
    M = Mesh.from_kuhn(ec.mesh)
    dh = DofHandler(M)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(M, dh)
    fixed = find_outer_boundary_dofs(M, dh, space)
    F = asm.assembleSource(space, torch.from_numpy(ec.js_dir * 3e6).cuda())
    Fdf = divergence_free_source(asm, space, F, fixed).cpu().numpy()
    edges = M.Edges.cpu().numpy().astype(np.int64)
    E = edges.shape[0]
    G = sp.csr_matrix((np.tile([-1.0, 1.0], E), (np.repeat(np.arange(E), 2), (edges - 1).ravel())),
                      shape=(E, M.NumNodes))
    bn = np.unique((edges[fixed] - 1).ravel())
    keep = np.setdiff1d(np.arange(M.NumNodes), bn)
    f = np.where(fixed, 0.0, F.cpu().numpy())
    assert np.linalg.norm(G[:, keep].T @ f) > 0.1 * np.linalg.norm(f)  # the raw load is not
    assert np.linalg.norm(G[:, keep].T @ Fdf) < 1e-8 * np.linalg.norm(f)
    assert np.all(Fdf[fixed] == 0.0)


def test_resume_continues_the_run(gpu):
    """solve(..., resume=snapshot) after step k reproduces steps k+1.. of the
    uninterrupted run bit for bit: the same Newton and PCG counts and the
    same solutions."""
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.transient import TransientSolver
    ec = ecore_box(10, 10, 8, seed=3)
    M = Mesh.from_kuhn(ec.mesh)
    dh = DofHandler(M)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(M, dh)
    fixed = find_outer_boundary_dofs(M, dh, space)
    lib = {1: create_linear(1.0), 2: marrocco(), 3: create_linear(1.0)}
    js = torch.from_numpy(ec.js_dir).cuda()
    src = lambda t: js * (3e7 * np.sin(2 * np.pi * 50 * t))
    steps = [1e-3] * 6
    full = TransientSolver(asm)
    sols, infos = full.solve(space, lib, {2: 2e6}, src, steps, fixed)
    first = TransientSolver(asm)
    snap = {}
    first.solve(space, lib, {2: 2e6}, src, steps[:3], fixed,
                on_step=lambda ctx: snap.update(first.snapshot(ctx)) if ctx["step"] == 3 else None)
    second = TransientSolver(asm)
    sols2, infos2 = second.solve(space, lib, {2: 2e6}, src, steps[3:], fixed, resume=snap)
    assert [i.step for i in infos2] == [4, 5, 6]
    for a, b, ia, ib in zip(sols[3:], sols2, infos[3:], infos2):
        assert (ia.newton_iters, ia.pcg_iters) == (ib.newton_iters, ib.pcg_iters)
        assert torch.equal(a, b)
