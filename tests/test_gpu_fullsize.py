"""Correctness at the benchmark sizes: the 128^3 Kuhn box's complex eddy-current
operator (12.6M tets, 242M nonzeros) and the 100x100x67 box's fused Newton set
(4.02M tets, 77.4M nonzeros), the two operators the benchmark times.

At these sizes the oracle cannot rebuild whole matrices in a test, so the
checks are size-independent properties over every entry, plus sampled
columns recomputed by the oracle's MEX-order kernels on just the elements
incident to each column (every contribution to a CSC column comes from the
elements around its edge):
- DOF map: every element's local edge i is the edge (min, max) of its two
  nodes (Mesh.m:118-135), the edge list is strictly lexicographic, and the
  counts are the exact Kuhn counts;
- symmetry of every matrix to 1e-13 of its largest entry;
- K G = 0 for the discrete gradient (curl grad = 0);
- Im(K + j w M_sigma) = w M_sigma, exactly zero on the columns no plate
  element touches;
- FP32_LOCAL against FP64 to 1e-6 of each matrix's largest entry, over all
  entries;
- sampled columns (and residual rows) against the oracle to 1e-12.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import mass_tables, stiffness_tables

pytestmark = pytest.mark.gpu


def _setup(km):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    return mesh, dh, space, asm


def _columns_of(colptr):
    import torch
    n = colptr.numel() - 1
    return torch.repeat_interleave(torch.arange(n, device=colptr.device), colptr.diff())


def _check_dof_map(mesh, dh, space, edges_expected):
    import torch
    from magpackage_amd.fem_tables import EDGE_NODES
    assert mesh.NumEdges == edges_expected
    E = mesh.Edges.long()                      # (E, 2), 1-based, lo < hi
    assert bool((E[:, 0] < E[:, 1]).all())
    key = E[:, 0] * (mesh.NumNodes + 1) + E[:, 1]
    assert bool((key[1:] > key[:-1]).all())    # sorted and unique (Mesh.m:118-128)
    T = mesh.T_dev.view(-1, 4).long()
    t2e = mesh.T2E.long()                      # 1-based edge ids
    for i, (a, b) in enumerate(EDGE_NODES):
        na, nb = T[:, a], T[:, b]
        ed = E[t2e[:, i] - 1]
        assert torch.equal(ed[:, 0], torch.minimum(na, nb))
        assert torch.equal(ed[:, 1], torch.maximum(na, nb))
        assert torch.equal(mesh.T2E_Sign[:, i].long(), torch.where(na < nb, 1, -1))
    # full-domain Nedelec: DOF = edge id, both 1-based (DofHandler.m:88-105)
    assert torch.equal(dh.DofMaps[space.toString()].long(), t2e)


def _check_symmetric(colptr, rowind, vals, tol, what):
    """vals [nnz] on a CSC pattern (rows sorted per column): |A - A^T| <= tol max|A|."""
    import torch
    n = colptr.numel() - 1
    nnz = int(colptr[-1])
    # sizes checked on the host before any gather on the device
    assert rowind.numel() == nnz and vals.numel() == nnz and not vals.is_complex(), what
    cols = _columns_of(colptr)
    r = rowind.long()
    k1 = cols * n + r                     # CSC order: ascending
    s2, perm = torch.sort(r * n + cols)   # the same entries in CSR order
    assert torch.equal(s2, k1), f"{what}: pattern is not symmetric"
    d = float((vals[perm] - vals).abs().max())
    m = float(vals.abs().max())
    print(f"{what}: max |A - A^T| / max |A| = {d / m:.1e}", flush=True)
    assert d <= tol * m, what


def _check_gradient_kernel(mesh, colptr, rowind, kvals, seed=5):
    """K v = 0 for v = G phi, phi a random nodal field."""
    import torch
    g = torch.Generator(device=kvals.device)
    g.manual_seed(seed)
    phi = torch.randn(mesh.NumNodes, dtype=torch.float64, device=kvals.device, generator=g)
    n = colptr.numel() - 1
    assert mesh.NumEdges == n and rowind.numel() == kvals.numel() == int(colptr[-1])
    E = mesh.Edges.long() - 1
    v = phi[E[:, 1]] - phi[E[:, 0]]
    cols = _columns_of(colptr)
    r = rowind.long()
    y = torch.zeros_like(v).index_add_(0, r, kvals * v[cols])
    s = torch.zeros_like(v).index_add_(0, r, kvals.abs() * v[cols].abs())
    ratio = float(y.abs().max() / s.max())
    print(f"|K G phi| / max(|K| |G phi|) = {ratio:.1e}", flush=True)
    assert ratio <= 1e-12


def _sample_columns(n, count, seed, extra=()):
    rng = np.random.default_rng(seed)
    s = set(int(c) for c in rng.choice(n, size=count, replace=False))
    s.update(int(c) for c in extra)
    s.update((0, n - 1))
    return np.array(sorted(s), dtype=np.int64)


def _incident(mesh, dh, space, sample):
    """Host arrays of the elements incident to the sampled columns, in the
    oracle's MEX layout (1-based T and Dofs, signs as doubles)."""
    import torch
    cell = dh.DofMaps[space.toString()]        # 1-based DOFs, 0 = inactive
    S = torch.from_numpy((sample + 1).astype(np.int32)).to(cell.device)
    eidx = torch.nonzero(torch.isin(cell, S).any(dim=1)).flatten()
    T = mesh.T_dev.view(-1, 4)[eidx].cpu().numpy().astype(np.float64)
    D = cell[eidx].cpu().numpy().astype(np.float64)
    Sg = mesh.T2E_Sign[eidx].cpu().numpy().astype(np.float64)
    return eidx.cpu().numpy(), T, D, Sg


def _engine_columns(colptr_h, rowind, vals, sample):
    """{column: (rows, values)} of the engine's CSC for the sampled columns."""
    import torch
    assert vals.numel() == rowind.numel() == int(colptr_h[-1]) and sample.max() < colptr_h.size - 1
    idx = np.concatenate([np.arange(colptr_h[j], colptr_h[j + 1]) for j in sample])
    it = torch.from_numpy(idx).to(vals.device)
    r = rowind[it].cpu().numpy()
    v = vals[it].cpu().numpy()
    out, o = {}, 0
    for j in sample:
        k = int(colptr_h[j + 1] - colptr_h[j])
        out[int(j)] = (r[o:o + k], v[o:o + k])
        o += k
    return out


def _compare_columns(eng, I, J, V, scale, tol, what):
    """Oracle triplets (1-based) summed per (row, column) against the engine's
    sampled columns; every oracle row must be in the engine's pattern."""
    worst = 0.0
    Jc = J.astype(np.int64) - 1
    order = np.argsort(Jc, kind="stable")
    Jc, Ic, Vc = Jc[order], I.astype(np.int64)[order] - 1, V[order]
    starts = np.searchsorted(Jc, list(eng.keys()), side="left")
    ends = np.searchsorted(Jc, list(eng.keys()), side="right")
    for (j, (rows, vals)), a, b in zip(eng.items(), starts, ends):
        ref = np.zeros(rows.size)
        pos = np.searchsorted(rows, Ic[a:b])
        assert np.all(pos < rows.size) and np.all(rows[np.minimum(pos, rows.size - 1)] == Ic[a:b]), \
            f"{what}: oracle row outside the engine's column {j}"
        np.add.at(ref, pos, Vc[a:b])
        worst = max(worst, float(np.abs(vals - ref).max()) if rows.size else 0.0)
    print(f"{what}: {len(eng)} sampled columns, max |diff| / max |A| = {worst / scale:.1e}",
          flush=True)
    assert worst <= tol * scale, what


@pytest.fixture(scope="module")
def plate128(gpu):
    from magpackage_amd.workloads import plate_box
    km, nu, sigma, omega = plate_box(128, seed=2)
    return km, nu, sigma, omega, _setup(km)


def test_kuhn128_dof_map(plate128):
    km, nu, sigma, omega, (mesh, dh, space, asm) = plate128
    _check_dof_map(mesh, dh, space, 14_827_904)
    assert asm.plan(space).nnz == 241_910_144


def test_kuhn128_complex_operator(plate128):
    import torch
    km, nu, sigma, omega, (mesh, dh, space, asm) = plate128
    plan = asm.plan(space)
    A = asm.assembleFrequencyOperator(space, nu, sigma, omega, drop_zeros=False)
    colptr, rowind = A.colptr, A.rowind
    assert A.values.is_complex() and A.values.numel() == plan.nnz
    re, im = A.values.real.contiguous(), A.values.imag.contiguous()
    _check_symmetric(colptr, rowind, re, 1e-13, "Re (K)")
    _check_symmetric(colptr, rowind, im, 1e-13, "Im (w M_sigma)")
    _check_gradient_kernel(mesh, colptr, rowind, re)
    # Im = w M_sigma, zero on the columns no plate element touches
    M = asm.assembleMassWeighted(space, sigma, drop_zeros=False).values
    dm = float((im - omega * M).abs().max())
    print(f"|Im - w M_sigma| / max = {dm / float(im.abs().max()):.1e}", flush=True)
    assert dm <= 1e-14 * float(im.abs().max())
    cell = dh.DofMaps[space.toString()]
    plate = torch.from_numpy(np.nonzero(sigma > 0)[0]).to(cell.device)
    touched = torch.zeros(plan.n, dtype=torch.bool, device=cell.device)
    touched[cell[plate].flatten().long() - 1] = True
    cols = _columns_of(colptr)
    assert bool((im[~touched[cols]] == 0).all())
    diag = rowind.long() == cols
    assert bool((im[diag & touched[cols]] > 0).all())
    assert bool((re[diag] > 0).all())
    # sampled columns against the oracle (MEX-order kernels + sparse())
    n = plan.n
    plate_cols = cell[plate[:64]].flatten().cpu().numpy() - 1
    sample = _sample_columns(n, 300, seed=11, extra=plate_cols)
    eidx, T, D, Sg = _incident(mesh, dh, space, sample)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    Ki, Kj, Kv = O.curl_curl(km.P, T, D, Sg, nu[eidx], qw, curl)
    Mi, Mj, Mv = O.mass(km.P, T, D, Sg, sigma[eidx], qm, val)
    ch = colptr.cpu().numpy()
    _compare_columns(_engine_columns(ch, rowind, re, sample), Ki, Kj, Kv,
                     float(re.abs().max()), 1e-12, "K columns")
    _compare_columns(_engine_columns(ch, rowind, im, sample), Mi, Mj, omega * Mv,
                     float(im.abs().max()), 1e-12, "w M_sigma columns")


def test_kuhn128_complex_fp32_local(plate128):
    """FP32-local mode over every entry of the 128^3 operator, against FP64."""
    km, nu, sigma, omega, (mesh, dh, space, asm) = plate128
    a64 = asm.assembleFrequencyOperator(space, nu, sigma, omega, drop_zeros=False).values
    a32 = asm.assembleFrequencyOperator(space, nu, sigma, omega, precision="fp32_local",
                                        drop_zeros=False).values
    assert a64.is_complex() and a32.is_complex() and a64.numel() == a32.numel()
    for name, part in (("Re", lambda v: v.real), ("Im", lambda v: v.imag)):
        d = float((part(a32) - part(a64)).abs().max())
        m = float(part(a64).abs().max())
        print(f"FP32_LOCAL {name}: {d / m:.1e}", flush=True)
        assert d <= 1e-6 * m


@pytest.fixture(scope="module")
def core100(gpu):
    from magpackage_amd.workloads import newton_core_box, newton_solution
    km, lib, sigma = newton_core_box()
    mesh, dh, space, asm = _setup(km)
    a = newton_solution(asm, space, dh.NumGlobalDofs, seed=3)
    return km, lib, sigma, a, (mesh, dh, space, asm)


def test_newton_set_fullsize(core100):
    """The fused Newton set on the 100x100x67 box: DOF map, symmetry of K, M,
    J, K G = 0, FP32_LOCAL against FP64 over every entry, and sampled columns
    of K(nu(B)), M_sigma, J and rows of R against the oracle."""
    from magpackage_amd.material import pack_material_data
    km, lib, sigma, a, (mesh, dh, space, asm) = core100
    _check_dof_map(mesh, dh, space, 4_760_467)
    plan = asm.plan(space)
    assert plan.nnz == 77_401_267
    s64 = asm.assembleNewtonSet(space, a, lib, sigma, precision="fp64")
    s32 = asm.assembleNewtonSet(space, a, lib, sigma, precision="fp32_local")
    colptr, rowind = plan.pattern()
    for name in ("K", "M", "J"):
        _check_symmetric(colptr, rowind, s64[name], 1e-13, name)
    _check_gradient_kernel(mesh, colptr, rowind, s64["K"])
    for name in ("K", "M", "J", "R"):
        d = float((s32[name] - s64[name]).abs().max())
        m = float(s64[name].abs().max())
        print(f"FP32_LOCAL {name}: {d / m:.1e}", flush=True)
        assert d <= 1e-6 * m, name
    # sampled columns against the oracle
    n = plan.n
    cell = dh.DofMaps[space.toString()]
    core = np.nonzero(km.RegionTags == 2)[0]
    extra = cell[core[:: max(1, core.size // 64)][:64]].flatten().cpu().numpy() - 1
    sample = _sample_columns(n, 300, seed=12, extra=extra)
    eidx, T, D, Sg = _incident(mesh, dh, space, sample)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    pk = pack_material_data(lib)
    ah = a.cpu().numpy()
    tags = km.RegionTags[eidx]
    Bo = O.element_b(km.P, T, D, Sg, ah, curl)
    nu, _ = O.eval_nu((Bo * Bo).sum(axis=1), tags, pk)
    Ji, Jj, Jv, Ri, Rv = O.jacobian(km.P, T, D, Sg, tags.astype(np.float64), ah, qw, curl, pk)
    Ki, Kj, Kv = O.curl_curl(km.P, T, D, Sg, nu, qw, curl)
    Mi, Mj, Mv = O.mass(km.P, T, D, Sg, sigma[eidx], qm, val)
    ch = colptr.cpu().numpy()
    for name, (I, J, V) in (("K", (Ki, Kj, Kv)), ("M", (Mi, Mj, Mv)), ("J", (Ji, Jj, Jv))):
        _compare_columns(_engine_columns(ch, rowind, s64[name], sample), I, J, V,
                         float(s64[name].abs().max()), 1e-12, f"{name} columns")
    # residual rows: every element containing edge j contributes to R_j
    R = s64["R"].cpu().numpy()
    Rref = np.zeros(n)
    np.add.at(Rref, Ri.astype(np.int64) - 1, Rv)
    d = float(np.abs(R[sample] - Rref[sample]).max())
    print(f"R rows: max |diff| / max |R| = {d / np.abs(R).max():.1e}", flush=True)
    assert d <= 1e-12 * np.abs(R).max()
