"""The allocation-free fused Newton set (mpk_csc_newton_set_noalloc) against
mpk_csc_newton_set, both called through the C ABI: identical K, M, J and R
bits in FP64 and FP32_LOCAL, on two calls in a row, the second on a smaller
mesh (the static scratch keeps the first call's contents, which must not
leak into the second), and the size limit refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(nx, ny, nz):
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.material import create_linear, marrocco
    from magpackage_amd.meshgen import kuhn_box, tag_box
    from magpackage_amd.workloads import CORE_SIGMA, newton_solution
    # unit box: the iron core (Marrocco, conducting) is inside at every size
    km = kuhn_box(nx, ny, nz, seed=7)
    tag_box(km, (0.3, 0.3, 0.2), (0.7, 0.7, 0.47), 2)
    lib = {1: create_linear(1.0), 2: marrocco()}
    sigma = np.where(km.RegionTags == 2, CORE_SIGMA, 0.0)
    assert (km.RegionTags == 2).any()
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    plan = asm.plan(space)
    a = newton_solution(asm, space, plan.n, seed=7)
    mat, keep = asm.material_tables(lib)
    sg = torch.from_numpy(sigma).cuda()
    return asm, mesh, plan, a, mat, keep, sg


def _run(fn, prob, prec):
    import torch
    from magpackage_amd import _lib as L
    asm, mesh, plan, a, mat, keep, sg = prob
    out = [torch.full((plan.nnz,), float("nan"), dtype=torch.float64, device="cuda")
           for _ in range(3)]
    out.append(torch.full((plan.n,), float("nan"), dtype=torch.float64, device="cuda"))
    qk = asm.stiffness_quadrature()
    qm = asm.mass_quadrature(weighted=True)
    rc = fn(plan.handle, L.ptr(mesh.P_dev), L.ptr(mesh.Tags_dev), L.ptr(a), C.byref(mat),
            L.ptr(sg), C.byref(qk), C.byref(qm), prec, *(L.ptr(o) for o in out),
            L.stream_ptr())
    L.check(rc)
    torch.cuda.synchronize()
    return out


def test_noalloc_matches_newton_set(gpu):
    import torch
    from magpackage_amd import _lib as L
    for prec in (L.PREC_FP64, L.PREC_FP32_LOCAL):
        for shape in ((40, 40, 27), (12, 12, 8)):    # second call: smaller mesh
            prob = _problem(*shape)
            print(f"precision {prec}, {shape}: {prob[2].ne} tets, {prob[2].n} DOFs", flush=True)
            ref = _run(L.lib.mpk_csc_newton_set, prob, prec)
            got = _run(L.lib.mpk_csc_newton_set_noalloc, prob, prec)
            for name, r, g in zip("KMJR", ref, got):
                assert bool(torch.isfinite(g).all()), name
                assert torch.equal(g.view(torch.int64), r.view(torch.int64)), (prec, shape, name)


def test_noalloc_size_limit(gpu):
    """A plan beyond the scratch is refused (MPK_ERR_LIMIT) before any launch;
    the same plan fits in FP32_LOCAL (24 B per element) and matches."""
    import torch
    from magpackage_amd import _lib as L
    from magpackage_amd._lib import MpkError
    prob = _problem(80, 80, 80)      # 3,072,000 tets: 147 MB of FP64 state
    with pytest.raises(MpkError, match="scratch") as e:
        _run(L.lib.mpk_csc_newton_set_noalloc, prob, L.PREC_FP64)
    assert e.value.code == L.MPK_ERR_LIMIT
    ref = _run(L.lib.mpk_csc_newton_set, prob, L.PREC_FP32_LOCAL)
    got = _run(L.lib.mpk_csc_newton_set_noalloc, prob, L.PREC_FP32_LOCAL)
    for name, r, g in zip("KMJR", ref, got):
        assert torch.equal(g.view(torch.int64), r.view(torch.int64)), name
