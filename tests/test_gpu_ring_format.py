"""The compact 8-byte ring records against the 16-byte ones: both formats
carry the same frames, orientation bits and in-column positions, so every
closed-form kernel (curl-curl, sigma-mass, K + j w M_sigma, the fused Newton
set in both of its passes, the batched frequency kernel) writes identical
bits with either.  The 16-byte format is chosen by MPK_RING_FORMAT=16 at plan
creation."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plans(km):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    out = {}
    for fmt in ("compact", "16"):
        old = os.environ.get("MPK_RING_FORMAT")
        if fmt == "16":
            os.environ["MPK_RING_FORMAT"] = "16"
        try:
            mesh = Mesh.from_kuhn(km) if hasattr(km, "num_elems") else km
            dh = DofHandler(mesh)
            space = FunctionSpace("Nedelec", 1)
            dh.distributeDofs(space)
            asm = Assembler(mesh, dh)
            plan = asm.plan(space)
        finally:
            if old is None:
                os.environ.pop("MPK_RING_FORMAT", None)
            else:
                os.environ["MPK_RING_FORMAT"] = old
        assert plan.ring_format == fmt
        out[fmt] = (asm, space)
    return out


def _same(a, b, what):
    import torch
    assert a.dtype == b.dtype and a.shape == b.shape, what
    if a.is_complex():
        a, b = torch.view_as_real(a), torch.view_as_real(b)
    assert torch.equal(a.contiguous().view(torch.int64), b.contiguous().view(torch.int64)), what


@pytest.mark.parametrize("mesh_kind", ["kuhn", "ecore_real"])
def test_formats_bitwise(gpu, mesh_kind):
    from magpackage_amd.material import MU0, create_linear, marrocco
    from magpackage_amd.meshgen import kuhn_box, tag_box
    if mesh_kind == "kuhn":
        km = kuhn_box(24, 20, 18, seed=4)
        tag_box(km, (0.3, 0.3, 0.2), (0.7, 0.7, 0.6), 2)
        tags = km.RegionTags
    else:
        from magpackage_amd.assembler import Mesh
        # the reference's E-core COMSOL mesh; its last region stands in for iron
        d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ecore_mesh.npz"))
        tags = np.where(d["tags"] == d["tags"].max(), 2, 1).astype(np.int32)
        km = Mesh(d["P"], d["T"], tags)
    plans = _plans(km)
    rng = np.random.default_rng(3)
    ne = len(tags)
    nu = np.where(tags == 2, rng.uniform(1, 3, ne) / (1000 * MU0), 1.0 / MU0)
    sigma = np.where(tags == 2, 2e6, 0.0)
    lib = {1: create_linear(1.0), 2: marrocco()}
    res = {}
    for fmt, (asm, space) in plans.items():
        from magpackage_amd.workloads import newton_solution
        r = {"K": asm.assembleStiffness(space, nu, drop_zeros=False).values,
             "M": asm.assembleMassWeighted(space, sigma, drop_zeros=False).values,
             "A": asm.assembleFrequencyOperator(space, nu, sigma, 2 * np.pi * 50,
                                                drop_zeros=False).values,
             "A32": asm.assembleFrequencyOperator(space, nu, sigma, 2 * np.pi * 50,
                                                  precision="fp32_local",
                                                  drop_zeros=False).values}
        a = newton_solution(asm, space, asm.plan(space).n, seed=5)
        for prec in ("fp64", "fp32_local"):
            for k, v in asm.assembleNewtonSet(space, a, lib, sigma, precision=prec).items():
                r[f"newton_{prec}_{k}"] = v
        res[fmt] = r
    for k in res["compact"]:
        _same(res["compact"][k], res["16"][k], f"{mesh_kind} {k}")
