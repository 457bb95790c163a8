"""Batched assembly (many meshes x frequencies, one launch) against the
single-mesh complex assembly, and its graph replay."""
import ctypes as C

import numpy as np
import pytest

from magpackage_amd.material import MU0
from magpackage_amd.meshgen import kuhn_box, tag_slab

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(gpu):
    import torch
    from magpackage_amd.batch import BatchAssembler
    shapes = [(3, 4, 2), (7, 5, 6), (10, 10, 10), (4, 9, 3)]
    meshes, nus, sigmas = [], [], []
    for i, s in enumerate(shapes):
        km = kuhn_box(*s, seed=50 + i)
        tag_slab(km, 2, 0.45, 0.55, 2)
        meshes.append(km)
        nus.append(np.random.default_rng(i).uniform(0.5, 2.0, km.num_elems) / MU0)
        sigmas.append(np.where(km.RegionTags == 2, 3.5e7, 0.0))
    omegas = 2 * np.pi * np.array([10.0, 330.0, 1e4])
    ba = BatchAssembler(meshes, nus, sigmas, omegas)
    return torch, ba, nus, sigmas, omegas


def _single(ba, i, nu, sg, omega):
    import torch
    from magpackage_amd import _lib as L
    plan = ba.plans[i]
    out = torch.empty(2 * plan.nnz, dtype=torch.float64, device="cuda")
    nd = torch.from_numpy(nu).cuda()
    sd = torch.from_numpy(sg).cuda()
    L.check(L.lib.mpk_csc_curl_curl_mass_complex(plan.handle, None, L.ptr(nd), L.ptr(sd),
                                                 float(omega), C.byref(ba.qk), C.byref(ba.qm),
                                                 L.PREC_FP64, L.ptr(out), L.stream_ptr()))
    return out


def test_batch_matches_single_mesh_bitwise(batch):
    torch, ba, nus, sigmas, omegas = batch
    refs = {(i, f): _single(ba, i, nus[i], sigmas[i], w)
            for i in range(len(nus)) for f, w in enumerate(omegas)}
    from magpackage_amd.batch import SCHEDULES
    for kb, bpc in SCHEDULES:
        ba.set_schedule(kb, bpc)
        for v in ba.vals:
            v.fill_(float("nan"))
        ba.run()
        torch.cuda.synchronize()
        for (i, f), ref in refs.items():
            got = ba.values(i, f)
            assert torch.equal(got.view(torch.int64), ref.view(torch.int64)), (kb, bpc, i, f)


def test_batch_graph_replay(batch):
    """Every replay, with eager work and synchronisation between replays,
    gives the eager result."""
    torch, ba, nus, sigmas, omegas = batch
    ba.set_schedule(3, 8)
    ba.run()
    torch.cuda.synchronize()
    eager = [v.clone() for v in ba.vals]
    ba.capture()
    for rep in range(4):
        for v in ba.vals:
            v.fill_(0.0)
        x = torch.randn(1000, device="cuda").sum()  # unrelated eager work
        torch.cuda.synchronize()
        ba.replay()
        torch.cuda.synchronize()
        for v, e in zip(ba.vals, eager):
            assert torch.equal(v.view(torch.int64), e.view(torch.int64)), rep


def test_autotune_and_latency(batch, tmp_path):
    torch, ba, nus, sigmas, omegas = batch
    cache = tmp_path / "tune.json"
    rec = ba.autotune(cache=cache, reps=1)
    from magpackage_amd.batch import SCHEDULES
    assert tuple(rec["schedule"]) == ba.schedule and len(rec["ms"]) == len(SCHEDULES)
    rec2 = ba.autotune(cache=cache)  # reuses the cached schedule
    assert rec2["schedule"] == rec["schedule"]
    lat = ba.per_mesh_latency(reps=1)
    assert lat.shape == (4,) and np.all(lat > 0)
