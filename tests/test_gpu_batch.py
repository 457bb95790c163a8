"""Batched assembly (many meshes x frequencies, one launch) against the
oracle's K + 1j w M_sigma and the single-mesh complex assembly, and its graph
replay."""
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
    ba.src_meshes = meshes
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


def _oracle_km(km, nu, sg):
    """K and M_sigma of one mesh from the oracle's MEX-order kernels + sparse()."""
    from oracle import oracle as O
    from tests.helpers import mass_tables, stiffness_tables
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    P, T, D, S = mi["P"], mi["T"], mi["Dofs"], mi["Signs"]
    K = O.sparse_scipy(*O.curl_curl(P, T, D, S, nu, qw, curl), n, n)
    M = O.sparse_scipy(*O.mass(P, T, D, S, sg, qm, val), n, n)
    return K, M


def _engine_csc(ba, i, f):
    import scipy.sparse as sp
    plan = ba.plans[i]
    colptr, rowind = plan.pattern()
    v = ba.values(i, f).view(__import__("torch").complex128).cpu().numpy()
    return sp.csc_matrix((v, rowind.cpu().numpy(), colptr.cpu().numpy()),
                         shape=(plan.n, plan.n))


def _check_against_oracle(A, K, M, w, tag):
    """Real part against K, imaginary part against w M_sigma, each to 1e-12 of
    its own largest entry (the engine keeps the structural pattern: entries
    the oracle's sparse() drops must be zero or rounding-level)."""
    dk = abs(A.real - K).max()
    dm = abs(A.imag - w * M).max()
    mk, mm = abs(K).max(), w * abs(M).max()
    im = f"{dm / mm:.1e}" if mm > 0 else f"{dm:.1e} (no conducting element: M_sigma = 0)"
    print(f"{tag}: |dRe| {dk / mk:.1e}, |dIm| {im}", flush=True)
    assert dk <= 1e-12 * mk and dm <= 1e-12 * mm, tag


def test_batch_against_oracle(batch):
    """Every (mesh, frequency) value set of the fixture batch against the
    oracle's K + 1j w M_sigma (FrequencySolver.m:127-131 without the
    regularisation term)."""
    torch, ba, nus, sigmas, omegas = batch
    ba.set_schedule(3, 0)
    ba.run()
    torch.cuda.synchronize()
    for i, km in enumerate(ba.src_meshes):
        K, M = _oracle_km(km, nus[i], sigmas[i])
        for f, w in enumerate(omegas):
            _check_against_oracle(_engine_csc(ba, i, f), K, M, w, f"mesh {i} f {f}")


def test_log_uniform_batch_against_oracle(gpu):
    """The log-uniform batch's own meshes and frequencies (256 Kuhn boxes of
    5e3-2e5 tets, seed 4; 32 frequencies log-spaced over 10 Hz - 10 kHz): its
    smallest and largest mesh, with two others, assembled as one batch under
    the autotuner's pick; value sets at the lowest, a middle and the highest
    frequency against the oracle."""
    import torch
    from magpackage_amd.batch import BatchAssembler, log_uniform_mesh, log_uniform_shapes
    shapes = log_uniform_shapes(256, seed=4)
    ne = [6 * a * b * c for (a, b, c), _ in shapes]
    pick = [int(np.argmin(ne)), int(np.argmax(ne)), 0, 1]
    meshes = [log_uniform_mesh(*shapes[i]) for i in pick]
    nus = [np.full(m.num_elems, 1.0 / MU0) for m in meshes]
    sigmas = [np.where(m.RegionTags == 2, 3.5e7, 0.0) for m in meshes]
    omegas = 2 * np.pi * np.logspace(1, 4, 32)
    ba = BatchAssembler(meshes, nus, sigmas, omegas)
    ba.autotune(reps=1)
    for v in ba.vals:
        v.fill_(float("nan"))
    ba.replay()
    torch.cuda.synchronize()
    for i, km in enumerate(meshes):
        K, M = _oracle_km(km, nus[i], sigmas[i])
        for f in (0, 17, 31):
            _check_against_oracle(_engine_csc(ba, i, f), K, M, omegas[f],
                                  f"mesh {pick[i]} ({km.num_elems} tets) f {omegas[f] / 2 / np.pi:.4g} Hz")
