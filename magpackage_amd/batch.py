"""Batched frequency-domain assembly: many meshes x many frequencies per launch.

Each mesh gets its own plan; ``BatchAssembler`` builds the engine's batch
(``mpk_batch_create``), lets an autotuner pick the schedule (ring gather
batch x grid shape, timed on the batch itself, cached by batch signature),
and launches the whole batch -- every mesh, every frequency -- as one HIP
graph replay.  Value set f of mesh m is K_m + 1j*omega_f*M_sigma,m on the
mesh's structural CSC pattern, what FrequencySolver.m:127-131 forms per
frequency (without the regularisation term).

No reference counterpart as one call: the reference assembles per mesh and
per frequency; its HBFEM kernel (src/mex/assemble_hbfem_kernel_mex.cpp:25-260)
stacks harmonic blocks of one mesh.
"""
from __future__ import annotations

import ctypes as C
import json
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib as L
from .assembler import Assembler, DofHandler, FunctionSpace, Mesh

# schedules the autotuner tries: (ring gather batch, blocks per CU; 0 = one block per tile)
SCHEDULES = tuple((kb, bpc) for kb in (2, 3, 4, 6) for bpc in (0, 8, 16))


class BatchAssembler:
    def __init__(self, meshes, nus, sigmas, omegas, precision: str = "fp64"):
        """meshes: objects with P (Nn, 3), T (Ne, 4) 1-based, RegionTags;
        nus / sigmas: per-element arrays (one per mesh); omegas: angular
        frequencies (at most 64)."""
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.omegas = np.ascontiguousarray(omegas, dtype=np.float64)
        self.nfreq = int(self.omegas.size)
        self.space = FunctionSpace("Nedelec", 1)
        self.meshes, self.plans, self.asms, self.vals, self._keep = [], [], [], [], []
        sp = L.stream_ptr()
        for km, nu, sg in zip(meshes, nus, sigmas):
            m = Mesh(km.P, km.T, km.RegionTags)
            dh = DofHandler(m)
            dh.distributeDofs(self.space)
            asm = Assembler(m, dh)
            plan = asm.plan(self.space)
            if not plan.ring_order:
                raise ValueError("batched assembly needs manifold meshes (ring-ordered plans)")
            L.check(L.lib.mpk_plan_set_coords(plan.handle, L.ptr(m.P_dev), sp))
            nu_d = torch.from_numpy(np.ascontiguousarray(nu, dtype=np.float64)).to(self.dev)
            sg_d = torch.from_numpy(np.ascontiguousarray(sg, dtype=np.float64)).to(self.dev)
            out = torch.empty(max(2 * self.nfreq * plan.nnz, 2), dtype=torch.float64,
                              device=self.dev)
            self.meshes.append(m)
            self.asms.append(asm)
            self.plans.append(plan)
            self.vals.append(out)
            self._keep += [nu_d, sg_d]
        self.items = (L.BatchItem * len(self.plans))()
        for i, (plan, out) in enumerate(zip(self.plans, self.vals)):
            self.items[i].plan = plan.handle.value
            self.items[i].nu = L.ptr(self._keep[2 * i])
            self.items[i].sigma = L.ptr(self._keep[2 * i + 1])
            self.items[i].vals_c128 = L.ptr(out)
        self.qk = self.asms[0].stiffness_quadrature()
        self.qm = self.asms[0].mass_quadrature(weighted=True)
        self.prec = L.PRECISIONS[precision]
        self.handle = self._create(self.items, len(self.plans))
        nt, nz, lds = C.c_int64(), C.c_int64(), C.c_int64()
        L.check(L.lib.mpk_batch_info(self.handle, C.byref(nt), C.byref(nz), C.byref(lds)))
        self.ntiles, self.nnz_total, self.lds_bytes = nt.value, nz.value, lds.value
        self.schedule = (3, 0)
        self.graph = None

    def _create(self, items, count):
        h = C.c_void_p()
        L.check(L.lib.mpk_batch_create(items, count, self.omegas.ctypes.data_as(C.c_void_p),
                                       self.nfreq, C.byref(self.qk), C.byref(self.qm), self.prec,
                                       L.stream_ptr(), C.byref(h)))
        return h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and getattr(L, "lib", None) is not None:
            L.lib.mpk_batch_destroy(h)

    # -------------------------------------------------------------- tuning
    @property
    def signature(self) -> str:
        ne = sum(m.NumElements for m in self.meshes)
        return f"{len(self.meshes)}m-{ne}e-{self.nnz_total}nz-{self.nfreq}f-p{self.prec}"

    def set_schedule(self, ring_batch: int, blocks_per_cu: int):
        L.check(L.lib.mpk_batch_schedule(self.handle, ring_batch, blocks_per_cu))
        self.schedule = (ring_batch, blocks_per_cu)
        self.graph = None

    def run(self):
        L.check(L.lib.mpk_batch_run(self.handle, L.stream_ptr()))

    def _time(self, reps: int) -> float:
        stream = torch.cuda.current_stream()
        self.run()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            self.run()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def autotune(self, cache: str | Path | None = None, reps: int = 2, log=None) -> dict:
        """Times every schedule of SCHEDULES on this batch and keeps the
        fastest; with ``cache`` (a JSON file), a schedule tuned before for the
        same batch signature is reused."""
        key = self.signature
        store = {}
        if cache is not None and Path(cache).exists():
            store = json.loads(Path(cache).read_text())
            if key in store:
                kb, bpc = store[key]["schedule"]
                self.set_schedule(kb, bpc)
                return store[key]
        times = {}
        for kb, bpc in SCHEDULES:
            self.set_schedule(kb, bpc)
            times[f"{kb},{bpc}"] = self._time(reps)
            if log:
                log(f"[autotune] ring batch {kb}, blocks/CU {bpc}: {times[f'{kb},{bpc}']:.3f} ms")
        best = min(times, key=times.get)
        kb, bpc = (int(v) for v in best.split(","))
        self.set_schedule(kb, bpc)
        rec = {"schedule": [kb, bpc], "ms": times}
        if cache is not None:
            store[key] = rec
            Path(cache).write_text(json.dumps(store, indent=1))
        return rec

    # ------------------------------------------------------------- running
    def capture(self):
        """Captures one batch launch into a HIP graph (torch.cuda.CUDAGraph)."""
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run()  # warm the launch path outside the capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def values(self, mesh: int, freq: int) -> torch.Tensor:
        """Complex128 values (interleaved doubles) of mesh `mesh`, frequency `freq`."""
        nnz = self.plans[mesh].nnz
        return self.vals[mesh][2 * freq * nnz: 2 * (freq + 1) * nnz]

    def per_mesh_latency(self, reps: int = 3) -> np.ndarray:
        """Milliseconds to assemble each mesh alone (all frequencies, one
        launch, the batch's schedule), one batch of one item per mesh."""
        out = []
        for i in range(len(self.plans)):
            one = (L.BatchItem * 1)(self.items[i])
            h = self._create(one, 1)
            try:
                L.check(L.lib.mpk_batch_schedule(h, *self.schedule))
                stream = torch.cuda.current_stream()
                L.check(L.lib.mpk_batch_run(h, L.stream_ptr()))
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    L.check(L.lib.mpk_batch_run(h, L.stream_ptr()))
                e1.record(stream)
                torch.cuda.synchronize()
                out.append(e0.elapsed_time(e1) / reps)
            finally:
                L.lib.mpk_batch_destroy(h)
        return np.asarray(out)


def log_uniform_shapes(count: int = 256, lo: float = 5e3, hi: float = 2e5, seed: int = 4):
    """Box shapes (nx, ny, nz) and mesh seeds of the log-uniform batch: tet
    counts log-uniform on [lo, hi] (seeded), each box nx x ny x nz cubes
    (6 tets per cube) close to a cube in shape."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(lo), np.log(hi), count))
    out = []
    for i, ne in enumerate(targets):
        c = max(ne / 6.0, 1.0)
        nx = max(int(round(c ** (1.0 / 3.0))), 2)
        ny = max(int(round((c / nx) ** 0.5)), 2)
        nz = max(int(round(c / (nx * ny))), 2)
        out.append(((nx, ny, nz), seed * 1000 + i))
    return out


def log_uniform_mesh(shape, mesh_seed: int):
    """One mesh of the log-uniform batch: a jittered Kuhn box with the
    conducting slab (tag 2, centroid z in [0.45, 0.55])."""
    from .meshgen import kuhn_box, tag_slab
    km = kuhn_box(*shape, seed=mesh_seed)
    tag_slab(km, 2, 0.45, 0.55, 2)
    return km


def log_uniform_meshes(count: int = 256, lo: float = 5e3, hi: float = 2e5, seed: int = 4):
    """The log-uniform batch's meshes (log_uniform_shapes, log_uniform_mesh)."""
    return [log_uniform_mesh(s, ms) for s, ms in log_uniform_shapes(count, lo, hi, seed)]
