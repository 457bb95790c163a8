#!/usr/bin/env python3
"""Benchmark of the MI355X assembly engine (one GPU).

Default workload: frequency-domain eddy-current operator K + 1j*omega*M_sigma
as a complex128 CSC matrix on a 128^3-cube jittered Kuhn box (12,582,912
tets, 14,827,904 edge DOFs), a conducting plate (centroid z in
[0.45, 0.55], sigma = 3.5e7 S/m), mu_r = 1 everywhere, 50 Hz.  One step =
loading the caller's coordinates into the plan + one assembly launch over
the whole mesh, inputs resident in HBM, the CSC pattern plan built once
before timing (its cost is reported separately).  ``value`` is the step
rate; the roofline figure divides the step's algorithmic bytes by the
assembly kernel's own duration (HIP events around that launch).

Prints ONE JSON line (rank 0).  The timed region is exactly --steps launches
between two device synchronisations; the kernel duration is measured with
HIP events on the stream the kernel runs on.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "assembled tets/s + % HBM roofline; nonlinear step time learned vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_freq(nn, ne, nnz):
    """Inputs counted once at stored width + complex128 CSC values written once:
    coordinates 24*Nn, connectivity 16*Ne, edge DOFs 24*Ne, nu and sigma 8*Ne each,
    values 16*nnz."""
    return 24 * nn + 16 * ne + 24 * ne + 8 * ne + 8 * ne + 16 * nnz


def freq_workload(n=128, seed=2):
    from magpackage_amd.material import MU0
    from magpackage_amd.meshgen import kuhn_box, tag_slab
    km = kuhn_box(n, n, n, seed=seed)
    tag_slab(km, 2, 0.45, 0.55, 2)
    nu = np.full(km.num_elems, 1.0 / MU0)
    sigma = np.where(km.RegionTags == 2, 3.5e7, 0.0)
    return km, nu, sigma, 2 * np.pi * 50.0


def cpu_baseline(sample_n: int, reps: int):
    """Reference path on the host: MEX-order COO kernels (oracle, OpenMP) +
    sparse() for K and M_sigma + K + 1j*omega*M on a bounded sample."""
    from oracle import oracle as O
    from magpackage_amd.fem_tables import get_quadrature_data, matlab_flat, nedelec_tet_p1
    km, nu, sigma, omega = freq_workload(sample_n, seed=2)
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    q_pts, qw = get_quadrature_data("tet", 2)
    val, curl = nedelec_tet_p1(q_pts)
    val, curl = matlab_flat(val), matlab_flat(curl)

    def run():
        t0 = time.perf_counter()
        I, J, V = O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl)
        K = O.sparse_scipy(I, J, V, n, n)
        del I, J, V
        I, J, V = O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sigma, qw, val)
        M = O.sparse_scipy(I, J, V, n, n)
        del I, J, V
        A = K + 1j * omega * M
        t1 = time.perf_counter()
        return t1 - t0, A.nnz

    run()  # warm-up
    times = [run()[0] for _ in range(reps)]
    t = float(np.median(times))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": km.num_elems / t,
        "unit": "tets/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"Kuhn {sample_n}^3 box ({km.num_elems} tets, same plate/sigma/50 Hz), "
                   f"COO kernels on {cores} OpenMP threads + serial sparse() + K+1j*w*M, "
                   f"median of {reps} after 1 warm-up: {t:.3f} s"),
    }


def read_traffic(workload: str):
    f = ROOT / "profiles" / f"traffic_{workload}.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cubes per axis of the Kuhn box")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp64_exact", "fp32_local"])
    ap.add_argument("--cpu-sample", type=int, default=80)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib", action="store_true",
                    help="after timing, run a plain 512 MiB elementwise copy (reads and writes "
                         "512 MiB) so HBM counters can be scaled against a known size")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine is measured on one GPU only; "
                                   "run with --gpus 1 (multi-GPU sharding is a design note)"}))
        sys.exit(2)

    import torch
    from magpackage_amd import _lib as L
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    km, nu, sigma, omega = freq_workload(args.n)
    log(f"[bench] mesh {args.n}^3: {km.num_elems} tets, {km.num_nodes} nodes "
        f"({time.perf_counter() - t0:.1f} s)")
    mesh = Mesh.from_kuhn(km)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    asm = Assembler(mesh, dh)
    plan = asm.plan(space)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    log(f"[bench] edges+dofs {1e3 * (t2 - t1):.1f} ms, plan {1e3 * (t3 - t2):.1f} ms: "
        f"{dh.NumGlobalDofs} dofs, nnz {plan.nnz}, max tile {plan.max_tile_nnz}")

    dev = torch.device("cuda", 0)
    nu_d = torch.from_numpy(nu).to(dev)
    sg_d = torch.from_numpy(sigma).to(dev)
    vals = torch.empty(2 * plan.nnz, dtype=torch.float64, device=dev)
    qk = asm.stiffness_quadrature()
    qm = asm.mass_quadrature(weighted=True)
    prec = L.PRECISIONS[args.precision]
    stream = torch.cuda.current_stream()
    sp = L.stream_ptr(stream)

    pd = L.ptr(mesh.P_dev)

    def step(ev_a=None, ev_b=None):
        # the caller's P goes into the plan's padded coordinate copy, then the
        # assembly kernel reads that copy (P argument NULL)
        L.check(L.lib.mpk_plan_set_coords(plan.handle, pd, sp))
        if ev_a is not None:
            ev_a.record(stream)
        L.check(L.lib.mpk_csc_curl_curl_mass_complex(
            plan.handle, None, L.ptr(nu_d), L.ptr(sg_d), float(omega), C.byref(qk),
            C.byref(qm), prec, L.ptr(vals), sp))
        if ev_b is not None:
            ev_b.record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    eva = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    evb = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    tw0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step(eva[i], evb[i])
    ev1.record(stream)
    torch.cuda.synchronize()
    tw1 = time.perf_counter()
    ms_total = ev0.elapsed_time(ev1)
    ms_per = ms_total / args.steps
    ms_kernel = float(np.mean([a.elapsed_time(b) for a, b in zip(eva, evb)]))
    ne, nn, nnz = km.num_elems, km.num_nodes, plan.nnz
    value = ne / (ms_per * 1e-3)
    alg = algorithmic_bytes_freq(nn, ne, nnz)
    achieved = alg / (ms_kernel * 1e-3) / 1e9
    log(f"[bench] {args.steps} steps: {ms_per:.4f} ms/step (events), assembly kernel "
        f"{ms_kernel:.4f} ms, wall {1e3 * (tw1 - tw0):.1f} ms, {value:.4g} tets/s, "
        f"{achieved:.0f} GB/s algorithmic")

    if args.calib:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 3 x 512 MiB read + 512 MiB write")
        del a, b

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline on a {args.cpu_sample}^3 sample ...")
        cpu = cpu_baseline(args.cpu_sample, args.cpu_reps)
        log(f"[bench] cpu: {cpu['value']:.4g} tets/s")

    workload = f"kuhn{args.n}_freq_k_jwsm_c128"
    traffic = read_traffic(workload) if args.n == 128 else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "fp32_local" else "f32-local/f64-acc",
        "data": "synthetic (seeded jittered Kuhn box)",
        "config": {
            "workload": workload,
            "operator": "K + 1j*omega*M_sigma, complex128 CSC",
            "tets": ne, "nodes": nn, "edge_dofs": int(dh.NumGlobalDofs), "nnz": int(nnz),
            "precision": args.precision, "frequency_hz": 50.0, "plate_sigma": 3.5e7,
            "plan_build_ms": round(1e3 * (t3 - t2), 1),
            "edges_dofs_ms": round(1e3 * (t2 - t1), 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": alg,
            "kernel_ms": ms_kernel,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
