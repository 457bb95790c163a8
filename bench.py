#!/usr/bin/env python3
"""Benchmark of the MI355X assembly engine (one GPU).

Default line (`python bench.py`): frequency-domain eddy-current operator
K + 1j*omega*M_sigma as a complex128 CSC matrix on a 128^3-cube jittered Kuhn
box (12,582,912 tets, 14,827,904 edge DOFs), a conducting plate (centroid z
in [0.45, 0.55], sigma = 3.5e7 S/m), mu_r = 1 everywhere, 50 Hz.  One step =
loading the caller's coordinates into the plan + one assembly launch over
the whole mesh, inputs resident in HBM, the CSC pattern plan built once
before timing (its cost is reported separately).  ``value`` is the step
rate; the roofline figure divides the step's algorithmic bytes by the
assembly kernel's own duration (HIP events around that launch).  The same
line carries, under extra keys:
  fp32_local          the operator's FP32-local mode (time, error vs FP64);
  newton_set          the fused K(nu(B)) + M_sigma + J + R set on the
                      100x100x67 box (4.02M tets), FP32-local, with its own
                      roofline, traffic and CPU baseline;
  nonlinear_step_time the E-core transient (983,040 tets), 200 backward-Euler
                      steps of 1 ms, every Newton linear system solved to
                      1e-10 (the reference solves it directly): plain Newton
                      against the learned warm start trained on steps 1-40,
                      over steps 41-200;
  nonlinear_step_time_inexact_newton
                      the same transient with the engine's inexact Newton
                      (Eisenstat-Walker forcing), 80 steps: plain Newton,
                      Newton from the extrapolated start and the learned warm
                      start over steps 41-80.

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
    from magpackage_amd.workloads import plate_box
    return plate_box(n, seed)


def algorithmic_bytes_newton(nn, ne, n, nnz):
    """Fused Newton set: coordinates 24*Nn, connectivity 16*Ne, edge DOFs
    24*Ne, sigma 8*Ne, tags 4*Ne, A 8*N in; K, M, J 8*nnz each and R 8*N out."""
    return 24 * nn + 16 * ne + 24 * ne + 8 * ne + 4 * ne + 8 * n + 3 * 8 * nnz + 8 * n


def newton_workload(nx=100, ny=100, nz=67, seed=3):
    """100x100x67-cube jittered Kuhn box (cube edge 0.01), an iron core
    (centroid in [0.3,0.7]x[0.3,0.7]x[0.2,0.47], Marrocco curve, sigma 2e6)
    in air (nu0 = 1/mu0)."""
    from magpackage_amd.workloads import newton_core_box
    return newton_core_box(nx, ny, nz, seed)


def cpu_baseline_newton(sample: int, reps: int):
    """Reference path on the host for the Newton set: the oracle's MEX-order
    Jacobian (J, R), curl-curl (K at the element nu) and mass kernels on
    OpenMP threads + sparse() of each (the oracle's threaded COO -> CSC, bit
    for bit the serial one), on a bounded sample of the workload."""
    from oracle import oracle as O
    from magpackage_amd.fem_tables import get_quadrature_data, matlab_flat, nedelec_tet_p1
    from magpackage_amd.material import pack_material_data
    nz = max(2, int(round(sample * 0.67)))
    km, lib, sigma = newton_workload(sample, sample, nz, seed=3)
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    packed = pack_material_data(lib)
    q_pts, qw = get_quadrature_data("tet", 2)
    val, curl = nedelec_tet_p1(q_pts)
    val, curl = matlab_flat(val), matlab_flat(curl)
    a = np.random.default_rng(3).normal(size=n) * 1e-3
    tags = km.RegionTags.astype(np.float64)

    def run():
        ts = 0.0
        t0 = time.perf_counter()
        I, J, V, Ri, Rv = O.jacobian(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, a, qw,
                                     curl, packed)
        u = time.perf_counter()
        Jm = O.sparse_par(I, J, V, n, n)
        keep = Ri > 0
        R = np.bincount(Ri[keep].astype(np.int64) - 1, weights=Rv[keep], minlength=n)
        ts += time.perf_counter() - u
        B = O.element_b(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], a, curl)
        nu, _ = O.eval_nu((B * B).sum(axis=1), km.RegionTags, packed)
        trip = O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], nu, qw, curl)
        u = time.perf_counter()
        K = O.sparse_par(*trip, n, n)
        ts += time.perf_counter() - u
        trip = O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sigma, qw, val)
        u = time.perf_counter()
        M = O.sparse_par(*trip, n, n)
        ts += time.perf_counter() - u
        del trip
        t1 = time.perf_counter()
        return t1 - t0, (t1 - t0) - ts, ts

    run()
    runs = [run() for _ in range(reps)]
    t = float(np.median([r[0] for r in runs]))
    tk = float(np.median([r[1] for r in runs]))
    tsp = float(np.median([r[2] for r in runs]))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": km.num_elems / t,
        "unit": "tets/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"Kuhn {sample}x{sample}x{nz} box ({km.num_elems} tets, same core and "
                   f"curve), Jacobian + B + curl-curl + mass kernels on {cores} OpenMP threads "
                   f"+ threaded sparse() of K, M, J, median of {reps} after 1 warm-up: {t:.3f} s "
                   f"(kernels {tk:.3f} s, sparse() {tsp:.3f} s)"),
    }


def run_newton(args, emit=True):
    """K(nu(B)) + M_sigma + J + R in one fused pass, FP32-local by default.
    Returns the result line (printed when ``emit``)."""
    import torch
    from magpackage_amd import _lib as L
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    km, lib, sigma = newton_workload()
    log(f"[bench] mesh 100x100x67: {km.num_elems} tets, {km.num_nodes} nodes "
        f"({time.perf_counter() - t0:.1f} s)")
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    plan = asm.plan(space)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    n = plan.n
    dev = torch.device("cuda", 0)
    # A ~ N(0, 1) (seed 3), scaled so that max |B_e| = 2.2 T
    from magpackage_amd.workloads import newton_solution
    a = newton_solution(asm, space, n, seed=3)
    bn = asm.computeElementB(a, space).norm(dim=1)
    log(f"[bench] plan {1e3 * (t3 - t2):.1f} ms: {n} dofs, nnz {plan.nnz}; "
        f"|B| in [{float(bn.min()):.3g}, {float(bn.max()):.3g}] T")
    mat, keep = asm.material_tables(lib)
    sg = torch.from_numpy(sigma).to(dev)
    kv = torch.empty(plan.nnz, dtype=torch.float64, device=dev)
    mv = torch.empty_like(kv)
    jv = torch.empty_like(kv)
    rv = torch.empty(n, dtype=torch.float64, device=dev)
    qk = asm.stiffness_quadrature()
    qm = asm.mass_quadrature(weighted=True)
    prec = L.PRECISIONS[args.precision]
    stream = torch.cuda.current_stream()
    sp = L.stream_ptr(stream)
    pd = L.ptr(mesh.P_dev)

    def step(ev_a=None, ev_b=None):
        L.check(L.lib.mpk_plan_set_coords(plan.handle, pd, sp))
        if ev_a is not None:
            ev_a.record(stream)
        L.check(L.lib.mpk_csc_newton_set(plan.handle, None, L.ptr(mesh.Tags_dev), L.ptr(a),
                                         C.byref(mat), L.ptr(sg), C.byref(qk), C.byref(qm), prec,
                                         L.ptr(kv), None if args.split_m else L.ptr(mv),
                                         L.ptr(jv), L.ptr(rv), sp))
        if args.split_m:
            L.check(L.lib.mpk_csc_mass(plan.handle, None, L.ptr(sg), C.byref(qm), prec,
                                       L.ptr(mv), sp))
        if ev_b is not None:
            ev_b.record(stream)

    ms_per, ms_kernel = timed_steps(step, stream, args.steps, args.warmup)
    torch.cuda.synchronize()
    ne, nn, nnz = km.num_elems, km.num_nodes, plan.nnz
    value = ne / (ms_per * 1e-3)
    alg = algorithmic_bytes_newton(nn, ne, n, nnz)
    achieved = alg / (ms_kernel * 1e-3) / 1e9
    log(f"[bench] {args.steps} steps: {ms_per:.4f} ms/step (events), "
        f"{'K+J+R pass + M pass' if args.split_m else 'fused K+M+J+R'} "
        f"{ms_kernel:.4f} ms, {value:.4g} tets/s, {achieved:.0f} GB/s algorithmic")
    # the same set in FP64: its kernel time and the FP32-local error per output
    err = None
    if args.precision == "fp32_local":
        out32 = [t.clone() for t in (kv, mv, jv, rv)]
        prec = L.PRECISIONS["fp64"]
        _, ms64 = timed_steps(step, stream, args.steps, args.warmup)
        err = {"fp64_kernel_ms": ms64}
        for name, a32, a64 in zip("KMJR", out32, (kv, mv, jv, rv)):
            err[f"max_rel_err_{name}"] = float((a32 - a64).abs().max() / a64.abs().max())
        log(f"[bench] fp64 set {ms64:.4f} ms; fp32-local errors " +
            " ".join(f"{k[-1]} {v:.1e}" for k, v in err.items() if k.startswith("max")))
        del out32
    if args.calib:
        calib_copy(dev)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline on a {args.cpu_sample}-cube sample ...")
        cpu = cpu_baseline_newton(args.cpu_sample, args.cpu_reps)
        log(f"[bench] cpu: {cpu['value']:.4g} tets/s")
    workload = "kuhn100x100x67_newton_kmjr"
    line = {
        "metric": METRIC, "value": value, "unit": "tets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-local/f64-acc",
        "data": "synthetic (seeded jittered Kuhn box, Gaussian A scaled to max |B| = 2.2 T)",
        "config": {"workload": workload,
                   "operator": "K(nu(B)) + M_sigma + J + R, fused, FP64 CSC values",
                   "tets": ne, "nodes": nn, "edge_dofs": n, "nnz": int(nnz),
                   "precision": args.precision, "plan_build_ms": round(1e3 * (t3 - t2), 1)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(workload),
                     "algorithmic_bytes": alg, "kernel_ms": ms_kernel},
        "cpu_baseline": cpu,
        "fp32_local_vs_fp64": err,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def run_ecore(args, emit=True):
    """Nonlinear transient on the synthetic E-core: plain Newton (x_prev start,
    the reference's choice) against the learned-reluctivity warm start (and
    Newton from the extrapolated start), same mesh, source, tolerances and
    linear solver.  Steps 1..--train-steps run once with plain Newton (every
    method takes them the same way while the network is untrained); they
    train the network, and every other method continues from that state."""
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.learned import NuPredictor, learned_warm_start, training_collector
    from magpackage_amd.material import create_linear, marrocco
    from magpackage_amd.meshgen import ecore_box
    from magpackage_amd.transient import TransientSolver, divergence_free_source

    torch.cuda.set_device(0)
    ec = ecore_box(*args.ecore_cubes)
    mesh = Mesh.from_kuhn(ec.mesh)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    fixed = find_outer_boundary_dofs(mesh, dh, space)
    lib = {1: create_linear(1.0), 2: marrocco(), 3: create_linear(1.0)}
    sigma = {2: 2e6}
    js = torch.from_numpy(ec.js_dir).cuda()
    amp, freq, dt = 3e6, 50.0, args.dt
    if args.raw_source:
        src = lambda t: js * (amp * np.sin(2 * np.pi * freq * t))
    else:
        # the winding's load made discretely divergence-free (one nodal solve)
        t0 = time.perf_counter()
        f_unit = divergence_free_source(asm, space, asm.assembleSource(space, js), fixed)
        log(f"[ecore] divergence-free source: {time.perf_counter() - t0:.1f} s")
        src = lambda t: f_unit * (amp * np.sin(2 * np.pi * freq * t))
    steps = [dt] * args.steps
    ne, n = ec.mesh.num_elems, dh.NumGlobalDofs
    log(f"[ecore] {ne} tets, {n} edge DOFs, {int(fixed.sum())} fixed; {args.steps} steps of "
        f"{dt:g} s, train on steps 1-{args.train_steps}")

    k0 = args.train_steps
    w = 2 * np.pi * freq
    kinds = [k for k in ("plain", "extrapolated", "learned")
             if not (k == "extrapolated" and args.no_extrap_run) and not (args.only and k != args.only)]
    pred, coll = None, None
    if "learned" in kinds:
        pred = NuPredictor(asm, lib, seed=0, drive=lambda t: (np.sin(w * t), np.cos(w * t)))
        coll = training_collector(pred, space, k0, args.train_iters, log=log)

    def run(kind: str, resume=None, on_step=None, nsteps=None):
        solver = TransientSolver(asm)
        solver.Verbose = 2 if args.verbose_newton else True
        solver.NewtonLinearTolerance = args.newton_linear_tol
        if args.aux_sweeps:
            solver.AuxSweeps = args.aux_sweeps
        hook = None
        if kind == "extrapolated":
            solver.InitialGuess = "extrapolated"
        if kind == "learned":
            hook = learned_warm_start(pred, space, eta_warm=args.warm_eta)
            solver.InitialGuess = "extrapolated"  # extrapolation stays the fallback
        run_steps = steps[k0:] if resume is not None else steps[:nsteps]
        snap = {}

        def step_hook(ctx):
            if on_step is not None:
                on_step(ctx)
            if resume is None and ctx["step"] == k0:
                snap.update(solver.snapshot(ctx))

        t0 = time.perf_counter()
        _, infos = solver.solve(space, lib, sigma, src, run_steps, fixed, warm_start=hook,
                                on_step=step_hook, resume=resume)
        torch.cuda.synchronize()
        return infos, time.perf_counter() - t0, snap

    res = {"eval_steps": f"{k0 + 1}-{args.steps}"}
    # Steps 1..k0 are the same for every method (the learned start is not yet
    # trained): plain Newton runs them once, its converged steps 1..k0 train
    # the network, and the other methods continue from its state after step
    # k0, so that all of them are compared over steps k0+1.. from one state.
    snap = None
    for kind in ["plain"] + [k for k in kinds if k != "plain"]:
        log(f"[ecore] {kind}")
        if kind == "plain":
            collect = None
            if coll is not None:
                collect = lambda ctx: coll(ctx) if ctx["step"] <= k0 else None
            infos, total, sn = run("plain", on_step=collect,
                                   nsteps=None if "plain" in kinds else k0)
            snap = sn or None
            if snap is None and len(kinds) > 1:
                raise RuntimeError(f"the run has no step {k0} to continue from")
            if "plain" not in kinds:
                continue
            ev_infos = infos[k0:]
        else:
            infos, total, _ = run(kind, resume=snap,
                                  on_step=coll if kind == "learned" else None)
            ev_infos = infos
        ev = lambda key: float(np.mean([getattr(i, key) for i in ev_infos]))
        res[kind] = {"ms_per_step": ev("wall_ms"), "newton": ev("newton_iters"),
                     "pcg": ev("pcg_iters"), "total_s": total,
                     "converged": all(i.converged for i in infos),
                     "newton_per_step": [i.newton_iters for i in infos],
                     "ms_per_step_all": [round(i.wall_ms, 1) for i in infos]}
        if kind == "learned":
            res[kind]["warm_used"] = float(np.mean([i.warm_start == "warm" for i in ev_infos]))
            res[kind]["train"] = pred.train_info
    if "plain" not in res or "learned" not in res:
        if emit:
            print(json.dumps({"ecore": res}), flush=True)
        return res
    plain, learn = res["plain"], res["learned"]
    speed = plain["ms_per_step"] / learn["ms_per_step"]
    for kind in ("plain", "extrapolated", "learned"):
        if kind in res:
            r = res[kind]
            log(f"[ecore] steps {k0 + 1}-{args.steps} {kind}: {r['ms_per_step']:.1f} ms/step, "
                f"{r['newton']:.2f} Newton, {r['pcg']:.0f} PCG iterations")
    log(f"[ecore] learned vs plain: {speed:.2f}x")
    res["speedup_learned_vs_plain"] = speed
    if "extrapolated" in res:
        # what the network adds over the extrapolated start it corrects
        res["speedup_learned_vs_extrapolated"] = (res["extrapolated"]["ms_per_step"]
                                                  / learn["ms_per_step"])
        res["speedup_extrapolated_vs_plain"] = (plain["ms_per_step"]
                                                / res["extrapolated"]["ms_per_step"])
        log(f"[ecore] learned vs extrapolated: {res['speedup_learned_vs_extrapolated']:.2f}x")
    res["config"] = {"workload": "ecore_transient_learned_vs_plain", "tets": ne, "edge_dofs": n,
                     "dt": dt, "steps": args.steps, "train_steps": k0, "rel_tol": 1e-6,
                     "newton_linear_solve": ("eisenstat-walker forcing"
                                             if args.newton_linear_tol is None else
                                             f"relative residual {args.newton_linear_tol:g}"),
                     "source": "raw" if args.raw_source else "divergence-free projection"}
    if emit:
        print(json.dumps({
            "metric": METRIC, "value": speed, "unit": "x (plain / learned time per step)",
            "n_gpus": 1, "steps": args.steps, "warmup": 0,
            "ms_per_step": learn["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 (MLP bf16/f32)",
            "data": "synthetic E-core (jittered Kuhn box), Marrocco iron, 50 Hz winding",
            "config": res["config"], "nonlinear": res,
        }), flush=True)
    return res


def run_batch(args):
    """256 jittered Kuhn boxes (tets log-uniform on [5e3, 2e5], seed 4) x 32
    frequencies (10 Hz - 10 kHz, log-spaced): K + 1j*w*M_sigma for all of them
    in one HIP graph launch, schedule from the autotuner; per-mesh latency
    from one launch per mesh alone."""
    import torch
    from magpackage_amd.batch import BatchAssembler, log_uniform_meshes
    from magpackage_amd.material import MU0
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    meshes = log_uniform_meshes(args.batch_meshes, seed=4)
    nus = [np.full(m.num_elems, 1.0 / MU0) for m in meshes]
    sigmas = [np.where(m.RegionTags == 2, 3.5e7, 0.0) for m in meshes]
    omegas = 2 * np.pi * np.logspace(1, 4, args.batch_freqs)
    ba = BatchAssembler(meshes, nus, sigmas, omegas, precision=args.precision or "fp64")
    torch.cuda.synchronize()
    ne = sum(m.num_elems for m in meshes)
    nn = sum(m.num_nodes for m in meshes)
    log(f"[batch] {len(meshes)} meshes, {ne} tets, nnz {ba.nnz_total}, {ba.ntiles} tiles, "
        f"{args.batch_freqs} frequencies, outputs {16 * args.batch_freqs * ba.nnz_total / 1e9:.1f}"
        f" GB; set-up {time.perf_counter() - t0:.1f} s")
    rec = ba.autotune(reps=2, log=log)
    log(f"[batch] autotuner picked ring batch {ba.schedule[0]}, blocks/CU {ba.schedule[1]}")
    ba.capture()
    stream = torch.cuda.current_stream()

    def step(ev_a=None, ev_b=None):
        if ev_a is not None:
            ev_a.record(stream)
        ba.replay()
        if ev_b is not None:
            ev_b.record(stream)

    ms_per, ms_kernel = timed_steps(step, stream, args.steps, args.warmup)
    lat = ba.per_mesh_latency(reps=2)
    alg = sum(24 * m.num_nodes + (16 + 24 + 8 + 8) * m.num_elems for m in meshes) + \
        16 * args.batch_freqs * ba.nnz_total
    achieved = alg / (ms_kernel * 1e-3) / 1e9
    value = ne / (ms_per * 1e-3)
    log(f"[batch] {ms_per:.3f} ms per batch launch: {value:.4g} tets/s "
        f"({value * args.batch_freqs:.4g} tet-frequencies/s), {achieved:.0f} GB/s algorithmic; "
        f"per-mesh latency p50 {np.percentile(lat, 50):.3f} ms, p99 {np.percentile(lat, 99):.3f} ms")
    print(json.dumps({
        "metric": METRIC, "value": value, "unit": "tets/s (each tet x 32 frequencies)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if (args.precision or "fp64") == "fp64" else "f32-local/f64-acc",
        "data": "synthetic (seeded jittered Kuhn boxes)",
        "config": {"workload": "batch256_freq32", "meshes": len(meshes), "tets": ne,
                   "nodes": nn, "nnz_total": int(ba.nnz_total), "frequencies": args.batch_freqs,
                   "schedule": list(ba.schedule), "autotune_ms": rec["ms"],
                   "per_mesh_latency_ms": {"p50": float(np.percentile(lat, 50)),
                                           "p99": float(np.percentile(lat, 99)),
                                           "max": float(lat.max())}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": alg,
                     "kernel_ms": ms_kernel},
    }), flush=True)


def run_mlp(args):
    """MFMA throughput of the reluctivity MLP kernels: inference on 1M rows and
    one training step (loss + backward) on a batch, HIP events around each
    launch on the launch stream."""
    import torch
    from magpackage_amd import _lib as L
    from magpackage_amd.learned import NPARAM, _glorot
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    p = torch.from_numpy(_glorot(0)).to(dev)
    wb = torch.empty(9216, dtype=torch.int16, device=dev)
    L.check(L.lib.mpk_mlp_sync_bf16(L.ptr(p), L.ptr(wb), L.stream_ptr()))
    stream = torch.cuda.current_stream()
    sp = L.stream_ptr(stream)
    out = {}
    # MFMA flops per row: 2 (16*64 + 64*64 + 64*64); backward: dH (2 layers)
    # and dW (3 layers) products, 2 (2*64*64 + 16*64 + 2*64*64)
    f_fwd = 2 * (16 * 64 + 2 * 64 * 64)
    f_bwd = 2 * (4 * 64 * 64 + 16 * 64)
    for name, n in (("forward", 1 << 20), ("loss_grad", 8192), ("loss_grad_1M", 1 << 20)):
        X = torch.randn(n, 16, device=dev, dtype=torch.float32)
        t = torch.randn(n, device=dev, dtype=torch.float32)
        y = torch.empty(n, device=dev, dtype=torch.float32)
        g = torch.empty(NPARAM, device=dev, dtype=torch.float32)
        loss = torch.empty(1, device=dev, dtype=torch.float32)
        ws = C.c_int64()
        L.check(L.lib.mpk_mlp_work_size(n, C.byref(ws)))
        work = torch.empty(ws.value, device=dev, dtype=torch.float32)

        def step(ev_a=None, ev_b=None):
            if ev_a is not None:
                ev_a.record(stream)
            if name == "forward":
                L.check(L.lib.mpk_mlp_forward(L.ptr(p), L.ptr(wb), L.ptr(X), n, L.ptr(y), sp))
            else:
                L.check(L.lib.mpk_mlp_loss_grad(L.ptr(p), L.ptr(wb), L.ptr(X), L.ptr(t), n,
                                                L.ptr(g), L.ptr(loss), L.ptr(work), sp))
            if ev_b is not None:
                ev_b.record(stream)

        _, ms = timed_steps(step, stream, args.steps, args.warmup)
        flops = n * (f_fwd if name == "forward" else f_fwd + f_bwd)
        tf = flops / (ms * 1e-3) / 1e12
        out[name] = {"rows": n, "ms": ms, "mfma_tflops": tf, "frac_of_2500": tf / 2500.0}
        log(f"[mlp] {name}: {n} rows, {ms:.4f} ms, {tf:.2f} TFLOP/s on MFMA "
            f"({100 * tf / 2500:.2f}% of 2.5 PFLOP/s dense bf16)")
    print(json.dumps({"mlp": out}), flush=True)


def timed_steps(step, stream, steps, warmup):
    """Runs warmup + timed steps; returns (ms per step over the timed region,
    mean ms of the bracketed kernel launches)."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    eva = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    evb = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ev0.record(stream)
    for i in range(steps):
        step(eva[i], evb[i])
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_per = ev0.elapsed_time(ev1) / steps
    ms_kernel = float(np.mean([x.elapsed_time(y) for x, y in zip(eva, evb)]))
    return ms_per, ms_kernel


def calib_copy(dev):
    import torch
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    log("[bench] calibration copy: 3 x 512 MiB read + 512 MiB write")


def cpu_baseline(sample_n: int, reps: int):
    """Reference path on the host: MEX-order COO kernels (oracle, OpenMP) +
    sparse() for K and M_sigma + K + 1j*omega*M on a bounded sample, the
    sparse() and the sum on the oracle's threads as well."""
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
        t1 = time.perf_counter()
        K = O.sparse_par(I, J, V, n, n)
        del I, J, V
        t2 = time.perf_counter()
        I, J, V = O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sigma, qw, val)
        t3 = time.perf_counter()
        M = O.sparse_par(I, J, V, n, n)
        del I, J, V
        A = O.csc_add_jw(K, M, omega, n)
        t4 = time.perf_counter()
        return t4 - t0, (t1 - t0) + (t3 - t2), (t2 - t1) + (t4 - t3)

    run()  # warm-up
    runs = [run() for _ in range(reps)]
    t = float(np.median([r[0] for r in runs]))
    tk = float(np.median([r[1] for r in runs]))
    ts = float(np.median([r[2] for r in runs]))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": km.num_elems / t,
        "unit": "tets/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"Kuhn {sample_n}^3 box ({km.num_elems} tets, same plate/sigma/50 Hz), "
                   f"COO kernels, sparse() and K+1j*w*M on {cores} OpenMP threads, "
                   f"median of {reps} after 1 warm-up: {t:.3f} s (COO kernels {tk:.3f} s, "
                   f"sparse() and K+1j*w*M {ts:.3f} s)"),
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
    ap.add_argument("--workload", default="freq128", choices=["freq128", "newton4m", "ecore", "mlp", "batch"],
                    help="freq128: K + 1j*w*M_sigma complex128 on the 128^3 box (default); "
                         "newton4m: fused K(nu(B)) + M_sigma + J + R on the 100x100x67 box")
    ap.add_argument("--n", type=int, default=128, help="cubes per axis of the Kuhn box")
    ap.add_argument("--dt", type=float, default=1e-4, help="ecore: time step (s)")
    ap.add_argument("--train-steps", type=int, default=40, help="ecore: steps used for training")
    ap.add_argument("--train-iters", type=int, default=20000, help="ecore: Adam iterations")
    ap.add_argument("--ecore-cubes", type=int, nargs=3, default=[64, 64, 40],
                    help="ecore: cubes per axis (64 64 40: 983,040 tets)")
    ap.add_argument("--batch-meshes", type=int, default=256, help="batch: number of meshes")
    ap.add_argument("--batch-freqs", type=int, default=32, help="batch: number of frequencies")
    ap.add_argument("--warm-eta", type=float, default=1e-2,
                    help="ecore: relative tolerance of the learned warm start's solve")
    ap.add_argument("--newton-linear-tol", type=float, default=None,
                    help="ecore: solve every Newton step's linear system to this relative "
                         "residual (1e-10: the reference's exact direct-solve regime) instead "
                         "of the Eisenstat-Walker forcing")
    ap.add_argument("--aux-sweeps", type=int, default=None,
                    help="ecore: damped Jacobi sweeps of the nodal correction in the PCG "
                         "preconditioner (default: the solver's, 2)")
    ap.add_argument("--no-nonlinear", action="store_true",
                    help="freq128: skip the learned-vs-plain nonlinear transients")
    ap.add_argument("--no-inexact", action="store_true",
                    help="freq128: skip the inexact-Newton learned-vs-plain transient")
    ap.add_argument("--no-newton", action="store_true",
                    help="freq128: skip the fused Newton set on the 100x100x67 box")
    ap.add_argument("--raw-source", action="store_true",
                    help="ecore: use the per-element winding current as is (not projected "
                         "onto divergence-free loads)")
    ap.add_argument("--verbose-newton", action="store_true",
                    help="ecore: print each step's residual history and PCG counts")
    ap.add_argument("--only", default=None, choices=["plain", "extrapolated", "learned"],
                    help="ecore: run one method only")
    ap.add_argument("--no-extrap-run", action="store_true",
                    help="ecore: skip the plain-Newton-from-extrapolation reference run")
    ap.add_argument("--split-m", action="store_true",
                    help="newton4m: K, J, R in one pass and M_sigma in a second one")
    ap.add_argument("--precision", default=None, choices=["fp64", "fp64_exact", "fp32_local"],
                    help="default fp64 (freq128) / fp32_local (newton4m)")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="cube count of the CPU baseline's box (default: the workload's own "
                         "size, 128 for freq128 and 100 for newton4m)")
    ap.add_argument("--cpu-reps", type=int, default=5)
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
    if args.workload == "mlp":
        return run_mlp(args)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "ecore":
        if args.steps == 20:
            args.steps = 200
        return run_ecore(args)
    if args.workload == "newton4m":
        args.precision = args.precision or "fp32_local"
        if args.precision == "fp64_exact":
            ap.error("the fused Newton set has fp64 and fp32_local modes")
        if args.cpu_sample is None:
            args.cpu_sample = 100
        return run_newton(args)
    args.precision = args.precision or "fp64"
    if args.cpu_sample is None:
        args.cpu_sample = args.n

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

    ms_per, ms_kernel = timed_steps(step, stream, args.steps, args.warmup)
    torch.cuda.synchronize()
    ne, nn, nnz = km.num_elems, km.num_nodes, plan.nnz
    value = ne / (ms_per * 1e-3)
    alg = algorithmic_bytes_freq(nn, ne, nnz)
    achieved = alg / (ms_kernel * 1e-3) / 1e9
    log(f"[bench] {args.steps} steps: {ms_per:.4f} ms/step (events), assembly kernel "
        f"{ms_kernel:.4f} ms, {value:.4g} tets/s, "
        f"{achieved:.0f} GB/s algorithmic")
    # the FP32-local mode of the same operator, timed the same way, with its
    # error against the FP64 values just computed
    fp32 = None
    if args.precision == "fp64":
        ref = vals.clone()
        prec = L.PRECISIONS["fp32_local"]
        ms32_step, ms32 = timed_steps(step, stream, args.steps, args.warmup)
        torch.cuda.synchronize()
        v32, v64 = vals.view(-1, 2), ref.view(-1, 2)
        errs = [float((v32[:, k] - v64[:, k]).abs().max() / v64[:, k].abs().max())
                for k in (0, 1)]
        ach32 = alg / (ms32 * 1e-3) / 1e9
        fp32 = {"ms_per_step": ms32_step, "kernel_ms": ms32, "value": ne / (ms32_step * 1e-3),
                "achieved_gbs": ach32, "frac": ach32 / HBM_PEAK_GBS,
                "max_rel_err_re": errs[0], "max_rel_err_im": errs[1]}
        log(f"[bench] fp32-local: kernel {ms32:.4f} ms ({ach32:.0f} GB/s), max rel err "
            f"Re {errs[0]:.1e} Im {errs[1]:.1e}")
        del ref, v32, v64

    if args.calib:
        calib_copy(dev)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline on a {args.cpu_sample}^3 sample ...")
        cpu = cpu_baseline(args.cpu_sample, args.cpu_reps)
        log(f"[bench] cpu: {cpu['value']:.4g} tets/s")

    ndofs = int(dh.NumGlobalDofs)
    del vals, nu_d, sg_d, plan, asm, dh, mesh, km
    torch.cuda.empty_cache()

    newton = None
    if not args.no_newton:
        # the fused K + M + J + R Newton set on the 100x100x67 box (FP32-local,
        # with its FP64 time and error), roofline, traffic and CPU baseline
        na = argparse.Namespace(**vars(args))
        na.precision, na.calib = "fp32_local", False
        na.cpu_sample = 100
        newton = run_newton(na, emit=False)
        for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline"):
            newton.pop(k, None)
        torch.cuda.empty_cache()

    nonlinear = None
    if not args.no_nonlinear:
        # learned-vs-plain nonlinear transient as specified: 200 backward-Euler
        # steps of 1 ms (ten 50 Hz periods), the network trained on steps
        # 1-40, plain Newton and the learned warm start compared over steps
        # 41-200, every Newton step's linear system solved to 1e-10 (the
        # reference solves it directly, with MUMPS)
        ea = argparse.Namespace(**vars(args))
        ea.steps, ea.train_steps, ea.dt = 200, 40, 1e-3
        ea.only, ea.no_extrap_run, ea.verbose_newton = None, True, False
        ea.newton_linear_tol = 1e-10
        r = run_ecore(ea, emit=False)
        nonlinear = {k: r[k] for k in ("eval_steps", "speedup_learned_vs_plain", "config")}
        for kind in ("plain", "learned"):
            nonlinear[kind] = {k: r[kind][k] for k in ("ms_per_step", "newton", "pcg", "converged")}
        nonlinear["learned"]["train"] = r["learned"]["train"]
        nonlinear["learned"]["warm_used"] = r["learned"]["warm_used"]
        torch.cuda.empty_cache()

    inexact = None
    if not args.no_nonlinear and not args.no_inexact:
        # the same transient with the engine's inexact Newton (each linear
        # solve only as accurate as the Eisenstat-Walker forcing asks), 80
        # steps: plain Newton, Newton from the extrapolated start and the
        # learned warm start over steps 41-80 (learned vs extrapolated is the
        # network's own gain)
        ea = argparse.Namespace(**vars(args))
        ea.steps, ea.train_steps, ea.dt = 80, 40, 1e-3
        ea.only, ea.no_extrap_run, ea.verbose_newton = None, False, False
        ea.newton_linear_tol = None
        r = run_ecore(ea, emit=False)
        inexact = {k: r[k] for k in ("eval_steps", "speedup_learned_vs_plain",
                                     "speedup_learned_vs_extrapolated",
                                     "speedup_extrapolated_vs_plain", "config")}
        for kind in ("plain", "extrapolated", "learned"):
            inexact[kind] = {k: r[kind][k] for k in ("ms_per_step", "newton", "pcg", "converged")}
        inexact["learned"]["warm_used"] = r["learned"]["warm_used"]

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
            "tets": ne, "nodes": nn, "edge_dofs": ndofs, "nnz": int(nnz),
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
        "fp32_local": fp32,
        "newton_set": newton,
        "nonlinear_step_time": nonlinear,
        "nonlinear_step_time_inexact_newton": inexact,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
