"""The learned warm start (magpackage_amd.learned) on a small E-core transient:
what the trained network gains over the methods without it.

Three runs of the same backward-Euler transient (Marrocco iron with
sigma = 2e6 S/m, a 50 Hz winding, dt = 1 ms, the same tolerances and linear
solver):
- plain: Newton from x_prev, the reference's choice (TransientSolver.m:159)
  with its Newton loop (:164-259);
- extrapolated: Newton from 2 x_prev - x_prev2;
- learned: the reluctivity network, trained on the GPU by the engine's own
  MLP kernels on the first TRAIN steps of this run, supplies the warm start
  (extrapolation remains the fallback, the residual picks).
Over the steps after training, the learned run takes fewer Newton iterations
per step than both other runs, and its converged fields agree with plain
Newton's to the Newton tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 40
TRAIN = 20


def _run(kind):
    import torch
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.boundary import find_outer_boundary_dofs
    from magpackage_amd.learned import NuPredictor, learned_warm_start, training_collector
    from magpackage_amd.material import create_linear, marrocco
    from magpackage_amd.meshgen import ecore_box
    from magpackage_amd.transient import TransientSolver, divergence_free_source
    ec = ecore_box(24, 24, 16)
    mesh = Mesh.from_kuhn(ec.mesh)
    dh = DofHandler(mesh)
    space = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(space)
    asm = Assembler(mesh, dh)
    fixed = find_outer_boundary_dofs(mesh, dh, space)
    lib = {1: create_linear(1.0), 2: marrocco(), 3: create_linear(1.0)}
    js = torch.from_numpy(ec.js_dir).cuda()
    f_unit = divergence_free_source(asm, space, asm.assembleSource(space, js), fixed)
    w = 2 * np.pi * 50.0
    src = lambda t: f_unit * (3e6 * np.sin(w * t))
    solver = TransientSolver(asm)
    solver.Verbose = True  # one line per step (progress)
    print(f"[learned test] {kind}: {ec.mesh.num_elems} tets", flush=True)
    hook = on_step = None
    if kind == "extrapolated":
        solver.InitialGuess = "extrapolated"
    if kind == "learned":
        pred = NuPredictor(asm, lib, seed=0, drive=lambda t: (np.sin(w * t), np.cos(w * t)))
        hook = learned_warm_start(pred, space)
        on_step = training_collector(pred, space, TRAIN, train_iters=6000)
        solver.InitialGuess = "extrapolated"
    sols, infos = solver.solve(space, lib, {2: 2e6}, src, [1e-3] * STEPS, fixed,
                               warm_start=hook, on_step=on_step)
    B = [asm.computeElementB(x, space) for x in sols[TRAIN:]]
    return infos, B


@pytest.fixture(scope="module")
def runs(gpu):
    return {k: _run(k) for k in ("plain", "extrapolated", "learned")}


def test_learned_warm_start_gains(runs):
    newton = {}
    for kind, (infos, _) in runs.items():
        assert all(i.converged for i in infos), kind
        ev = infos[TRAIN:]
        newton[kind] = float(np.mean([i.newton_iters for i in ev]))
        pcg = float(np.mean([i.pcg_iters for i in ev]))
        print(f"{kind}: steps {TRAIN + 1}-{STEPS}: {newton[kind]:.2f} Newton, {pcg:.0f} PCG "
              f"per step; Newton per step {[i.newton_iters for i in ev]}", flush=True)
    used = np.mean([i.warm_start == "warm" for i in runs["learned"][0][TRAIN:]])
    print(f"learned: warm start chosen on {100 * used:.0f}% of the steps", flush=True)
    assert used >= 0.5
    assert newton["learned"] < newton["extrapolated"]
    assert newton["learned"] < newton["plain"]


def test_learned_matches_plain_fields(runs):
    """Converged B of the learned run against plain Newton's, step by step:
    both stop at a residual 1e-6 below the step's first (plain) residual."""
    import torch
    worst = 0.0
    for bl, bp in zip(runs["learned"][1], runs["plain"][1]):
        worst = max(worst, float((bl - bp).abs().max() / bp.abs().max()))
    print(f"max |B_learned - B_plain| / max |B_plain| = {worst:.1e}", flush=True)
    assert worst <= 1e-5
    for bl, bp in zip(runs["extrapolated"][1], runs["plain"][1]):
        assert float((bl - bp).abs().max() / bp.abs().max()) <= 1e-5
    assert torch.isfinite(runs["learned"][1][-1]).all()
