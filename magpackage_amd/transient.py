"""Backward-Euler nonlinear transient driver on the GPU engine.

Mirror of TransientSolver.solve (src/solver/TransientSolver.m:38-267,
274-311), pure-A formulation (no scalar potential space): per time step
with source F(t),

    R(x) = R_nu(x) + M_sigma (x - x_prev) / dt - F + reg * M_reg * x
    J(x) = J_nu(x) + M_sigma / dt + reg * M_reg

with the regularisation scale reg = RegScaleRel * mean|diag(J_nu + M_sigma/dt)|
(positive entries) fixed at the first Newton iteration of the first step
(:150-158), Dirichlet DOFs removed as BoundaryCondition.applyDirichlet does
(:170), convergence when |R| < Tolerance or |R| / |R_first| < RelTolerance
(:173-195), and the reference's backtracking line search (:207-237).

Every matrix lives on the plan's structural CSC pattern on the device:
J_nu and R_nu come from mpk_csc_jacobian, M_sigma and M_reg from
mpk_csc_mass (once), the source from mpk_csc_source; the linear solve is the
engine's Jacobi-PCG on the Dirichlet-applied operator (the reference's MUMPS
direct solver is not available), with an inexact-Newton forcing term.

A ``warm_start`` hook may supply the first Newton iterate of a step (the
learned reluctivity predictor does); the plain iterate x_prev stays the
fallback and the residual decides.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .assembler import Assembler, expand_material_map


def divergence_free_source(assembler: Assembler, space, F, fixed, tol: float = 1e-10):
    """Removes the discrete-gradient part of an assembled source vector:
    F_df = F - M G phi with (G^T M G) phi = G^T F over the free nodes, M the
    unweighted edge mass and G the edge-node incidence (edge (a < b):
    phi_b - phi_a).  Then G^T F_df = 0: the source is the load of a
    divergence-free current, as a winding's current is.  A current density
    sampled per element (a rectangular loop's direction field) is not
    discretely divergence-free; the reference's direct solver absorbs the
    gradient part through the regularisation, an iterative solver stalls on
    it.  Set-up work on the host (scipy CG on the nodal system)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    mesh = assembler.Mesh
    plan = assembler.plan(space)
    m = assembler.assembleMass(space, drop_zeros=False)
    M = m.to_scipy()
    edges = mesh.Edges.cpu().numpy().astype(np.int64)
    emap = assembler.DofHandler.EntityMaps[space.toString()].cpu().numpy()
    n = plan.n
    fixed = np.asarray(fixed, dtype=bool)
    act = emap > 0
    rows = emap[act] - 1
    E = edges[act]
    G = sp.csr_matrix((np.tile([-1.0, 1.0], len(rows)), (np.repeat(rows, 2), (E - 1).ravel())),
                      shape=(n, mesh.NumNodes))
    bnodes = np.unique((E[fixed[rows]] - 1).ravel())
    keep = np.setdiff1d(np.arange(mesh.NumNodes), bnodes)
    Gi = G[:, keep]
    f = F.detach().cpu().numpy() if torch.is_tensor(F) else np.asarray(F, dtype=np.float64)
    f = np.where(fixed, 0.0, f)
    Ln = (Gi.T @ M @ Gi).tocsr()
    rhs = Gi.T @ f
    phi, info = spla.cg(Ln, rhs, rtol=tol, maxiter=20000, M=sp.diags(1.0 / Ln.diagonal()))
    if info != 0:
        raise RuntimeError(f"nodal projection did not converge (info {info})")
    fdf = f - M @ (Gi @ phi)
    fdf[fixed] = 0.0
    return torch.from_numpy(fdf).to(torch.device("cuda", torch.cuda.current_device()))


def build_aux_space(assembler: Assembler, space_key: str, fixed, c_vals, colptr, rowind,
                    sweeps: int = 2, omega: float = 2.0 / 3.0):
    """Gradient space for mpk_pcg_aux_solve: G over the free nodes (nodes not on
    a fixed edge) and Ln = G^T C G, C the system's mass part on the plan's
    pattern (c_vals; the curl part of the system vanishes on gradients).  Host
    set-up with scipy.  Returns (mpk_aux_space struct, device arrays it points
    to, which the caller keeps alive)."""
    import scipy.sparse as sp
    mesh = assembler.Mesh
    edges = mesh.Edges.cpu().numpy().astype(np.int64)
    emap = assembler.DofHandler.EntityMaps[space_key].cpu().numpy().astype(np.int64)
    fixed = (fixed.cpu().numpy() if torch.is_tensor(fixed) else np.asarray(fixed)).astype(bool)
    n = fixed.size
    act = emap > 0
    rows = emap[act] - 1
    E = edges[act] - 1
    dnodes = np.zeros(mesh.NumNodes, dtype=bool)
    dnodes[E[fixed[rows]].ravel()] = True
    free = np.nonzero(~dnodes)[0]
    free_id = np.full(mesh.NumNodes, -1, dtype=np.int64)
    free_id[free] = np.arange(free.size)
    nv = free.size
    en = np.full((n, 2), -1, dtype=np.int32)
    en[rows, 0] = free_id[E[:, 0]]
    en[rows, 1] = free_id[E[:, 1]]
    ma, mb = en[rows, 0] >= 0, en[rows, 1] >= 0
    G = sp.csr_matrix((np.concatenate([-np.ones(int(ma.sum())), np.ones(int(mb.sum()))]),
                       (np.concatenate([rows[ma], rows[mb]]),
                        np.concatenate([en[rows, 0][ma], en[rows, 1][mb]]))), shape=(n, nv))
    Gc = G.tocsc()
    Gc.sort_indices()
    refs = (Gc.indices.astype(np.int64) << 1) | (Gc.data > 0)
    Cm = sp.csc_matrix((c_vals.cpu().numpy(), rowind.cpu().numpy(), colptr.cpu().numpy()),
                       shape=(n, n))
    Ln = (G.T @ Cm @ G).tocsr()
    Ln.sort_indices()
    dev = c_vals.device
    keep = dict(
        en=torch.from_numpy(en.reshape(-1)).to(dev),
        nptr=torch.from_numpy(Gc.indptr.astype(np.int32)).to(dev),
        nref=torch.from_numpy(refs.astype(np.int32)).to(dev),
        lp=torch.from_numpy(Ln.indptr.astype(np.int32)).to(dev),
        li=torch.from_numpy(Ln.indices.astype(np.int32)).to(dev),
        lv=torch.from_numpy(Ln.data.astype(np.float64)).to(dev))
    aux = L.AuxSpace()
    aux.num_nodes = nv
    aux.edge_nodes = L.ptr(keep["en"])
    aux.node_ptr = L.ptr(keep["nptr"])
    aux.node_refs = L.ptr(keep["nref"])
    aux.ln_ptr = L.ptr(keep["lp"])
    aux.ln_idx = L.ptr(keep["li"])
    aux.ln_vals = L.ptr(keep["lv"])
    aux.sweeps = int(sweeps)
    aux.omega = float(omega)
    ws = C.c_int64()
    L.check(L.lib.mpk_pcg_aux_work_size(n, nv, C.byref(ws)))
    keep["work"] = torch.empty(ws.value, dtype=torch.float64, device=dev)
    return aux, keep


@dataclass
class StepInfo:
    step: int
    time: float
    newton_iters: int
    pcg_iters: int
    line_search_evals: int
    residual: float
    rel_residual: float
    converged: bool
    wall_ms: float
    warm_start: str = "x_prev"
    history: list = field(default_factory=list)
    pcg_per_solve: list = field(default_factory=list)


class TransientSolver:
    """TransientSolver with the reference's properties and defaults
    (src/solver/TransientSolver.m:10-27)."""

    def __init__(self, assembler: Assembler):
        self.Assembler = assembler
        self.Dt = 0.01
        self.MaxIterations = 50
        self.Tolerance = 1e-5
        self.RelTolerance = 1e-6
        self.UseLineSearch = True
        self.MaxLineSearchIters = 10
        self.Relaxation = 1.0
        self.AutoRegularization = True
        self.RegScaleRel = 1e-5
        # linear solver (engine): Jacobi-PCG with forcing term
        # eta_k = min(max(min(eta_max, gamma (|R_k| / |R_k-1|)^2),
        #                 eta_floor * target / |R_k|), 0.5)
        self.LinearEtaMax = 1e-2
        self.LinearGamma = 0.9
        self.LinearEtaFloor = 0.5
        # None: the forcing above (inexact Newton).  A number: every Newton
        # step's linear system is solved to that relative residual, as the
        # reference's direct solve (MUMPS) solves it exactly
        self.NewtonLinearTolerance = None
        self.LinearMaxIter = 20000
        self.LinearCheckEvery = 10
        # solveAV: relative CG tolerance of the A-V Newton solves
        self.LinearRelTolAV = 1e-12
        # "aux": Jacobi + gradient-space correction (mpk_pcg_aux_solve);
        # "jacobi": plain Jacobi-PCG
        self.LinearPreconditioner = "aux"
        # nodal correction p(Ln) = AuxSweeps damped Jacobi sweeps; one sweep is
        # the diagonal omega Dn^-1, with no Ln product per iteration (983k-tet
        # E-core, PCG to 1e-10: 8200 iterations per step against 7880 with two
        # sweeps, 0.99 s against 1.07 s per step)
        self.AuxSweeps = 1
        self.AuxOmega = 2.0 / 3.0
        self.Precision = "fp64"
        # first Newton iterate of a step: "previous" (x_prev, the reference's
        # choice, :159) or "extrapolated" (2 x_prev - x_prev2)
        self.InitialGuess = "previous"
        self.Verbose = False

    # ------------------------------------------------------------------ setup
    def _setup(self, space, matLib, sigmaMap, fixedDofs):
        asm = self.Assembler
        self._space_key = space.toString()
        dev = torch.device("cuda", torch.cuda.current_device())
        plan = asm.plan(space)
        self.plan = plan
        self.n = n = plan.n
        self.dev = dev
        self.sp = L.stream_ptr()
        mesh = asm.Mesh
        L.check(L.lib.mpk_plan_set_coords(plan.handle, L.ptr(mesh.P_dev), self.sp))
        self.prec = L.PRECISIONS[self.Precision]
        self.mat, self._mat_keep = asm.material_tables(matLib)
        ntag = int(self.mat.num_tags)
        if int(mesh.RegionTags.min()) < 1 or int(mesh.RegionTags.max()) > ntag:
            raise L.MpkError(L.MPK_ERR_INDEX, "region tags outside material table")
        self.tags = mesh.Tags_dev
        self.qk = asm.stiffness_quadrature()
        sigma = expand_material_map(sigmaMap, mesh.RegionTags)
        self.sigma = torch.from_numpy(np.ascontiguousarray(sigma)).to(dev)
        self.m_sigma = torch.empty(max(plan.nnz, 1), dtype=torch.float64, device=dev)
        qm = asm.mass_quadrature(weighted=True)
        L.check(L.lib.mpk_csc_mass(plan.handle, None, L.ptr(self.sigma), C.byref(qm), self.prec,
                                   L.ptr(self.m_sigma), self.sp))
        self.m_reg = None
        if self.AutoRegularization:
            self.m_reg = torch.empty_like(self.m_sigma)
            qr = asm.mass_quadrature(weighted=False)
            L.check(L.lib.mpk_csc_mass(plan.handle, None, None, C.byref(qr), self.prec,
                                       L.ptr(self.m_reg), self.sp))
        colptr, rowind = plan.pattern()
        self.colptr = colptr.to(torch.int32).contiguous()
        self.rowind = rowind.contiguous()
        cols = torch.repeat_interleave(torch.arange(n, device=dev),
                                       (colptr[1:] - colptr[:-1]).to(torch.int64))
        self.diag_pos = torch.nonzero(self.rowind.to(torch.int64) == cols).flatten()
        fixed = np.zeros(n, dtype=np.uint8)
        if fixedDofs is not None:
            fx = np.asarray(fixedDofs)
            if fx.dtype == bool:
                fixed[: min(len(fx), n)] = fx[:n]
            else:
                idx = fx.astype(np.int64).ravel()
                idx = idx[(idx >= 1) & (idx <= n)]
                fixed[idx - 1] = 1
        self.fixed = torch.from_numpy(fixed).to(dev)
        self.free = (self.fixed == 0).to(torch.float64)
        ws = C.c_int64()
        L.check(L.lib.mpk_pcg_work_size(n, C.byref(ws)))
        self.work = torch.empty(ws.value, dtype=torch.float64, device=dev)
        self.jv = torch.empty_like(self.m_sigma)
        self.jsys = torch.empty_like(self.m_sigma)
        self.cvals = torch.empty_like(self.m_sigma)
        self.res = torch.empty(n, dtype=torch.float64, device=dev)
        self.tmp = torch.empty(n, dtype=torch.float64, device=dev)
        self.reg_abs = 0.0
        self._c_key = None
        self._aux_key = None
        self._aux = None
        self.timers = {}

    # -------------------------------------------------------------- kernels
    def _tic(self):
        if self.Verbose > 1:
            torch.cuda.synchronize()
        return time.perf_counter()

    def _toc(self, name, t0):
        if self.Verbose > 1:
            torch.cuda.synchronize()
            self.timers[name] = self.timers.get(name, 0.0) + 1e3 * (time.perf_counter() - t0)

    def _jacobian(self, x, calc_j: bool):
        L.check(L.lib.mpk_csc_jacobian(self.plan.handle, None, L.ptr(self.tags), L.ptr(x),
                                       C.byref(self.mat), C.byref(self.qk), int(calc_j),
                                       self.prec, L.ptr(self.jv) if calc_j else None,
                                       L.ptr(self.res), self.sp))

    def _spmv(self, vals, x, y):
        L.check(L.lib.mpk_spmv_sym(L.ptr(self.colptr), L.ptr(self.rowind), L.ptr(vals),
                                   self.n, L.ptr(x), L.ptr(y), self.sp))

    def _combine(self, out, c0, v0, c1=0.0, v1=None, c2=0.0, v2=None):
        L.check(L.lib.mpk_csc_combine(self.plan.nnz, c0, L.ptr(v0), c1, L.ptr(v1), c2, L.ptr(v2),
                                      0.0, None, L.ptr(out), self.sp))

    def _constant_part(self, dt):
        """C = M_sigma / dt + reg * M_reg (the x-linear part of R and J)."""
        key = (dt, self.reg_abs)
        if key != self._c_key:
            if self.m_reg is not None and self.reg_abs > 0:
                self._combine(self.cvals, 1.0 / dt, self.m_sigma, self.reg_abs, self.m_reg)
            else:
                self._combine(self.cvals, 1.0 / dt, self.m_sigma)
            self._c_key = key

    def residual(self, x, b, dt):
        """Dirichlet-masked R(x) = R_nu(x) + C x - b and its norm."""
        t0 = self._tic()
        self._jacobian(x, calc_j=False)
        self._constant_part(dt)
        self._spmv(self.cvals, x, self.tmp)
        r = (self.res + self.tmp - b) * self.free
        nr = float(torch.linalg.vector_norm(r))
        self._toc("residual", t0)
        return r, nr

    def _build_aux(self):
        self._aux = build_aux_space(self.Assembler, self._space_key, self.fixed, self.cvals,
                                    self.colptr, self.rowind, self.AuxSweeps, self.AuxOmega)

    def _linear_solve(self, rhs, eta):
        t0 = self._tic()
        try:
            return self._linear_solve_impl(rhs, eta)
        finally:
            self._toc("pcg", t0)

    def _linear_solve_impl(self, rhs, eta):
        dx = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        it = C.c_int()
        rel = C.c_double()
        if self.LinearPreconditioner == "aux":
            if self._aux_key != self._c_key:
                self._build_aux()
                self._aux_key = self._c_key
            aux, keep = self._aux
            L.check(L.lib.mpk_pcg_aux_solve(L.ptr(self.colptr), L.ptr(self.rowind),
                                            L.ptr(self.jsys), self.n, L.ptr(self.fixed),
                                            C.byref(aux), L.ptr(rhs), L.ptr(dx),
                                            L.ptr(keep["work"]), float(eta), self.LinearMaxIter,
                                            self.LinearCheckEvery, C.byref(it), C.byref(rel),
                                            self.sp))
            return dx, it.value
        L.check(L.lib.mpk_pcg_solve(L.ptr(self.colptr), L.ptr(self.rowind), L.ptr(self.jsys),
                                    self.n, L.ptr(self.fixed), L.ptr(rhs), L.ptr(dx),
                                    L.ptr(self.work), float(eta), self.LinearMaxIter,
                                    self.LinearCheckEvery, C.byref(it), C.byref(rel), self.sp))
        return dx, it.value

    def solve_linear_system(self, vals, rhs, eta):
        """PCG on a caller-assembled matrix of the plan's pattern (Dirichlet applied)."""
        saved = self.jsys
        self.jsys = vals
        try:
            return self._linear_solve(rhs * self.free, eta)
        finally:
            self.jsys = saved

    # ---------------------------------------------------------------- solve
    def solve(self, space_A, matLib, sigmaMap, sourceFunc, timeSteps, fixedDofs_A,
              init_A=None, warm_start=None, on_step=None, resume=None):
        """Runs the time steps; returns (solutions [device tensors], infos).

        sourceFunc(t) gives the source map (Assembler.expandSourceMap formats
        or an (Ne, 3) device tensor), or an assembled source vector (a 1-D
        device tensor of NumGlobalDofs entries).  warm_start(ctx) may return a
        candidate first iterate for the step; on_step(ctx) is called after
        each step.  resume (a snapshot(ctx) taken in an on_step hook of an
        earlier run on the same problem) continues that run after the
        snapshot's step: timeSteps are then the steps after it, and the
        previous two solutions, the time and the regularisation are the
        snapshot's.
        """
        self._setup(space_A, matLib, sigmaMap, fixedDofs_A)
        n = self.n
        x_prev = torch.zeros(n, dtype=torch.float64, device=self.dev)
        if init_A is not None:
            a0 = torch.as_tensor(init_A, dtype=torch.float64, device=self.dev).flatten()
            x_prev[: min(n, a0.numel())] = a0[:n]
        x_prev2 = x_prev.clone()
        results, infos = [], []
        t_now = 0.0
        step0 = 0
        if resume is not None:
            if resume["x"].numel() != n:
                raise ValueError("resume snapshot is for another problem size")
            x_prev = resume["x"].clone()
            x_prev2 = resume["x_prev"].clone()
            t_now = float(resume["t"])
            step0 = int(resume["step"])
            self.reg_abs = float(resume["reg_abs"])
        for step, dt in enumerate(np.asarray(timeSteps, dtype=np.float64).ravel(),
                                  start=step0 + 1):
            torch.cuda.synchronize()
            w0 = time.perf_counter()
            t_now += float(dt)
            src = sourceFunc(t_now)
            if torch.is_tensor(src) and src.dim() == 1:
                F = src.to(torch.float64)
            else:
                F = self.Assembler.assembleSource(space_A, src)
            # b = F + M_sigma x_prev / dt
            self._spmv(self.m_sigma, x_prev, self.tmp)
            b = F + self.tmp / float(dt)
            ctx = dict(step=step, t=t_now, dt=float(dt), x_prev=x_prev, x_prev2=x_prev2, b=b,
                       solver=self)
            x = x_prev.clone()
            label = "x_prev"
            rel_ref = None
            extrap = self.InitialGuess == "extrapolated" and step > 2
            if extrap and warm_start is None:
                # the relative test stays the one plain Newton applies: against
                # the residual at x_prev (the reference's first iterate, :159)
                _, rel_ref = self.residual(x_prev, b, dt)
                x, label = 2.0 * x_prev - x_prev2, "extrapolated"
            if warm_start is not None and self.reg_abs > 0:
                # the residual decides between warm start, extrapolation or x_prev
                _, r_prev = self.residual(x_prev, b, dt)
                ctx["r_prev"] = r_prev
                rel_ref = r_prev  # the plain method's first residual
                t0 = self._tic()
                cand = warm_start(ctx)
                self._toc("warm_start", t0)
                if cand is not None:
                    best, best_r = x, r_prev
                    if extrap:
                        xe = 2.0 * x_prev - x_prev2
                        _, r_e = self.residual(xe, b, dt)
                        if r_e < best_r:
                            best, best_r, label = xe, r_e, "extrapolated"
                    _, r_cand = self.residual(cand, b, dt)
                    if r_cand < best_r:
                        best, best_r, label = cand, r_cand, "warm"
                    x = best
            info = self._newton(step, t_now, float(dt), x, b, rel_ref)
            info.warm_start = label
            if "warm_pcg" in ctx:
                info.pcg_iters += int(ctx["warm_pcg"])
                info.pcg_per_solve.insert(0, int(ctx["warm_pcg"]))
            x = info.x
            del info.x
            torch.cuda.synchronize()
            info.wall_ms = 1e3 * (time.perf_counter() - w0)
            infos.append(info)
            results.append(x.clone())
            if on_step is not None:
                ctx.update(x=x, info=info)
                on_step(ctx)
            x_prev2, x_prev = x_prev, x
            if self.Verbose:
                print(f"step {step} t={t_now:.4g}: {info.newton_iters} Newton, "
                      f"{info.pcg_iters} PCG, res {info.residual:.3e} ({info.rel_residual:.2e}) "
                      f"[{label}] {info.wall_ms:.1f} ms", file=sys.stderr, flush=True)
                if self.Verbose > 1:
                    print("    ms: " + " ".join(f"{k} {v:.1f}" for k, v in self.timers.items()),
                          file=sys.stderr, flush=True)
                    self.timers = {}
                    ref = info.history[0] if rel_ref is None else rel_ref
                    print("    residuals/R0 " + " ".join(f"{h / ref:.1e}" for h in info.history) +
                          f" | PCG per solve {info.pcg_per_solve}", file=sys.stderr, flush=True)
        return results, infos

    def snapshot(self, ctx):
        """State after the step of an on_step ctx, for solve(..., resume=):
        the step and its time, its solution and the previous one, and the
        regularisation fixed at the first step."""
        return {"step": int(ctx["step"]), "t": float(ctx["t"]), "x": ctx["x"].clone(),
                "x_prev": ctx["x_prev"].clone(), "reg_abs": float(self.reg_abs)}

    def solveAV(self, space_A, space_V, matLib, sigmaMap, sourceFunc, timeSteps, fixedDofs_A,
                  fixedDofs_V=None, init_A=None, init_V=None):
        """TransientSolver.solve with the scalar-potential space (src/solver/
        TransientSolver.m:35-267, evaluateSystemState :287-311): unknowns
        [A; V], V on the nodes of the conducting regions.

            R = R_nu + M_sigma (x - x_prev)/dt - F + C V + [C^T (A - A_prev)/dt + L V]
                + reg M_reg x
            J = J_nu + M_sigma/dt + C + C^T/dt + L + reg M_reg

        with C = assembleCoupling(A, V, sigma) and L = assembleScalarLaplacian(V,
        sigma), reg from the mean positive |diag J| at the first iteration of the
        first step, V grounded at the given DOFs or the first V DOF. With
        V = W/dt the Newton matrix is J_nu + reg M_reg + (1/dt)[[M_sigma, C],
        [C^T, L]] in (A, W): symmetric positive definite once grounded, so the
        engine's auxiliary-space CG solves it (to LinearRelTolAV) on a merged CSC
        pattern (the plan's edge block, C, C^T, L). Linear materials skip the line
        search and stop below 1e-9 after a step, as the reference does.
        Returns (solutions [N device tensors], infos)."""
        import scipy.sparse as sp
        asm = self.Assembler
        dh = asm.DofHandler
        mesh = asm.Mesh
        dh.distributeDofs(space_A)
        nA = dh.SpaceLocalSizes[space_A.toString()]
        from .frequency import FrequencySolver
        tags_c = FrequencySolver.getConductingRegions(sigmaMap)
        if not tags_c or space_V is None:
            raise ValueError("solveAV needs a scalar-potential space and a conducting region; "
                             "use solve() for the edge space alone")
        kv = space_V.toString()
        if kv not in dh.DofMaps:
            _, offset_V = dh.distributeDofs(space_V, tags_c)
        else:
            offset_V = dh.SpaceOffsets[kv]
        N = dh.NumGlobalDofs
        from .nonlinear import _index_list
        fv = _index_list(fixedDofs_V, N)
        fixed = np.zeros(N, dtype=bool)
        fixed[_index_list(fixedDofs_A, N)] = True
        if fv.size:
            fixed[fv] = True
        else:
            fixed[offset_V] = True
        self._setup(space_A, matLib, sigmaMap, fixed)
        if self.n != N:
            raise RuntimeError(f"plan has {self.n} DOFs, NumGlobalDofs is {N}")
        dev = self.dev
        Cm = asm.assembleCoupling(space_A, space_V, sigmaMap)
        Lm = asm.assembleScalarLaplacian(space_V, sigmaMap)
        pcp, pri = self.plan.pattern()
        Cs = sp.csc_matrix((Cm.values.cpu().numpy(), Cm.rowind.cpu().numpy(),
                            Cm.colptr.cpu().numpy()), shape=(N, N))
        CT = Cs.T.tocsc()
        CT.sort_indices()
        Cs.sort_indices()
        blocks = [(pcp.cpu().numpy().astype(np.int64), pri.cpu().numpy()),
                  (Cs.indptr.astype(np.int64), Cs.indices), (CT.indptr.astype(np.int64), CT.indices),
                  (Lm.colptr.cpu().numpy().astype(np.int64), Lm.rowind.cpu().numpy())]
        mcp, mri, pos = _merged_pattern(N, blocks)
        nnz_m = int(mcp[-1])
        colptr = torch.from_numpy(mcp).to(dev)
        rowind = torch.from_numpy(mri).to(dev)
        pos_d = [torch.from_numpy(p).to(dev) for p in pos]
        c_vals = torch.from_numpy(Cs.data).to(dev)
        ct_vals = torch.from_numpy(CT.data).to(dev)
        l_vals = Lm.values.to(torch.float64)
        # symmetric C/C^T/L values for the residual's products
        cl = torch.zeros(nnz_m, dtype=torch.float64, device=dev)
        cl[pos_d[1]] = c_vals
        cl[pos_d[2]] = ct_vals
        cl[pos_d[3]] = l_vals
        lvv_diag = Lm.to_scipy().diagonal()[nA:]
        sysv = torch.zeros(nnz_m, dtype=torch.float64, device=dev)
        cvals_m = torch.zeros(nnz_m, dtype=torch.float64, device=dev)
        fixed_d = self.fixed
        free = self.free
        linear = all(getattr(m, "Type", "Linear") == "Linear" for m in dict(matLib).values())
        x_prev = torch.zeros(N, dtype=torch.float64, device=dev)
        if init_A is not None:
            a0 = torch.as_tensor(init_A, dtype=torch.float64, device=dev).flatten()
            x_prev[: min(nA, a0.numel())] = a0[:nA]
        if init_V is not None:
            v0 = torch.as_tensor(init_V, dtype=torch.float64, device=dev).flatten()
            x_prev[offset_V: offset_V + v0.numel()] = v0
        y1 = torch.empty(N, dtype=torch.float64, device=dev)
        y2 = torch.empty_like(y1)
        mx = torch.empty_like(y1)

        def spmv_m(vals, x, y):
            L.check(L.lib.mpk_spmv_sym(L.ptr(colptr), L.ptr(rowind), L.ptr(vals), N, L.ptr(x),
                                       L.ptr(y), self.sp))

        def residual(x, dt, F, calc_j):
            self._jacobian(x, calc_j)
            dx_ = x - x_prev
            self._spmv(self.m_sigma, dx_, mx)
            R = self.res + mx / dt - F
            xv = torch.zeros_like(x)
            xv[nA:] = x[nA:]
            spmv_m(cl, xv, y1)                 # [C V; L V]
            xa = torch.zeros_like(x)
            xa[:nA] = dx_[:nA]
            spmv_m(cl, xa, y2)                 # [0; C^T (A - A_prev)] (+ C 0)
            y2[:nA] = 0.0
            R = R + y1 + y2 / dt
            if self.AutoRegularization and self.reg_abs > 0:
                self._spmv(self.m_reg, x, mx)
                R = R + self.reg_abs * mx
            return R

        results, infos = [], []
        t_now = 0.0
        aux_key = None
        for step, dt in enumerate(np.asarray(timeSteps, dtype=np.float64).ravel(), start=1):
            dt = float(dt)
            t_now += dt
            src = sourceFunc(t_now)
            F = (src.to(torch.float64) if torch.is_tensor(src) and src.dim() == 1
                 else asm.assembleSource(space_A, src))
            x = x_prev.clone()
            info = StepInfo(step, t_now, 0, 0, 0, math.inf, math.inf, False, 0.0)
            initial = 1.0
            for it in range(1, self.MaxIterations + 1):
                R = residual(x, dt, F, True)
                if self.reg_abs == 0 and self.AutoRegularization:
                    self._combine(self.jsys, 1.0, self.jv, 1.0 / dt, self.m_sigma)
                    d = np.concatenate([self.jsys[self.diag_pos].abs().cpu().numpy(),
                                        np.abs(lvv_diag)])
                    d = d[d > 0]
                    avg = float(d.mean()) if d.size else 1.0
                    self.reg_abs = (avg if math.isfinite(avg) and avg != 0 else 1.0) * self.RegScaleRel
                    self._spmv(self.m_reg, x, mx)
                    R = R + self.reg_abs * mx
                Rb = R * free
                norm_res = float(torch.linalg.vector_norm(Rb))
                info.history.append(norm_res)
                if it == 1:
                    initial = norm_res if norm_res >= 1e-20 else 1.0
                info.newton_iters = it
                info.residual, info.rel_residual = norm_res, norm_res / initial
                if not math.isfinite(norm_res):
                    break  # a NaN/Inf residual ends the step unconverged
                if norm_res < self.Tolerance or norm_res / initial < self.RelTolerance:
                    info.converged = True
                    break
                # scaled symmetric Newton matrix on the merged pattern
                self._combine(self.cvals, 1.0 / dt, self.m_sigma, self.reg_abs, self.m_reg)
                self._combine(self.jsys, 1.0, self.jv, 1.0, self.cvals)
                sysv.zero_()
                sysv[pos_d[0]] = self.jsys[: pos_d[0].numel()]
                sysv[pos_d[1]] = c_vals / dt
                sysv[pos_d[2]] = ct_vals / dt
                sysv[pos_d[3]] = l_vals / dt
                key = (dt, self.reg_abs)
                if key != aux_key:
                    cvals_m.zero_()
                    cvals_m[pos_d[0]] = self.cvals[: pos_d[0].numel()]
                    aux, keep = build_aux_space(asm, space_A.toString(), fixed_d, cvals_m, colptr,
                                                rowind, self.AuxSweeps, self.AuxOmega)
                    aux_key = key
                w = torch.zeros(N, dtype=torch.float64, device=dev)
                itc, rel = C.c_int(), C.c_double()
                L.check(L.lib.mpk_pcg_aux_solve(L.ptr(colptr), L.ptr(rowind), L.ptr(sysv), N,
                                                L.ptr(fixed_d), C.byref(aux), L.ptr(-Rb), L.ptr(w),
                                                L.ptr(keep["work"]), self.LinearRelTolAV,
                                                self.LinearMaxIter, self.LinearCheckEvery,
                                                C.byref(itc), C.byref(rel), self.sp))
                info.pcg_iters += int(itc.value)
                info.pcg_per_solve.append(int(itc.value))
                if not bool(torch.isfinite(w).all()):
                    break  # the reference stops on a NaN step (TransientSolver.m:201-204)
                dxv = w.clone()
                dxv[nA:] = w[nA:] / dt       # V = W / dt
                alpha = self.Relaxation
                if self.UseLineSearch and not linear:
                    for k in range(self.MaxLineSearchIters + 1):
                        a_try = alpha * (0.5 ** k)
                        xt = x + a_try * dxv
                        rt = float(torch.linalg.vector_norm(residual(xt, dt, F, False) * free))
                        info.line_search_evals += 1
                        if rt < norm_res or (it == 1 and k < 3):
                            alpha = a_try
                            break
                        if rt > norm_res * 10:
                            alpha = alpha * 0.1
                x = x + alpha * dxv
                if norm_res > 1e20:
                    raise RuntimeError("Divergence detected (Res > 1e20). Check BCs or Time Step.")
                if linear and norm_res < 1e-9:
                    info.converged = True
                    break
            infos.append(info)
            results.append(x.clone())
            x_prev = x
        return results, infos

    def _newton(self, step, t, dt, x, b, rel_ref):
        info = StepInfo(step, t, 0, 0, 0, math.inf, math.inf, False, 0.0)
        initial_res = rel_ref
        prev_norm = None
        target = None
        for it in range(1, self.MaxIterations + 1):
            t0 = self._tic()
            self._jacobian(x, calc_j=True)
            if self.reg_abs == 0 and self.AutoRegularization:
                self._combine(self.jsys, 1.0, self.jv, 1.0 / dt, self.m_sigma)
                d = self.jsys[self.diag_pos].abs()
                d = d[d > 0]
                avg = float(d.mean()) if d.numel() else 1.0
                if not math.isfinite(avg) or avg == 0:
                    avg = 1.0
                self.reg_abs = avg * self.RegScaleRel
            self._constant_part(dt)
            self._combine(self.jsys, 1.0, self.jv, 1.0, self.cvals)
            self._spmv(self.cvals, x, self.tmp)
            r = (self.res + self.tmp - b) * self.free
            norm_res = float(torch.linalg.vector_norm(r))
            self._toc("jacobian", t0)
            info.history.append(norm_res)
            if initial_res is None:
                initial_res = norm_res if norm_res >= 1e-20 else 1.0
            rel = norm_res / initial_res
            info.newton_iters = it
            info.residual, info.rel_residual = norm_res, rel
            if norm_res < self.Tolerance or rel < self.RelTolerance:
                info.converged = True
                break
            if norm_res > 1e20:
                raise RuntimeError("Divergence detected (Res > 1e20). Check BCs or Time Step.")
            target = max(self.Tolerance, self.RelTolerance * initial_res)
            # Eisenstat-Walker forcing, never tighter than what reaching the
            # target from this residual needs
            if self.NewtonLinearTolerance is not None:
                eta = float(self.NewtonLinearTolerance)
            else:
                eta = self.LinearEtaMax
                if prev_norm is not None:
                    eta = min(eta, self.LinearGamma * (norm_res / prev_norm) ** 2)
                eta = min(max(eta, self.LinearEtaFloor * target / norm_res), 0.5)
            dx, nit = self._linear_solve(-r, eta)
            info.pcg_iters += nit
            info.pcg_per_solve.append(nit)
            if not bool(torch.isfinite(dx).all()):
                break
            alpha = self.Relaxation
            if self.UseLineSearch:
                for k in range(self.MaxLineSearchIters + 1):
                    a_try = alpha * (0.5 ** k)
                    x_try = x + a_try * dx
                    _, r_try = self.residual(x_try, b, dt)
                    info.line_search_evals += 1
                    if r_try < norm_res or (it == 1 and k < 3):
                        alpha = a_try
                        break
                    if r_try > norm_res * 10:
                        alpha = alpha * 0.1
            x = x + alpha * dx
            prev_norm = norm_res
        info.x = x
        return info


def _merged_pattern(n, blocks):
    """Merged CSC pattern of disjoint blocks. blocks: list of (colptr, rowind)
    int64 host arrays over the same n x n index space, supports pairwise
    disjoint. Returns (colptr int32, rowind int32, [slot index per block
    entry]), all host arrays."""
    import scipy.sparse as sp
    mats, offs, tot = [], [], 0
    for cp, ri in blocks:
        nnz = int(cp[-1])
        mats.append(sp.csc_matrix((np.arange(tot + 1, tot + nnz + 1, dtype=np.float64), ri, cp),
                                  shape=(n, n)))
        offs.append(tot)
        tot += nnz
    Msum = mats[0]
    for m in mats[1:]:
        Msum = Msum + m
    Msum = Msum.tocsc()
    Msum.sort_indices()
    if Msum.nnz != tot:
        raise RuntimeError("block patterns overlap")
    src = Msum.data.astype(np.int64) - 1
    slot_of = np.empty(tot, dtype=np.int64)
    slot_of[src] = np.arange(tot)
    pos = [slot_of[o:o + int(cp[-1])] for o, (cp, _) in zip(offs, blocks)]
    return Msum.indptr.astype(np.int32), Msum.indices.astype(np.int32), pos
