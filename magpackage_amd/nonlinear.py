"""Static nonlinear magnetostatic driver on the GPU engine.

Mirror of NonlinearSolver (src/solver/NonlinearSolver.m:1-276): the same
properties and defaults (:3-16), ``solve`` (:23-222) and ``evaluateSystem``
(:241-275). Per Newton iteration:

    R(x) = R_nu(x) - F + eps * M_reg * x
    J(x) = J_nu(x) + eps * M_reg

Here R_nu and J_nu come from the Jacobian kernel, F is the source vector, and
M_reg is the unweighted edge mass. eps = RegScale * max|diag J_sys| is taken
at the initial iterate (:127-136). Dirichlet DOFs are removed as
BoundaryCondition.applyDirichlet does. The line search follows :150-173: the
first iteration always takes the relaxed step, and later ones halve it until
the residual (fixed rows zeroed *before* eps M x is added) drops. The loop
stops when |R| < Tolerance or |alpha dx| < 1e-10 (:183-191).

The scalar-potential branch (``space_V`` with a conducting region) follows
evaluateSystem's DC equations (:249-265):
- rows of A read R_nu + C_AV V - F, and rows of V read L_VV V;
- C_AV is the (A rows, V columns) block only, so the Newton matrix is block
  upper-triangular;
- with V grounded (the given fixed DOFs, else the first V DOF, :47-51), the V
  rows force dV = 0. V therefore stays zero from a zero start, and A solves
  the edge-space problem alone.
The branch enters A only through eps, whose maximum diagonal then includes
L_VV's. The engine builds L_VV (``Assembler.assembleScalarLaplacian``) for
that maximum and treats the V DOFs as fixed in its edge-space solve. A
nonzero initial V is rejected.

The linear solve is the engine's auxiliary-space PCG on the Dirichlet-applied
operator. It is solved to a relative residual of ``LinearRelTol`` (1e-10), so
that the steps are those of the reference's direct solve to that accuracy;
MUMPS is not available. ``MatrixK_Add`` is kept as a property: the
reference's evaluateSystem does not use it (:266-268).
"""
from __future__ import annotations

import math
import sys

import numpy as np
import torch

from . import _lib as L
from .assembler import Assembler
from .frequency import FrequencySolver
from .transient import TransientSolver


def _index_list(fixed, n: int) -> np.ndarray:
    """0-based indices of a logical mask or a 1-based index list, within n."""
    if fixed is None:
        return np.zeros(0, dtype=np.int64)
    a = fixed.detach().cpu().numpy() if torch.is_tensor(fixed) else np.asarray(fixed)
    if a.size == 0:
        return np.zeros(0, dtype=np.int64)
    if a.dtype == bool:
        idx = np.nonzero(a.ravel())[0]
    else:
        idx = a.astype(np.int64).ravel() - 1
    return idx[(idx >= 0) & (idx < n)]


class NonlinearSolver:
    """NonlinearSolver with the reference's properties and defaults."""

    def __init__(self, assembler: Assembler):
        self.Assembler = assembler
        self.MaxIterations = 100
        self.Tolerance = 1e-6
        self.Relaxation = 1.0
        self.UseLineSearch = True
        self.MaxLineSearchIters = 10
        self.MatrixK_Add = None
        self.AutoRegularization = True
        self.RegScale = 1e-6
        # engine linear solver
        self.LinearRelTol = 1e-10
        self.LinearMaxIter = 20000
        self.LinearCheckEvery = 10
        self.LinearPreconditioner = "aux"
        self.Precision = "fp64"
        self.Verbose = False

    getConductingRegions = staticmethod(FrequencySolver.getConductingRegions)

    def solve(self, space_A, space_V, matLib, sigmaMap, sourceMap, fixedDofs_A,
              fixedDofs_V=None, x_init=None):
        """Returns (A_sol, V_sol, info): device tensors of the edge DOFs and of
        the V DOFs (None without a conducting scalar-potential space), and a
        dict with Converged, Iterations, FinalResidual, Solution (the full
        [NumGlobalDofs] vector) and the engine's LinearIterations per step."""
        asm = self.Assembler
        dh = asm.DofHandler
        _, offset_A = dh.distributeDofs(space_A)
        num_A = dh.SpaceLocalSizes[space_A.toString()]
        if num_A == 0:
            raise RuntimeError("NonlinearSolver:NoDoFs Active DoFs is 0. "
                               "Check Mesh initialization.")
        if offset_A != 0:
            raise ValueError("the edge space must hold the first DOFs (A_sol = x(1:num_A))")
        enable_V = space_V is not None
        offset_V = -1
        auto_ground = None
        if enable_V:
            tags = self.getConductingRegions(sigmaMap)
            if not tags:
                enable_V = False
            else:
                kv = space_V.toString()
                if kv not in dh.DofMaps:
                    _, offset_V = dh.distributeDofs(space_V, tags)
                else:
                    offset_V = dh.SpaceOffsets[kv]
                fv = _index_list(fixedDofs_V, 1 << 62)
                if fv.size == 0:
                    auto_ground = offset_V  # 0-based offset_V + 1
        N = dh.NumGlobalDofs
        fixed_idx = _index_list(fixedDofs_A, N)
        if enable_V:
            idx_V = _index_list(fixedDofs_V, N)
            if auto_ground is not None:
                idx_V = np.unique(np.append(idx_V, auto_ground))
            fixed_idx = np.concatenate([fixed_idx, idx_V[idx_V < N]])
        fixed = np.zeros(N, dtype=bool)
        fixed[fixed_idx] = True
        # engine: the edge-space system, every DOF outside it held fixed
        eng_fixed = fixed.copy()
        eng_fixed[num_A:] = True
        eng = TransientSolver(asm)
        eng.AutoRegularization = self.AutoRegularization
        eng.LinearPreconditioner = self.LinearPreconditioner
        eng.LinearMaxIter = self.LinearMaxIter
        eng.LinearCheckEvery = self.LinearCheckEvery
        eng.Precision = self.Precision
        eng._setup(space_A, matLib, np.zeros(asm.Mesh.NumElements), eng_fixed)
        if eng.n != N:
            raise RuntimeError(f"plan has {eng.n} DOFs, NumGlobalDofs is {N}")
        dev = eng.dev
        F = asm.assembleSource(space_A, sourceMap).to(torch.float64)
        x = torch.zeros(N, dtype=torch.float64, device=dev)
        if x_init is not None:
            x0 = torch.as_tensor(x_init, dtype=torch.float64, device=dev).flatten()
            k = min(x0.numel(), N)
            x[:k] = x0[:k]
        if bool((x[num_A:] != 0).any()):
            raise NotImplementedError("a nonzero initial scalar potential is not supported: "
                                      "the engine solves the edge space with V = 0")
        free = torch.from_numpy(~fixed).to(dev).to(torch.float64)
        fixed_dev = torch.from_numpy(fixed).to(dev)
        mreg_x = torch.empty(N, dtype=torch.float64, device=dev)

        def evaluate(xv, calc_j):
            eng._jacobian(xv, calc_j)
            return eng.res - F

        def add_reg(r, xv):
            if self.AutoRegularization:
                eng._spmv(eng.m_reg, xv, mreg_x)
                r = r + eps * mreg_x
            return r

        # initial evaluation and the regularisation scale
        eps = 0.0
        R = evaluate(x, True)
        if self.AutoRegularization:
            ref = float(eng.jv[eng.diag_pos].abs().max()) if eng.diag_pos.numel() else 0.0
            if enable_V:
                lvv = asm.assembleScalarLaplacian(space_V, sigmaMap)
                cols = torch.repeat_interleave(
                    torch.arange(N, device=dev), (lvv.colptr[1:] - lvv.colptr[:-1]))
                d = lvv.values[lvv.rowind.to(torch.int64) == cols].abs()
                if d.numel():
                    ref = max(ref, float(d.max()))
            if ref == 0:
                ref = 1.0
            eps = ref * self.RegScale
        eng.reg_abs = eps
        eng._constant_part(1.0)  # eps * M_reg: the aux preconditioner's mass part

        def system(xv, r):
            if self.AutoRegularization:
                eng._combine(eng.jsys, 1.0, eng.jv, eps, eng.m_reg)
            else:
                eng._combine(eng.jsys, 1.0, eng.jv)
            return add_reg(r, xv) * free

        Rb = system(x, R)
        norm_res = float(torch.linalg.vector_norm(Rb))
        converged = False
        history = [norm_res]
        lin_iters = []
        it = 0
        for it in range(1, self.MaxIterations + 1):
            dx, nit = eng._linear_solve(-Rb, self.LinearRelTol)
            lin_iters.append(nit)
            alpha = self.Relaxation
            if self.UseLineSearch:
                res_prev = norm_res
                for k in range(self.MaxLineSearchIters + 1):
                    a_try = alpha * (0.5 ** k)
                    x_try = x + a_try * dx
                    R_try = evaluate(x_try, False)
                    R_try = R_try.masked_fill(fixed_dev, 0.0)
                    R_try = add_reg(R_try, x_try)
                    res_try = float(torch.linalg.vector_norm(R_try))
                    if res_try < res_prev or (it == 1 and k < 3):
                        alpha = a_try
                        break
            x = x + alpha * dx
            norm_step = float(torch.linalg.vector_norm(alpha * dx))
            R = evaluate(x, True)
            Rb = system(x, R)
            norm_res = float(torch.linalg.vector_norm(Rb))
            history.append(norm_res)
            if self.Verbose:
                print(f"  iter {it}: alpha {alpha:.3g}, |R| {norm_res:.3e}, "
                      f"|dx| {norm_step:.3e}, PCG {nit}", file=sys.stderr, flush=True)
            if norm_res < self.Tolerance or norm_step < 1e-10 or not math.isfinite(norm_res):
                converged = math.isfinite(norm_res)
                break
        info = dict(Converged=converged, Iterations=it, FinalResidual=norm_res, Solution=x,
                    History=history, LinearIterations=lin_iters, RegScaleAbs=eps)
        A_sol = x[:num_A]
        V_sol = x[offset_V:] if enable_V else None
        return A_sol, V_sol, info
