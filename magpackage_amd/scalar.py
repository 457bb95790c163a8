"""Scalar-potential (DC conduction) solver on the GPU engine.

Mirror of ScalarSolver (src/solver/ScalarSolver.m:1-58):
- S = assembleScalarLaplacian(space, sigmaMap), on the GPU;
- the Dirichlet values are lifted (F = -S V_fixed);
- the system is BoundaryCondition.applyDirichlet's: fixed rows and columns
  removed, a unit diagonal, and F_sys(bcNodes) = bcValues.
The reference hands that system to MUMPS in symmetric mode. Here the engine's
Jacobi-PCG solves the free block to a relative residual of ``LinearRelTol``,
and the fixed entries are set to their values. That is the same vector the
unit-diagonal rows give.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib as L
from .assembler import Assembler


class ScalarSolver:
    def __init__(self, assembler: Assembler):
        self.Assembler = assembler
        self.LinearRelTol = 1e-12
        self.LinearMaxIter = 200000
        self.LinearCheckEvery = 10

    def solve(self, space, sigmaMap, bcNodes, bcValues):
        """Returns (V_sol [NumGlobalDofs] device tensor, Info with SolveTime,
        Iterations and RelResidual). bcNodes are 1-based DOF numbers."""
        S = self.Assembler.assembleScalarLaplacian(space, sigmaMap)
        n = S.shape[0]
        dev = S.values.device
        idx = np.asarray(bcNodes, dtype=np.int64).ravel() - 1
        vals = np.broadcast_to(np.asarray(bcValues, dtype=np.float64).ravel(), idx.shape)
        if idx.size and (idx.min() < 0 or idx.max() >= n):
            raise IndexError("bcNodes outside the DOF range")
        fixed = np.zeros(n, dtype=np.uint8)
        fixed[idx] = 1
        vfix = np.zeros(n)
        vfix[idx] = vals
        colptr = S.colptr.to(torch.int32).contiguous()
        rowind = S.rowind.contiguous()
        sv = S.values.contiguous()
        sp = L.stream_ptr()
        vfix_d = torch.from_numpy(vfix).to(dev)
        Sv = torch.empty(n, dtype=torch.float64, device=dev)
        L.check(L.lib.mpk_spmv_sym(L.ptr(colptr), L.ptr(rowind), L.ptr(sv), n, L.ptr(vfix_d),
                                   L.ptr(Sv), sp))
        F = -Sv
        fixed_d = torch.from_numpy(fixed).to(dev)
        ws = C.c_int64()
        L.check(L.lib.mpk_pcg_work_size(n, C.byref(ws)))
        work = torch.empty(ws.value, dtype=torch.float64, device=dev)
        x = torch.empty(n, dtype=torch.float64, device=dev)
        it, rel = C.c_int(), C.c_double()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.check(L.lib.mpk_pcg_solve(L.ptr(colptr), L.ptr(rowind), L.ptr(sv), n, L.ptr(fixed_d),
                                    L.ptr(F), L.ptr(x), L.ptr(work), float(self.LinearRelTol),
                                    int(self.LinearMaxIter), int(self.LinearCheckEvery),
                                    C.byref(it), C.byref(rel), sp))
        torch.cuda.synchronize()
        t_solve = time.perf_counter() - t0
        L.check_krylov(it, rel, float(self.LinearRelTol), "ScalarSolver.solve")
        x = torch.where(fixed_d.bool(), vfix_d, x)
        info = dict(SolveTime=t_solve, Iterations=int(it.value), RelResidual=float(rel.value))
        return x, info
