"""Linear solver front end on the GPU engine.

Mirror of LinearSolver (src/solver/LinearSolver.m:1-157). It keeps the
properties, ``solve(A, b)``, ``factorize`` / ``solveFromFactors`` /
``clearFactors``, and the NaN/Inf checks (:23-24). The reference chooses
MUMPS, MATLAB's backslash or GMRES. This engine has no sparse direct
factorisation, so every method runs a Krylov solve on the device to a relative
residual of ``KrylovTolerance`` (1e-12 by default):
- CG with Jacobi preconditioning for real symmetric matrices;
- COCG (unconjugated CG) with Jacobi preconditioning for complex symmetric
  ones, the systems the reference hands MUMPS in symmetric mode.
An unsymmetric matrix raises NotImplementedError: it would need MUMPS' LU.
``factorize`` uploads the matrix once and returns a handle that
``solveFromFactors`` reuses.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


class LinearSolver:
    def __init__(self, method: str = "Auto"):
        self.Method = method
        self.Tolerance = 1e-6      # the reference's GMRES tolerance (kept)
        self.MaxIter = 1000
        self.MumpsSymmetry = 0
        self.MumpsICNTL = {"i14": 100, "i7": 5}
        self.KrylovTolerance = 1e-12
        self.KrylovMaxIter = 200000
        self.SymmetryTolerance = 1e-12
        self.LastIterations = 0
        self.LastRelResidual = float("nan")

    # ------------------------------------------------------------------
    def _upload(self, A):
        """Device CSC (int32 colptr / rowind, real and imaginary values) of a
        square symmetric matrix: a scipy sparse matrix or the engine's
        CscMatrix."""
        import scipy.sparse as sp
        if hasattr(A, "to_scipy"):
            A = A.to_scipy()
        A = sp.csc_matrix(A)
        n, m = A.shape
        if n != m:
            raise ValueError("A must be square")
        data = A.data
        if not np.all(np.isfinite(data)):
            raise ValueError("Matrix A contains NaN/Inf")
        D = (A - A.T).tocsc()
        scale = np.abs(data).max() if data.size else 0.0
        if D.nnz and np.abs(D.data).max() > self.SymmetryTolerance * max(scale, 1e-300):
            raise NotImplementedError("unsymmetric systems need a sparse LU (MUMPS), which "
                                      "this engine does not provide")
        A.sort_indices()
        dev = _dev()
        h = dict(n=n, is_complex=np.iscomplexobj(data),
                 colptr=torch.from_numpy(A.indptr.astype(np.int32)).to(dev),
                 rowind=torch.from_numpy(A.indices.astype(np.int32)).to(dev),
                 re=torch.from_numpy(np.ascontiguousarray(data.real, dtype=np.float64)).to(dev))
        if h["is_complex"]:
            h["im"] = torch.from_numpy(np.ascontiguousarray(data.imag,
                                                            dtype=np.float64)).to(dev)
        ws = C.c_int64()
        if h["is_complex"]:
            L.check(L.lib.mpk_cocg_aux_work_size(n, 0, C.byref(ws)))
        else:
            L.check(L.lib.mpk_pcg_work_size(n, C.byref(ws)))
        h["work"] = torch.empty(max(ws.value, 1), dtype=torch.float64, device=dev)
        return h

    def _solve_one(self, h, b):
        n = h["n"]
        dev = h["re"].device
        it, rel = C.c_int(), C.c_double()
        if h["is_complex"] or np.iscomplexobj(b):
            im = h.get("im")
            if im is None:
                im = torch.zeros_like(h["re"])
            bd = torch.from_numpy(np.ascontiguousarray(b, dtype=np.complex128)).to(dev)
            x = torch.empty(n, dtype=torch.complex128, device=dev)
            work = h["work"]
            if not h["is_complex"]:
                ws = C.c_int64()
                L.check(L.lib.mpk_cocg_aux_work_size(n, 0, C.byref(ws)))
                work = torch.empty(ws.value, dtype=torch.float64, device=dev)
            L.check(L.lib.mpk_cocg_aux_solve(L.ptr(h["colptr"]), L.ptr(h["rowind"]),
                                             L.ptr(h["re"]), L.ptr(im), n, None, None,
                                             L.ptr(bd), L.ptr(x), L.ptr(work),
                                             self.KrylovTolerance, self.KrylovMaxIter, 10,
                                             C.byref(it), C.byref(rel), L.stream_ptr()))
        else:
            bd = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).to(dev)
            x = torch.empty(n, dtype=torch.float64, device=dev)
            L.check(L.lib.mpk_pcg_solve(L.ptr(h["colptr"]), L.ptr(h["rowind"]), L.ptr(h["re"]),
                                        n, None, L.ptr(bd), L.ptr(x), L.ptr(h["work"]),
                                        self.KrylovTolerance, self.KrylovMaxIter, 10,
                                        C.byref(it), C.byref(rel), L.stream_ptr()))
        self.LastIterations = int(it.value)
        self.LastRelResidual = float(rel.value)
        # an unconverged Krylov solve is an error, as a failed factorisation
        # would be for the direct solvers this front end stands in for
        L.check_krylov(it, rel, self.KrylovTolerance, "LinearSolver.solve")
        return x.cpu().numpy()

    # ------------------------------------------------------------------
    def solve(self, A, b):
        """x = A \\ b for a symmetric (real or complex) sparse A; b [n] or [n, k]."""
        if str(self.Method).upper() not in ("AUTO", "MUMPS", "BACKSLASH", "DIRECT", "ITERATIVE"):
            raise ValueError(f"Unknown method: {self.Method}")
        bb = b.detach().cpu().numpy() if torch.is_tensor(b) else np.asarray(b)
        if not np.all(np.isfinite(bb)):
            raise ValueError("RHS b contains NaN/Inf")
        h = self._upload(A)
        return self.solveFromFactors(h, bb)

    def factorize(self, A):
        return self._upload(A)

    def solveFromFactors(self, mumpsID, b, A=None):
        bb = b.detach().cpu().numpy() if torch.is_tensor(b) else np.asarray(b)
        if bb.ndim == 1:
            return self._solve_one(mumpsID, bb)
        return np.stack([self._solve_one(mumpsID, bb[:, k]) for k in range(bb.shape[1])],
                        axis=1)

    def clearFactors(self, mumpsID):
        mumpsID.clear()
