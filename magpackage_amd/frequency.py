"""Time-harmonic solver on the GPU engine (src/solver/FrequencySolver.m:1-241).

``FrequencySolver(assembler, freq).solve(space_A, space_V, matLib, sigmaMap,
sourceMap, fixedDofs_A, fixedDofs_V)`` follows the reference's solve.

Operator. Per outer iteration (FrequencySolver.m:116-139) it forms
    S = K(nu) + 1j*omega*M_sigma + eps*M_reg,
    eps = 1e-6 * max|diag K|,
with M_reg the unweighted edge mass. M_sigma is there only when a scalar space
is given and some region conducts (:63-77, :96-100), as in the reference.

Nonlinear materials. The outer loop re-evaluates the effective reluctivity
nu(|B|^2) per element from the current complex solution (:205-238). It
relaxes the update adaptively (:141-170) and stops when the relative change
falls below Tolerance.

Linear solve. The reference hands S to MUMPS in symmetric mode (:34-38).
Here S is kept as R + 1j*I on the plan's structural pattern and solved by
``mpk_cocg_aux_solve``. That is conjugate-orthogonal CG preconditioned by
the auxiliary-space operator of P = R + I. The set-up of P's gradient-space
matrix is host scipy, as for the transient driver.

A-V systems: the A* gauge. With a scalar space the reference solves the
A-V system: its A block, the coupling C, and the scalar Laplacian L of the
conducting region, with one node grounded (:98-100, :130-133). Here the same
physics is solved in the A* gauge, A* = A + grad(V_mod), V_mod = V/(1j*omega).
- Substitution turns the A-V equations into the single equation
  curl nu curl A* + 1j*omega*sigma*A* = Js. The gradients of the nodal space
  lie in the edge space, so the discrete solutions correspond exactly, up to
  the eps*M_reg term. That term acts on A here and on A* in the reference.
- B = curl A* = curl A, and the eddy current
  -sigma*(1j*omega*A + grad V) = -1j*omega*sigma*A* are the same fields.
- The returned scalar potential is therefore zero: V_sol has one entry per
  node of the conducting regions, as the reference's does.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from .assembler import Assembler, FunctionSpace, _dev
from .material import MU0, ppval


def _fixed_mask(fixed, n: int) -> np.ndarray:
    """Logical mask or 1-based index list -> bool [n] (FrequencySolver.m:105-116)."""
    if fixed is None:
        return np.zeros(n, dtype=bool)
    f = fixed.cpu().numpy() if torch.is_tensor(fixed) else np.asarray(fixed)
    if f.dtype == bool:
        out = np.zeros(n, dtype=bool)
        out[: min(n, f.size)] = f.ravel()[:n]
        return out
    idx = f.ravel().astype(np.int64)
    idx = idx[(idx >= 1) & (idx <= n)]
    out = np.zeros(n, dtype=bool)
    out[idx - 1] = True
    return out


def _mat_of(matLib, tag):
    if matLib is None:
        return None
    for k in (tag, float(tag), str(tag)):
        if k in matLib:
            return matLib[k]
    return None


class FrequencySolver:
    def __init__(self, assembler: Assembler, freq: float):
        self.Assembler = assembler
        self.Frequency = float(freq)
        self.MaxIterations = 200
        self.Tolerance = 1e-5
        self.Relaxation = 0.5
        self.MinRelaxation = 0.05
        self.MaxRelaxation = 1.0
        self.RelaxationGrow = 1.05
        self.RelaxationShrink = 0.6
        # linear solve (COCG) settings
        self.LinearTolerance = 1e-10
        self.LinearMaxIterations = 20000
        self.LinearCheckEvery = 25
        self.AuxSweeps = 2
        self.AuxOmega = 2.0 / 3.0
        self.LinearIterations = []
        self.LinearResiduals = []
        self.Iterations = 0

    @staticmethod
    def getConductingRegions(sigmaMap) -> list:
        """Tags whose conductivity exceeds 1e-12 (FrequencySolver.m:185-200)."""
        tags = []
        if isinstance(sigmaMap, dict):
            for k, v in sigmaMap.items():
                if np.max(np.abs(np.atleast_1d(v))) > 1e-12:
                    tags.append(int(float(k)))
        elif sigmaMap is not None:
            a = np.asarray(sigmaMap, dtype=np.float64).ravel()
            tags = [int(i) + 1 for i in np.nonzero(a > 1e-12)[0]]
        return tags

    # ------------------------------------------------------------------
    def _effective_nu(self, matLib, A_curr, space_key: str) -> np.ndarray:
        """updateEffectiveNu (FrequencySolver.m:205-238): per element,
        Nu_Linear for linear materials (1/mu0 without a material),
        ppval(SplineNu, |B|^2) for nonlinear ones, |B| the 2-norm of the
        complex element B."""
        mesh = self.Assembler.Mesh
        tags = mesh.RegionTags
        nu = np.zeros(mesh.NumElements)
        nonlin = []
        for t in np.unique(tags):
            m = _mat_of(matLib, int(t))
            sel = tags == t
            if m is None:
                nu[sel] = 1.0 / (4 * math.pi * 1e-7)
            elif str(m.Type).lower() == "linear":
                nu[sel] = m.Nu_Linear
            else:
                nonlin.append((sel, m))
        if nonlin:
            space = FunctionSpace("Nedelec", 1)
            Bre = self.Assembler.computeElementB(A_curr.real.contiguous(), space)
            Bim = self.Assembler.computeElementB(A_curr.imag.contiguous(), space)
            bsq = ((Bre * Bre).sum(1) + (Bim * Bim).sum(1)).cpu().numpy()
            for sel, m in nonlin:
                if m.SplineNu is not None:
                    nu[sel] = ppval(m.SplineNu, bsq[sel])
                else:
                    nu[sel] = m.nu0
        return nu

    def _aux(self, fixed_t, c_vals, colptr, rowind, key):
        from .transient import build_aux_space
        aux, keep = build_aux_space(self.Assembler, key, fixed_t, c_vals, colptr, rowind,
                                    self.AuxSweeps, self.AuxOmega)
        ws = C.c_int64()
        L.check(L.lib.mpk_cocg_aux_work_size(len(fixed_t), aux.num_nodes, C.byref(ws)))
        keep["work"] = torch.empty(ws.value, dtype=torch.float64, device=c_vals.device)
        return aux, keep

    def solve(self, space_A, space_V, matLib, sigmaMap, sourceMap, fixedDofs_A,
              fixedDofs_V=None):
        """Returns (A_sol, V_sol): complex128 device tensors, A_sol [num_A],
        V_sol [nodes of the conducting regions] (zero, the A* gauge) or None."""
        asm = self.Assembler
        mesh = asm.Mesh
        omega = 2 * math.pi * self.Frequency
        if omega == 0:
            omega = 1e-9
        dh = asm.DofHandler
        _, offset_A = dh.distributeDofs(space_A)
        if offset_A != 0:
            raise ValueError("the edge space must be the first space distributed")
        key = space_A.toString()
        num_A = dh.SpaceLocalSizes[key]
        enable_V = False
        num_V = 0
        if space_V is not None:
            cond = self.getConductingRegions(sigmaMap)
            if cond:
                enable_V = True
                num_V = int(np.unique(mesh.T[np.isin(mesh.RegionTags, cond)]).size)
        plan = asm.plan(space_A)
        n = plan.n
        dev = _dev()
        M_reg = asm.assembleMass(space_A, drop_zeros=False).values
        M_sig = (asm.assembleMassWeighted(space_A, sigmaMap, drop_zeros=False).values
                 if enable_V else None)
        F = asm.assembleSource(space_A, sourceMap).clone()
        b = torch.zeros(n, dtype=torch.complex128, device=dev)
        b.real.copy_(F)
        fixed = _fixed_mask(fixedDofs_A, n)
        fixed_t = torch.from_numpy(fixed.astype(np.uint8)).to(dev)
        colptr64, rowind = plan.pattern()
        colptr = colptr64.to(torch.int32).contiguous()
        rowind = rowind.contiguous()
        cols = torch.repeat_interleave(torch.arange(n, device=dev),
                                       (colptr64[1:] - colptr64[:-1]))
        diag_pos = torch.nonzero(rowind.to(torch.int64) == cols).flatten()
        has_nonlinear = False
        if isinstance(matLib, dict):
            has_nonlinear = any(str(m.Type).lower() == "nonlinear" for m in matLib.values())
        max_iter = int(has_nonlinear) * (self.MaxIterations - 1) + 1
        nnz = plan.nnz
        R = torch.empty(nnz, dtype=torch.float64, device=dev)
        I = torch.zeros(nnz, dtype=torch.float64, device=dev)
        Cm = torch.empty(nnz, dtype=torch.float64, device=dev)
        if enable_V:
            L.check(L.lib.mpk_csc_combine(nnz, omega, L.ptr(M_sig), 0.0, None, 0.0, None, 0.0,
                                          None, L.ptr(I), L.stream_ptr()))
        sol = torch.zeros(n, dtype=torch.complex128, device=dev)
        A_curr = torch.zeros(num_A, dtype=torch.complex128, device=dev)
        prev_diff = 1e9
        alpha = self.Relaxation
        self.LinearIterations, self.LinearResiduals = [], []
        for it in range(1, max_iter + 1):
            nu = self._effective_nu(matLib, A_curr, key)
            K = asm.assembleStiffness(space_A, nu, drop_zeros=False).values
            ref = float(K[diag_pos].abs().max()) if diag_pos.numel() else 1.0
            if ref == 0.0:
                ref = 1.0
            eps = ref * 1e-6
            L.check(L.lib.mpk_csc_combine(nnz, 1.0, L.ptr(K), eps, L.ptr(M_reg), 0.0, None, 0.0,
                                          None, L.ptr(R), L.stream_ptr()))
            L.check(L.lib.mpk_csc_combine(nnz, eps, L.ptr(M_reg), 1.0, L.ptr(I), 0.0, None, 0.0,
                                          None, L.ptr(Cm), L.stream_ptr()))
            aux, keep = self._aux(fixed_t, Cm, colptr, rowind, key)
            x = torch.zeros(n, dtype=torch.complex128, device=dev)
            iters, rel = C.c_int(0), C.c_double(0.0)
            L.check(L.lib.mpk_cocg_aux_solve(
                L.ptr(colptr), L.ptr(rowind), L.ptr(R), L.ptr(I), n, L.ptr(fixed_t),
                C.byref(aux), L.ptr(b), L.ptr(x), L.ptr(keep["work"]), self.LinearTolerance,
                self.LinearMaxIterations, self.LinearCheckEvery, C.byref(iters), C.byref(rel),
                L.stream_ptr()))
            self.LinearIterations.append(iters.value)
            self.LinearResiduals.append(rel.value)
            if it == 1:
                new = x
                diff = 1.0
                alpha = self.Relaxation
            else:
                delta = x - sol
                diff = float(torch.linalg.vector_norm(delta) /
                             (torch.linalg.vector_norm(x) + 1e-10))
                if diff < prev_diff:
                    alpha = min(alpha * self.RelaxationGrow, self.MaxRelaxation)
                else:
                    alpha = max(alpha * self.RelaxationShrink, self.MinRelaxation)
                new = sol + alpha * delta
            self.Iterations = it
            if it > 1 and diff < self.Tolerance:
                sol = new
                break
            prev_diff = diff
            sol = new
            A_curr = sol[:num_A]
        A_sol = sol[:num_A]
        V_sol = torch.zeros(num_V, dtype=torch.complex128, device=dev) if enable_V else None
        return A_sol, V_sol


def divergence_free_part(assembler: Assembler, space, W: torch.Tensor, fixed) -> torch.Tensor:
    """W~ = W - M G Ln^-1 G^T W over the free DOFs (M the unit edge mass, G the
    edge-node incidence over nodes off the Dirichlet edges, Ln = G^T M G; one
    sparse direct nodal solve on the host): the part of a winding vector that
    drives no gradient field.  Fixed entries are zero."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    mesh = assembler.Mesh
    key = space.toString()
    n = int(W.numel())
    fixed = np.asarray(fixed, dtype=bool)
    m = assembler.assembleMass(space, drop_zeros=False).to_scipy().tocsr()
    edges = mesh.Edges.cpu().numpy().astype(np.int64) - 1
    emap = assembler.DofHandler.EntityMaps[key].cpu().numpy().astype(np.int64)
    act = emap > 0
    rows = emap[act] - 1
    E = edges[act]
    dn = np.zeros(mesh.NumNodes, dtype=bool)
    dn[E[fixed[rows]].ravel()] = True
    free = np.nonzero(~dn)[0]
    fid = np.full(mesh.NumNodes, -1, dtype=np.int64)
    fid[free] = np.arange(free.size)
    a, b = fid[E[:, 0]], fid[E[:, 1]]
    ma, mb = a >= 0, b >= 0
    G = sp.csr_matrix((np.r_[-np.ones(int(ma.sum())), np.ones(int(mb.sum()))],
                       (np.r_[rows[ma], rows[mb]], np.r_[a[ma], b[mb]])), shape=(n, free.size))
    w = np.where(fixed, 0.0, W.detach().cpu().numpy())
    phi = spla.splu((G.T @ m @ G).tocsc()).solve(G.T @ w)
    wt = w - m @ (G @ phi)
    wt[fixed] = 0.0
    return torch.from_numpy(wt).to(W.device)


class FrequencyCoupledSolver(FrequencySolver):
    """FrequencyCoupledSolver (src/solver/FrequencyCoupledSolver.m:1-322): the
    time-harmonic field driven by one voltage source through a winding and an
    external impedance Z = R + 1j w L.  Its monolithic system (:107-121)
        [S  -W] [A]   [ 0]
        [-1j w W^T  -Z] [I] = [-V],   S = K(nu) + 1j w M_sigma + eps M_reg,
    eps = RegScaleRel x max|K| (AutoRegularization), is solved exactly by
    elimination: v = S^-1 W (COCG, Dirichlet rows and columns removed, W
    without its fixed entries as applyDirichlet leaves it), then
    I = V / (Z + 1j w W^T v) and A = v I.  Scalar potential: the A* gauge of
    FrequencySolver (the coil and the conductors share no nodes in the
    reference's set-ups, so W^T A* = W^T A).  The outer relaxation loop for
    nonlinear materials is the reference's, over [A; I] (the reference's
    vector also holds V_mod).

    A winding vector with a gradient part (a direction field that is not
    discretely divergence-free) drives a gradient field that only the eps
    mass holds.  Its share of the flux linkage depends on eps and on how the
    gauge treats conductors, so there the current differs from the A-V
    system's.  ProjectWinding = True uses the divergence-free part
    (divergence_free_part), with which the two gauges agree to O(eps).
    Default False: the winding vector as assembled."""

    def __init__(self, assembler, freq: float):
        super().__init__(assembler, freq)
        self.MaxIterations = 100
        self.Tolerance = 1e-5
        self.AutoRegularization = True
        self.RegScaleRel = 1e-6
        self.ProjectWinding = False

    def solve(self, space_A, space_V, matLib, sigmaMap, circuitProps, windingObj, fixedDofs_A,
              fixedDofs_V=None, init_A=None):
        """Returns (A_sol complex128 device [num_A], V_sol (zeros, or None),
        I_sol complex, info {Converged, FinalDiff, Iterations})."""
        asm = self.Assembler
        mesh = asm.Mesh
        omega = 2 * math.pi * self.Frequency
        if omega == 0:
            omega = 1e-9
        dh = asm.DofHandler
        _, offset_A = dh.distributeDofs(space_A)
        if offset_A != 0:
            raise ValueError("the edge space must be the first space distributed")
        key = space_A.toString()
        num_A = dh.SpaceLocalSizes[key]
        cond = self.getConductingRegions(sigmaMap)
        enable_V = bool(cond) and space_V is not None
        num_V = int(np.unique(mesh.T[np.isin(mesh.RegionTags, cond)]).size) if enable_V else 0
        plan = asm.plan(space_A)
        n, nnz = plan.n, plan.nnz
        dev = _dev()
        M_sig = (asm.assembleMassWeighted(space_A, sigmaMap, drop_zeros=False).values
                 if enable_V else None)
        M_reg = (asm.assembleMass(space_A, drop_zeros=False).values
                 if self.AutoRegularization else None)
        fixed = _fixed_mask(fixedDofs_A, n)
        fixed_t = torch.from_numpy(fixed.astype(np.uint8)).to(dev)
        W = asm.assembleWinding(space_A, windingObj)
        if self.ProjectWinding:
            W = divergence_free_part(asm, space_A, W, fixed)
        Wf = (W * torch.from_numpy((~fixed).astype(np.float64)).to(dev)).to(torch.complex128)
        colptr64, rowind = plan.pattern()
        colptr = colptr64.to(torch.int32).contiguous()
        rowind = rowind.contiguous()
        is_linear = not (isinstance(matLib, dict) and any(
            str(m.Type).lower() == "nonlinear" for m in matLib.values()))
        max_iter = 1 if is_linear else self.MaxIterations
        Z = complex(circuitProps.R + 1j * omega * circuitProps.L)
        Vs = complex(circuitProps.V_source)
        R = torch.empty(nnz, dtype=torch.float64, device=dev)
        Im = torch.zeros(nnz, dtype=torch.float64, device=dev)
        Cm = torch.empty(nnz, dtype=torch.float64, device=dev)
        if enable_V:
            L.check(L.lib.mpk_csc_combine(nnz, omega, L.ptr(M_sig), 0.0, None, 0.0, None, 0.0,
                                          None, L.ptr(Im), L.stream_ptr()))
        sol = torch.zeros(n + 1, dtype=torch.complex128, device=dev)
        if init_A is not None:
            ia = torch.as_tensor(init_A).to(dev, torch.complex128).reshape(-1)
            sol[: min(ia.numel(), num_A)] = ia[:num_A]
        alpha, prev_diff, diff = self.Relaxation, 1e9, 1.0
        self.LinearIterations, self.LinearResiduals = [], []
        it = 0
        for it in range(1, max_iter + 1):
            nu = self._effective_nu(matLib, sol[:num_A], key)
            K = asm.assembleStiffness(space_A, nu, drop_zeros=False).values
            eps = 0.0
            if M_reg is not None:
                ref = float(K.abs().max()) if K.numel() else 1.0
                eps = (ref if ref != 0.0 else 1.0) * self.RegScaleRel
                L.check(L.lib.mpk_csc_combine(nnz, 1.0, L.ptr(K), eps, L.ptr(M_reg), 0.0, None,
                                              0.0, None, L.ptr(R), L.stream_ptr()))
                L.check(L.lib.mpk_csc_combine(nnz, eps, L.ptr(M_reg), 1.0, L.ptr(Im), 0.0, None,
                                              0.0, None, L.ptr(Cm), L.stream_ptr()))
            else:
                R.copy_(K)
                Cm.copy_(Im)
            aux, keep = self._aux(fixed_t, Cm, colptr, rowind, key)
            v = torch.zeros(n, dtype=torch.complex128, device=dev)
            iters, rel = C.c_int(0), C.c_double(0.0)
            L.check(L.lib.mpk_cocg_aux_solve(
                L.ptr(colptr), L.ptr(rowind), L.ptr(R), L.ptr(Im), n, L.ptr(fixed_t),
                C.byref(aux), L.ptr(Wf), L.ptr(v), L.ptr(keep["work"]), self.LinearTolerance,
                self.LinearMaxIterations, self.LinearCheckEvery, C.byref(iters), C.byref(rel),
                L.stream_ptr()))
            self.LinearIterations.append(iters.value)
            self.LinearResiduals.append(rel.value)
            I = Vs / (Z + 1j * omega * complex(torch.dot(Wf, v)))
            target = torch.cat([v * I, torch.tensor([I], dtype=torch.complex128, device=dev)])
            if it == 1:
                new, diff, alpha = target, 1.0, self.Relaxation
            else:
                delta = target - sol
                diff = float(torch.linalg.vector_norm(delta) /
                             (torch.linalg.vector_norm(target) + 1e-10))
                if diff < prev_diff:
                    alpha = min(alpha * self.RelaxationGrow, self.MaxRelaxation)
                else:
                    alpha = max(alpha * self.RelaxationShrink, self.MinRelaxation)
                new = sol + alpha * delta
            if it > 1 and diff < self.Tolerance:
                sol = new
                break
            prev_diff = diff
            sol = new
        self.Iterations = it
        V_sol = torch.zeros(num_V, dtype=torch.complex128, device=dev) if enable_V else None
        info = {"Converged": (it < max_iter) or (diff < self.Tolerance), "FinalDiff": diff,
                "Iterations": it}
        return sol[:num_A], V_sol, complex(sol[n]), info


__all__ = ["FrequencySolver", "FrequencyCoupledSolver", "divergence_free_part", "MU0"]
