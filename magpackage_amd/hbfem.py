"""Harmonic balance (HBFEM) host pieces: the alternating frequency-time
transform (src/utils/AFT.m:19-75) and the assembly entry point that the
reference's HBFEM solvers call (HBFEMSolver.m:78,96: Assembler.assembleHBFEM,
absent from the reference's Assembler) on the engine's HBFEM kernel
(mpk_hbfem_coo, the restatement of src/mex/assemble_hbfem_kernel_mex.cpp).
"""
from __future__ import annotations

import math
import warnings

import numpy as np


class AFT:
    """D_matrix [Nt x K] = exp(1j h theta) (frequency -> time), P_matrix
    [K x Nt] = f_h exp(-1j h theta)^T with f_h = 2/Nt (1/Nt for h = 0), theta =
    2 pi (0..Nt-1)/Nt.  Default Nt = max(32, next power of two >= 2 max(h) + 1);
    a given Nt below 2 max(h) + 1 is raised to it with a warning."""

    def __init__(self, harmonics, n_time_steps=None, base_freq: float = 50.0):
        self.Harmonics = np.asarray(harmonics, dtype=np.float64).ravel()
        self.NumHarmonics = int(self.Harmonics.size)
        min_steps = int(2 * self.Harmonics.max() + 1)
        if n_time_steps is None:
            self.NumTimeSteps = max(32, 1 << max(0, math.ceil(math.log2(min_steps))))
        elif np.size(n_time_steps) != 1:
            warnings.warn("TimeSteps argument is a vector: using the default steps")
            self.NumTimeSteps = max(32, 1 << max(0, math.ceil(math.log2(min_steps))))
        elif int(n_time_steps) < min_steps:
            warnings.warn(f"TimeSteps ({int(n_time_steps)}) is insufficient for the harmonics: "
                          f"adjusted to {min_steps}")
            self.NumTimeSteps = min_steps
        else:
            self.NumTimeSteps = int(n_time_steps)
        self.BaseFreq = float(base_freq)
        N = self.NumTimeSteps
        theta = np.arange(N) * (2 * np.pi / N)
        self.D_matrix = np.exp(1j * np.outer(theta, self.Harmonics))
        f = np.where(self.Harmonics == 0, 1.0 / N, 2.0 / N)
        self.P_matrix = f[:, None] * np.exp(-1j * np.outer(self.Harmonics, theta))

    def freq2time(self, coeffs):
        return np.real(self.D_matrix @ coeffs)

    def time2freq(self, vals):
        return self.P_matrix @ vals

    def jacobian_scalings(self) -> np.ndarray:
        """Block scalings of the DC-reluctivity Jacobian: 1 for h = 0, 2 for AC
        (assemble_hbfem_kernel.m:28-31)."""
        return np.where(self.Harmonics > 0, 2.0, 1.0)


# ---------------------------------------------------------------------------
# Harmonic-balance Newton solvers on the GPU engine
# ---------------------------------------------------------------------------

def _csc_parts(mat):
    """(colptr int32, rowind int32, vals float64) device tensors of a
    CscMatrix (this package's) or a scipy sparse matrix."""
    import torch
    from .assembler import _dev
    dev = _dev()
    if hasattr(mat, "colptr"):
        return (mat.colptr.to(torch.int32).contiguous(), mat.rowind.to(torch.int32).contiguous(),
                mat.values.to(torch.float64).contiguous())
    m = mat.tocsc()
    m.sort_indices()
    return (torch.from_numpy(m.indptr.astype(np.int32)).to(dev),
            torch.from_numpy(m.indices.astype(np.int32)).to(dev),
            torch.from_numpy(m.data.astype(np.float64)).to(dev))


class _HarmonicBlocks:
    """Per-harmonic real Jacobian blocks J_h (+ K_add) of the block-diagonal
    HBFEM Jacobian, in CSC on the device, with their linear solves: the
    engine's auxiliary-space CG for real right-hand sides and COCG (real
    matrix, complex vectors) for complex ones.  The auxiliary space is built
    from K_add (the regularising mass-type term; Jacobi only without it)."""

    def __init__(self, solver, nd, H, fixed_np, kadd):
        import ctypes as C
        import torch
        from . import _lib as L
        from .assembler import _dev
        from .transient import build_aux_space
        self.s, self.nd, self.H = solver, nd, H
        self.dev = _dev()
        self.fixed = torch.from_numpy(fixed_np.astype(np.uint8)).to(self.dev)
        self.free = (self.fixed == 0).to(torch.float64)
        self.kadd = kadd
        self.aux, self.keep, nv = None, {}, 0
        self.kadd_coo = None
        if kadd is not None:
            cp, ri, kv = kadd
            cols = torch.repeat_interleave(torch.arange(nd, device=self.dev, dtype=torch.float64),
                                           (cp[1:] - cp[:-1]).to(torch.int64))
            self.kadd_coo = (ri.to(torch.float64) + 1.0, cols + 1.0, kv)
            self.aux, self.keep = build_aux_space(solver.Assembler, solver._space_key, fixed_np,
                                                  kv, cp, ri, solver.AuxSweeps, solver.AuxOmega)
            nv = self.aux.num_nodes
        ws = C.c_int64()
        L.check(L.lib.mpk_cocg_aux_work_size(nd, nv, C.byref(ws)))
        self.cwork = torch.empty(ws.value, dtype=torch.float64, device=self.dev)
        L.check(L.lib.mpk_pcg_aux_work_size(nd, nv, C.byref(ws)))
        self.pwork = torch.empty(ws.value, dtype=torch.float64, device=self.dev)
        self.blocks = []
        self.pcg_iterations = 0

    def set_jacobian(self, J_tri):
        """Splits the global triplets into the H blocks and builds their CSC
        (sparse() semantics: duplicates summed in triplet order, zeros dropped)."""
        import ctypes as C
        import torch
        from . import _lib as L
        nd = self.nd
        I, J, V = J_tri["I"], J_tri["J"], J_tri["V"]
        self.blocks = []
        for h in range(self.H):
            off = float(h * nd)
            m = (I > off) & (I <= off + nd)
            parts_i, parts_j, parts_v = [I[m] - off], [J[m] - off], [V[m]]
            if self.kadd_coo is not None:
                parts_i.append(self.kadd_coo[0])
                parts_j.append(self.kadd_coo[1])
                parts_v.append(self.kadd_coo[2])
            bi = torch.cat(parts_i).contiguous()
            bj = torch.cat(parts_j).contiguous()
            bv = torch.cat(parts_v).contiguous()
            nt = int(bi.numel())
            colptr = torch.empty(nd + 1, dtype=torch.int64, device=self.dev)
            rowind = torch.empty(max(nt, 1), dtype=torch.int32, device=self.dev)
            vals = torch.empty(max(nt, 1), dtype=torch.float64, device=self.dev)
            nnz = C.c_int64(0)
            L.check(L.lib.mpk_sparse_coo_to_csc(L.ptr(bi), L.ptr(bj), L.ptr(bv), nt, nd, nd, 1,
                                                L.ptr(colptr), L.ptr(rowind), L.ptr(vals), nt,
                                                C.byref(nnz), L.stream_ptr()))
            k = int(nnz.value)
            self.blocks.append((colptr.to(torch.int32).contiguous(), rowind[:k].contiguous(),
                                vals[:k].contiguous(), torch.zeros(max(k, 1), dtype=torch.float64,
                                                                    device=self.dev)))

    def solve_complex(self, h, b):
        """x = J_h^-1 b (Dirichlet rows/columns removed), b complex [nd]."""
        import ctypes as C
        import torch
        from . import _lib as L
        cp, ri, v, z = self.blocks[h]
        x = torch.empty(self.nd, dtype=torch.complex128, device=self.dev)
        bb = (b * self.free).to(torch.complex128).contiguous()
        it, rel = C.c_int(), C.c_double()
        L.check(L.lib.mpk_cocg_aux_solve(L.ptr(cp), L.ptr(ri), L.ptr(v), L.ptr(z), self.nd,
                                         L.ptr(self.fixed),
                                         C.byref(self.aux) if self.aux is not None else None,
                                         L.ptr(bb), L.ptr(x), L.ptr(self.cwork),
                                         self.s.LinearTolerance, self.s.LinearMaxIterations, 10,
                                         C.byref(it), C.byref(rel), L.stream_ptr()))
        self.pcg_iterations += it.value
        return x

    def solve_real(self, h, b):
        import ctypes as C
        import torch
        from . import _lib as L
        cp, ri, v, _ = self.blocks[h]
        x = torch.empty(self.nd, dtype=torch.float64, device=self.dev)
        bb = (b * self.free).contiguous()
        it, rel = C.c_int(), C.c_double()
        if self.aux is not None:
            L.check(L.lib.mpk_pcg_aux_solve(L.ptr(cp), L.ptr(ri), L.ptr(v), self.nd,
                                            L.ptr(self.fixed), C.byref(self.aux), L.ptr(bb),
                                            L.ptr(x), L.ptr(self.keep["work"]),
                                            self.s.LinearTolerance, self.s.LinearMaxIterations,
                                            10, C.byref(it), C.byref(rel), L.stream_ptr()))
        else:
            L.check(L.lib.mpk_pcg_solve(L.ptr(cp), L.ptr(ri), L.ptr(v), self.nd,
                                        L.ptr(self.fixed), L.ptr(bb), L.ptr(x), L.ptr(self.pwork),
                                        self.s.LinearTolerance, self.s.LinearMaxIterations, 10,
                                        C.byref(it), C.byref(rel), L.stream_ptr()))
        self.pcg_iterations += it.value
        return x


class HBFEMSolver:
    """HBFEMSolver (src/solver/HBFEMSolver.m:1-185): Newton on the harmonic-
    balance residual R(X) - F (+ K_add X) with the block-diagonal
    DC-reluctivity Jacobian of assembleHBFEM, backtracking line search, stop
    on |Res(free)| / |F| < Tolerance.  The reference factors the Jacobian with
    MUMPS; here each harmonic's real block is solved on the GPU (COCG for the
    complex right-hand side, auxiliary-space preconditioner from MatrixK_Add).
    Same properties and defaults."""

    def __init__(self, assembler, aft):
        self.Assembler = assembler
        self.AFT = aft
        self.MaxIter = 50
        self.Tolerance = 1e-4
        self.MaxBacktracks = 20
        self.BacktrackFactor = 0.5
        self.MinStepSize = 1e-6
        self.MatrixK_Add = None
        self.LinearTolerance = 1e-10
        self.LinearMaxIterations = 40000
        self.AuxSweeps = 2
        self.AuxOmega = 2.0 / 3.0

    def _setup(self, space, fixedDofs):
        import torch
        from .assembler import _dev
        asm = self.Assembler
        self._space_key = space.toString()
        nd = int(asm.DofHandler.NumGlobalDofs)
        H = int(self.AFT.NumHarmonics)
        fx = np.zeros(nd, dtype=bool)
        if fixedDofs is not None:
            f = fixedDofs.cpu().numpy() if torch.is_tensor(fixedDofs) else np.asarray(fixedDofs)
            if f.dtype == bool:
                fx[: min(nd, f.size)] = f.ravel()[:nd]
            else:
                idx = f.astype(np.int64).ravel()
                fx[idx[(idx >= 1) & (idx <= nd)] - 1] = True
        kadd = _csc_parts(self.MatrixK_Add) if self.MatrixK_Add is not None else None
        self._blocks = _HarmonicBlocks(self, nd, H, fx, kadd)
        self._free = torch.from_numpy((~fx).astype(np.float64)).to(_dev())
        return nd, H

    def _kadd_times(self, X):
        """K_add X column by column (symmetric CSC product on the device)."""
        import torch
        from . import _lib as L
        if self._blocks.kadd is None:
            return torch.zeros_like(X)
        cp, ri, kv = self._blocks.kadd
        out = torch.empty_like(X)
        for part in ("real", "imag"):
            src = getattr(X, part).t().contiguous()      # [H, nd]
            dst = torch.empty_like(src)
            for h in range(X.shape[1]):
                L.check(L.lib.mpk_spmv_sym(L.ptr(cp), L.ptr(ri), L.ptr(kv), X.shape[0],
                                           L.ptr(src[h]), L.ptr(dst[h]), L.stream_ptr()))
            getattr(out, part).copy_(dst.t())
        return out

    def _masked_norm(self, R) -> float:
        import torch
        return float(torch.linalg.vector_norm(R * self._free[:, None]))

    def solve(self, space, matLibData, sourceMaps, fixedDofs, x0=None):
        """Returns (X_harmonics [NumGlobalDofs x H] complex128 device tensor,
        info {Iterations, Residuals, Converged, LinearIterations})."""
        import torch
        from .assembler import _dev
        asm = self.Assembler
        nd, H = self._setup(space, fixedDofs)
        dev = _dev()
        X = (torch.zeros((nd, H), dtype=torch.complex128, device=dev) if x0 is None
             else torch.as_tensor(x0).to(dev, torch.complex128).clone())
        F = torch.zeros((nd, H), dtype=torch.complex128, device=dev)
        for k in range(H):
            if sourceMaps is not None and k < len(sourceMaps) and sourceMaps[k] is not None:
                F[:, k] = asm.assembleSource(space, sourceMaps[k]).to(torch.complex128)
        norm_base = float(torch.linalg.vector_norm(F))
        if norm_base < 1e-10:
            norm_base = 1.0

        def residual(Xc, with_j):
            J_tri, R = asm.assembleHBFEM(space, Xc, self.AFT, matLibData, with_j)
            return J_tri, R + self._kadd_times(Xc) - F

        _, res = residual(X, False)
        cur = self._masked_norm(res)
        hist = [cur]
        converged = False
        it = 0
        for it in range(1, self.MaxIter + 1):
            J_tri, res = residual(X, True)
            self._blocks.set_jacobian(J_tri)
            dX = torch.stack([self._blocks.solve_complex(h, -res[:, h]) for h in range(H)], dim=1)
            alpha, accepted = 1.0, False
            for _ in range(self.MaxBacktracks + 1):
                Xt = X + alpha * dX
                _, rt = residual(Xt, False)
                new = self._masked_norm(rt)
                if not np.isfinite(new):
                    new = float("inf")
                if new < cur or alpha < self.MinStepSize:
                    X, cur, accepted = Xt, new, True
                    break
                alpha *= self.BacktrackFactor
            if not accepted:
                # as HBFEMSolver.m:161-164: a forced small step, its norm taken
                # over every row of R - F
                X = X + 1e-3 * dX
                _, R = asm.assembleHBFEM(space, X, self.AFT, matLibData, False)
                cur = float(torch.linalg.vector_norm(R - F))
            hist.append(cur)
            if cur / norm_base < self.Tolerance:
                converged = True
                break
        return X, {"Iterations": it, "Residuals": hist, "Converged": converged,
                   "LinearIterations": self._blocks.pcg_iterations}


class HBFEMCoupledSolver(HBFEMSolver):
    """HBFEMCoupledSolver (src/solver/HBFEMCoupledSolver.m:1-354): the
    harmonic-balance field coupled to one voltage-driven winding through a
    resistor, unknowns A [nd x H] and the current phasors I [H].  Field rows
    R(A) - C I_h (+ K_add A), circuit rows (R I_h + 1j w_h C^T A_h - V_h) x
    row_scale, row_scale = mean|field residual| / |R| at the start, clamped to
    [1e3, 1e9]; backtracking as HBFEMSolver, stop on |Res(free)| / |Res_0|.
    The monolithic Jacobian is block-diagonal over harmonics in the field and
    couples each harmonic to its own current only, so the Newton step is
    exact harmonic by harmonic: u = J_h^-1 Res_F,h, v = J_h^-1 C (C without
    its fixed entries, as applyDirichlet leaves it), then
    dI_h = (-Res_C,h + 1j w_h C^T u) / (1j w_h C^T v + R), dA_h = -u + v dI_h;
    the row scale cancels out of the step."""

    def __init__(self, assembler, aft, winding, R):
        super().__init__(assembler, aft)
        self.WindingObj = winding
        self.CircuitR = float(R)

    def solve(self, space, matLibData, V_harmonics, fixedDofs, x0=None):
        """Returns (X_sol [nd x H] complex128 device tensor, I_sol [H] complex
        numpy, info {Iterations, Converged, Residuals, LinearIterations})."""
        import torch
        from .assembler import _dev
        asm = self.Assembler
        nd, H = self._setup(space, fixedDofs)
        dev = _dev()
        w = 2 * np.pi * self.AFT.BaseFreq * self.AFT.Harmonics
        Vh = np.asarray(V_harmonics, dtype=np.complex128).ravel()
        A = (torch.zeros((nd, H), dtype=torch.complex128, device=dev) if x0 is None
             else torch.as_tensor(x0).to(dev, torch.complex128).clone())
        I = np.zeros(H, dtype=np.complex128)
        Cv = asm.assembleWinding(space, self.WindingObj).to(torch.float64)
        Cc = Cv.to(torch.complex128)
        Cfree = Cv * self._free
        Rc = self.CircuitR

        def residuals(Ac, Ic, with_j):
            J_tri, R = asm.assembleHBFEM(space, Ac, self.AFT, matLibData, with_j)
            rf = R - Cc[:, None] * torch.from_numpy(Ic).to(dev)[None, :] + self._kadd_times(Ac)
            psi = (Cc[:, None] * Ac).sum(dim=0).cpu().numpy()
            rcir = Rc * Ic + 1j * w * psi - Vh
            return J_tri, rf, rcir

        _, rf, rcir = residuals(A, I, False)
        mag = float(rf.abs().mean())
        if mag < 1e-5:
            mag = 1.0
        base = abs(Rc) if abs(Rc) >= 1e-10 else 1.0
        self.RowScale = rs = min(max(mag / base, 1e3), 1e9)

        def total_norm(rf_, rc_):
            return math.sqrt(self._masked_norm(rf_) ** 2 + float(np.sum(np.abs(rc_ * rs) ** 2)))

        cur = total_norm(rf, rcir)
        norm_base = cur if cur >= 1e-5 else 1.0
        hist = [cur]
        converged = False
        it = 0
        for it in range(1, self.MaxIter + 1):
            J_tri, rf, rcir = residuals(A, I, True)
            self._blocks.set_jacobian(J_tri)
            dA = torch.empty_like(A)
            dI = np.zeros(H, dtype=np.complex128)
            for h in range(H):
                u = self._blocks.solve_complex(h, rf[:, h])
                v = self._blocks.solve_real(h, Cfree)
                cu = complex(torch.dot(Cfree.to(torch.complex128), u))
                cv = float(torch.dot(Cfree, v))
                dI[h] = (-rcir[h] + 1j * w[h] * cu) / (1j * w[h] * cv + Rc)
                dA[:, h] = -u + v.to(torch.complex128) * complex(dI[h])
            alpha, accepted = 1.0, False
            for _ in range(self.MaxBacktracks + 1):
                At, It = A + alpha * dA, I + alpha * dI
                _, rft, rct = residuals(At, It, False)
                new = total_norm(rft, rct)
                if new < cur or alpha < self.MinStepSize:
                    A, I, cur, accepted = At, It, new, True
                    break
                alpha *= self.BacktrackFactor
            if not accepted:
                A, I = A + 1e-3 * dA, I + 1e-3 * dI
            hist.append(cur)
            if cur / norm_base < self.Tolerance:
                converged = True
                break
        return A, I, {"Iterations": it, "Converged": converged, "Residuals": hist,
                      "LinearIterations": self._blocks.pcg_iterations}
