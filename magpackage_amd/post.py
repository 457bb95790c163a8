"""Field post-processing: element B and point probes (src/post/PostProcessor.m).

``PostProcessor(assembler)`` mirrors the reference's class for the magnetic
field:
- ``computeElementB(A_sol)`` (:64-126): B = curl A per element. It runs on the
  GPU (``mpk_element_b``), once per real column, so complex solutions and
  several columns work.
- ``probeB(A_sol, point)`` (:403-441): B in the tet that holds the point, or 0
  outside the mesh.
- ``probeLine(A_sol, start, end, numPts, "B")`` (:505-532).
- ``computeMagneticEnergy(A_sol, MatLibData)`` (:544-611, :691-760): with
  linear materials only, 0.5 Re(A^H K A) with K assembled on the GPU. Otherwise
  it sums over elements the volume times the energy density at the element's B.
  That density is 0.5 nu B^2 for linear materials, and Simpson's rule for the
  integral of H dB over [0, |B|] with H = nu(b^2) b for nonlinear ones. The
  reference sums it over the 4 quadrature points of a constant B, with weights
  adding up to the volume. nu comes from the engine's B-H evaluation
  (mpk_eval_nu, the MEX's evaluate_nu_derivative). MaterialLib.evaluate,
  which the reference calls here, resets nu to nu0 below 1e-6 rather than
  below 1e-7: a B-H table never reaches that range.
- ``probeSolution(A_sol, point)`` (:443-454, :612-635): A interpolated with the
  edge basis at the point. The reference's element's basis values are mapped by
  the covariant Piola map, x_ref = J^-1 (x - p1) and N = N_ref J^-1.
- ``probeTotalCurrent(point, A_sol, V_sol, omega, sigmaMap, sourceMap)``
  (:456-503): J = Js + sigma (-1j omega A - grad V) in conducting elements, Js
  elsewhere. With the A* gauge that FrequencySolver returns, V_sol is zero and
  J_eddy = -1j omega sigma A*. ``regions`` (an extension) restricts the point
  location to the given tags, e.g. to probe a conductor on its surface.

The reference locates points with MATLAB's ``triangulation.pointLocation``.
Here every tet's affine map is inverted once on the host. A point's reference
coordinates then give its barycentric coordinates, and the first tet (lowest
index) whose coordinates are all >= -1e-12 holds the point. A point on a face
shared by two tets can land in either of them in the reference; B is constant
per element, so such points may differ by the jump across that face.
"""
from __future__ import annotations

import numpy as np
import torch

from .assembler import Assembler, FunctionSpace


class PostProcessor:
    def __init__(self, assembler: Assembler):
        self.Assembler = assembler
        self.Mesh = assembler.Mesh
        self.DofHandler = assembler.DofHandler
        P = np.asarray(self.Mesh.P, dtype=np.float64)
        T = np.asarray(self.Mesh.T, dtype=np.int64) - 1
        self._p0 = P[T[:, 0]]
        J = np.stack([P[T[:, 1]] - self._p0, P[T[:, 2]] - self._p0, P[T[:, 3]] - self._p0],
                     axis=2)  # (Ne, 3, 3), columns = edge vectors from node 1
        self._Jinv = np.linalg.inv(J)
        self._B_cache = None

    # ------------------------------------------------------------------
    def computeCentroids(self) -> np.ndarray:
        """(3, Ne) tet centroids (:52-62)."""
        P = np.asarray(self.Mesh.P)
        T = np.asarray(self.Mesh.T) - 1
        return ((P[T[:, 0]] + P[T[:, 1]] + P[T[:, 2]] + P[T[:, 3]]) / 4.0).T

    def computeElementB(self, A_sol, spaceName: str | None = None) -> torch.Tensor:
        """(3, Ne, Ncols) device tensor of B per element, complex if A_sol is."""
        a = A_sol if torch.is_tensor(A_sol) else torch.from_numpy(np.asarray(A_sol))
        if a.dim() == 1:
            a = a[:, None]
        space = FunctionSpace("Nedelec", 1)
        if spaceName is not None and spaceName != space.toString():
            raise ValueError(f"unsupported space {spaceName}")
        cols = []
        for k in range(a.shape[1]):
            col = a[:, k]
            if col.is_complex():
                br = self.Assembler.computeElementB(col.real.contiguous(), space)
                bi = self.Assembler.computeElementB(col.imag.contiguous(), space)
                cols.append(torch.complex(br, bi).T)
            else:
                cols.append(self.Assembler.computeElementB(col.contiguous(), space).T)
        return torch.stack(cols, dim=2)

    def locate(self, points, regions=None) -> np.ndarray:
        """0-based tet index holding each point (N, 3), -1 outside the mesh
        (or outside the given region tags)."""
        pts = np.atleast_2d(np.asarray(points, dtype=np.float64))
        out = np.full(pts.shape[0], -1, dtype=np.int64)
        allowed = (None if regions is None
                   else np.isin(np.asarray(self.Mesh.RegionTags), np.atleast_1d(regions)))
        for i, p in enumerate(pts):
            x = np.einsum("eij,ej->ei", self._Jinv, p[None, :] - self._p0)
            lam_min = np.minimum(x.min(axis=1), 1.0 - x.sum(axis=1))
            ok = lam_min >= -1e-12
            if allowed is not None:
                ok &= allowed
            hit = np.nonzero(ok)[0]
            if hit.size:
                out[i] = hit[0]
        return out

    def _probe_edge_field(self, sol: np.ndarray, e: int, point) -> np.ndarray:
        """internalProbeNedelec (:612-635): sum_k sol[dof_k] s_k N_k(point)."""
        from .fem_tables import nedelec_tet_p1
        xr = self._Jinv[e] @ (np.asarray(point, dtype=np.float64) - self._p0[e])
        val, _ = nedelec_tet_p1(xr.reshape(3, 1))           # (6, 3, 1)
        N = val[:, :, 0] @ self._Jinv[e]                    # rows: physical basis
        key = FunctionSpace("Nedelec", 1).toString()
        dofs = self.DofHandler.DofMaps[key][e].cpu().numpy().astype(np.int64)
        signs = self.Mesh.T2E_Sign[e].cpu().numpy().astype(np.float64)
        coef = np.where(dofs > 0, sol[np.maximum(dofs, 1) - 1], 0.0) * signs
        return coef @ N

    def probeSolution(self, A_sol, point):
        """(A_val (3,), elem_idx 1-based; NaN and zeros outside the mesh)."""
        e = int(self.locate(point)[0])
        if e < 0:
            return np.zeros(3), float("nan")
        a = A_sol.detach().cpu().numpy() if torch.is_tensor(A_sol) else np.asarray(A_sol)
        return self._probe_edge_field(a, e, point), e + 1

    def probeTotalCurrent(self, point, A_sol, V_sol, omega, sigmaMap, sourceMap,
                          regions=None):
        """(J_total (3,), elem_idx 1-based).  V_sol must be None or zero: only
        the edge field's gauge (A*) is solved here."""
        from .assembler import expand_material_map, expand_source_map
        e = int(self.locate(point, regions)[0])
        if e < 0:
            return np.zeros(3), float("nan")
        if V_sol is not None:
            v = V_sol.detach().cpu().numpy() if torch.is_tensor(V_sol) else np.asarray(V_sol)
            if np.any(v != 0):
                raise NotImplementedError("a nonzero scalar potential is not supported")
        tags = np.asarray(self.Mesh.RegionTags)
        js = expand_source_map(sourceMap, tags)[e]
        sigma = expand_material_map(sigmaMap, tags)[e]
        if abs(sigma) <= 1e-12:
            return js.astype(np.complex128), e + 1
        a = A_sol.detach().cpu().numpy() if torch.is_tensor(A_sol) else np.asarray(A_sol)
        A = self._probe_edge_field(a, e, point)
        return js + sigma * (-1j * omega * A), e + 1

    def probeB(self, A_sol, point):
        """(B_val, elem_idx): B (3,) or (3, Ncols) at the point, elem_idx
        1-based (NaN outside, where B is 0), as the reference returns."""
        e = int(self.locate(point)[0])
        a = A_sol if torch.is_tensor(A_sol) else torch.from_numpy(np.asarray(A_sol))
        ncols = 1 if a.dim() == 1 else a.shape[1]
        if e < 0:
            return np.zeros((3, ncols)).squeeze(), float("nan")
        B = self.computeElementB(a)[:, e, :].cpu().numpy()
        return (B[:, 0] if ncols == 1 else B), e + 1

    def computeMagneticEnergy(self, A_sol, MatLibData) -> float:
        """Magnetic energy [J] of the field A_sol (real or complex)."""
        import ctypes as C
        from . import _lib as L
        mats = dict(MatLibData)
        space = FunctionSpace("Nedelec", 1)
        a = A_sol if torch.is_tensor(A_sol) else torch.from_numpy(np.asarray(A_sol))
        dev = self.Mesh.P_dev.device
        a = a.to(dev)
        if all(m.Type == "Linear" for m in mats.values()):
            K = self.Assembler.assembleStiffness(space, {t: m.Nu_Linear for t, m in mats.items()})
            colptr = K.colptr.to(torch.int32).contiguous()
            n = K.shape[0]
            parts = [a.real, a.imag] if a.is_complex() else [a]
            energy = 0.0
            for v in parts:
                x = v.to(torch.float64).contiguous()
                y = torch.empty(n, dtype=torch.float64, device=dev)
                L.check(L.lib.mpk_spmv_sym(L.ptr(colptr), L.ptr(K.rowind), L.ptr(K.values), n,
                                           L.ptr(x), L.ptr(y), L.stream_ptr()))
                energy += 0.5 * float(x @ y)   # Re(A^H K A) for symmetric real K
            return energy
        B = self.computeElementB(a)[:, :, 0]
        bmag = torch.sqrt((B.abs() ** 2).sum(dim=0)).to(torch.float64).contiguous()
        ne = bmag.numel()
        mat, keep = self.Assembler.material_tables(mats)
        tags = self.Mesh.Tags_dev
        nu_half = torch.empty(ne, dtype=torch.float64, device=dev)
        nu_full = torch.empty_like(nu_half)
        for s_, out in (((0.5 * bmag) ** 2, nu_half), (bmag ** 2, nu_full)):
            s_ = s_.contiguous()
            L.check(L.lib.mpk_eval_nu(L.ptr(s_), L.ptr(tags), ne, C.byref(mat), L.ptr(out),
                                      None, L.stream_ptr()))
        lin = np.zeros(int(self.Mesh.RegionTags.max()) + 1, dtype=bool)
        for t, m in mats.items():
            if int(t) < lin.size and m.Type == "Linear":
                lin[int(t)] = True
        is_lin = torch.from_numpy(lin[self.Mesh.RegionTags]).to(dev)
        simpson = (bmag / 6.0) * (4.0 * nu_half * (0.5 * bmag) + nu_full * bmag)
        dens = torch.where(is_lin, 0.5 * nu_full * bmag ** 2, simpson)
        vol = torch.from_numpy(1.0 / np.abs(np.linalg.det(self._Jinv)) / 6.0).to(dev)
        return float((vol * dens).sum())

    def probeLine(self, A_sol, startPt, endPt, numPts: int, type_: str = "B") -> np.ndarray:
        """(numPts, 3) samples on the segment (linspace per coordinate)."""
        if str(type_).upper() != "B":
            raise NotImplementedError("only 'B' probes are provided")
        s = np.asarray(startPt, dtype=np.float64)
        e = np.asarray(endPt, dtype=np.float64)
        pts = np.stack([np.linspace(s[k], e[k], numPts) for k in range(3)], axis=1)
        idx = self.locate(pts)
        a = A_sol if torch.is_tensor(A_sol) else torch.from_numpy(np.asarray(A_sol))
        B = self.computeElementB(a)[:, :, 0].cpu().numpy()   # (3, Ne)
        out = np.zeros((numPts, 3), dtype=B.dtype)
        ok = idx >= 0
        out[ok] = B[:, idx[ok]].T
        return out
