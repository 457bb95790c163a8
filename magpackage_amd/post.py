"""Field post-processing: element B and point probes (src/post/PostProcessor.m).

``PostProcessor(assembler)`` mirrors the reference's class for the magnetic
field:
- ``computeElementB(A_sol)`` (:64-126): B = curl A per element. It runs on the
  GPU (``mpk_element_b``), once per real column, so complex solutions and
  several columns work.
- ``probeB(A_sol, point)`` (:403-441): B in the tet that holds the point, or 0
  outside the mesh.
- ``probeLine(A_sol, start, end, numPts, "B")`` (:505-532).
- ``computeElementJ(A_sol, V_sol, omega, sigmaMap, sourceMap)`` (:128-275): J at
  the element centroids, on the device.
- ``mapElementsToNodes`` / ``combineRegionData`` (:277-401): per-region nodal
  averages of element values, and their combination over regions.
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

    def computeElementJ(self, A_sol, V_sol, omega, sigmaMap, sourceMap) -> torch.Tensor:
        """(3, Ne) complex device tensor of the current density at the element
        centroids (:128-275): J = Js + sigma (-1j omega A - grad V) where
        |sigma| > 1e-12, Js elsewhere. A is the edge field at the centroid
        (reference basis values at (1/4, 1/4, 1/4) mapped by J^-T), and grad V
        the P1 gradient of V_sol (given on the Lagrange_P1 DOFs, numbered from
        that space's offset), taken only where all four of the element's V
        DOFs exist. The reference's lookup rules are kept. A numeric sigma
        vector is indexed by the element when it has at least e entries, else
        by the tag. A [3 x Ne] source is per element, a map per tag."""
        from .fem_tables import lagrange_tet_p1, nedelec_tet_p1
        dev = self.Mesh.P_dev.device
        ne = self.Mesh.NumElements
        tags = np.asarray(self.Mesh.RegionTags, dtype=np.int64)
        sig = element_sigma(sigmaMap, tags)
        src = np.asarray(sourceMap) if not isinstance(sourceMap, dict) else None
        if src is not None and src.ndim == 2 and src.shape[1] == ne:
            js = src.T.astype(np.float64)
        else:
            js = np.zeros((ne, 3))
            for t, v in (sourceMap.items() if isinstance(sourceMap, dict) else []):
                js[tags == int(t)] = np.asarray(v, dtype=np.float64).ravel()[:3]
        cond = np.abs(sig) > 1e-12
        J = torch.from_numpy(js.astype(np.complex128)).to(dev)
        if cond.any():
            ctr = np.full((3, 1), 0.25)
            val = torch.from_numpy(nedelec_tet_p1(ctr)[0][:, :, 0]).to(dev)      # (6, 3)
            grad = torch.from_numpy(lagrange_tet_p1(ctr)[1][:, :, 0]).to(dev)    # (4, 3)
            invJt = torch.from_numpy(np.transpose(self._Jinv, (0, 2, 1)).copy()).to(dev)
            key = FunctionSpace("Nedelec", 1).toString()
            dofs = self.DofHandler.DofMaps[key].to(torch.int64)
            signs = self.Mesh.T2E_Sign.to(torch.float64)
            a = A_sol if torch.is_tensor(A_sol) else torch.from_numpy(np.asarray(A_sol))
            a = a.to(dev).to(torch.complex128)
            coef = torch.where(dofs > 0, a[(dofs - 1).clamp(min=0)], 0) * signs
            A_ref = coef @ val.to(torch.complex128)                               # (Ne, 3)
            A_c = torch.einsum("eij,ej->ei", invJt.to(torch.complex128), A_ref)
            gV = torch.zeros_like(A_c)
            if V_sol is not None:
                kv = FunctionSpace("Lagrange", 1).toString()
                v = V_sol if torch.is_tensor(V_sol) else torch.from_numpy(np.asarray(V_sol))
                v = v.to(dev).to(torch.complex128)
                off = self.DofHandler.SpaceOffsets.get(kv, 0)
                loc = self.DofHandler.DofMaps[kv].to(torch.int64) - off
                ok = ((loc > 0) & (loc <= v.numel())).all(dim=1)
                vc = v[(loc - 1).clamp(0, max(v.numel() - 1, 0))]
                g_ref = vc @ grad.to(torch.complex128)
                gV = torch.where(ok[:, None], torch.einsum("eij,ej->ei",
                                                           invJt.to(torch.complex128), g_ref), 0)
            sg = torch.from_numpy(sig).to(dev)
            eddy = sg[:, None] * (-1j * omega * A_c - gV)
            J = torch.where(torch.from_numpy(cond).to(dev)[:, None], J + eddy, J)
        return J.T

    def mapElementsToNodes(self, Val_elems) -> dict:
        """{tag: (dims, Nn) nodal values}: per region, each node's value is the
        mean of the element values of that region's elements around it, 0 at
        nodes outside the region (:277-315)."""
        v = Val_elems if torch.is_tensor(Val_elems) else torch.from_numpy(np.asarray(Val_elems))
        if v.dim() == 1:
            v = v[None, :]
        ne, nn = self.Mesh.NumElements, self.Mesh.NumNodes
        if v.shape[1] != ne:
            raise ValueError(f"PostProcessor:DimensionMismatch Val_elems columns ({v.shape[1]}) "
                             f"must match mesh elements ({ne}).")
        dev = self.Mesh.P_dev.device
        v = v.to(dev)
        T0 = self.Mesh.T_dev.view(ne, 4).to(torch.int64) - 1
        tags = self.Mesh.Tags_dev
        out = {}
        for tag in np.unique(np.asarray(self.Mesh.RegionTags)):
            m = tags == int(tag)
            Ts = T0[m].reshape(-1)
            vs = v[:, m].repeat_interleave(4, dim=1)
            sums = torch.zeros((v.shape[0], nn), dtype=v.dtype, device=dev).index_add_(1, Ts, vs)
            cnt = torch.zeros(nn, dtype=torch.float64, device=dev).index_add_(
                0, Ts, torch.ones(Ts.numel(), dtype=torch.float64, device=dev))
            out[int(tag)] = torch.where(cnt > 0, sums / cnt.clamp(min=1).to(sums.dtype), 0)
        return out

    def combineRegionData(self, RegionDataMap: dict, targetRegions) -> torch.Tensor:
        """Nodal values of the target regions in order, a later region's nonzero
        nodes overwriting earlier ones (:360-401)."""
        nn = self.Mesh.NumNodes
        dev = self.Mesh.P_dev.device
        present = [RegionDataMap[int(t)] for t in np.atleast_1d(targetRegions)
                   if int(t) in RegionDataMap]
        dims = present[0].shape[0] if present else 1
        dt = present[0].dtype if present else torch.float64
        out = torch.zeros((dims, nn), dtype=dt, device=dev)
        for vals in present:
            mask = (vals != 0)[0] if dims == 1 else vals.abs().sum(dim=0) > 0
            out[:, mask] = vals[:, mask]
        return out

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


def element_sigma(sigmaMap, tags) -> np.ndarray:
    """Per-element sigma as computeElementJ looks it up (PostProcessor.m:185-194):
    a map by tag (0 when absent); a numeric vector by the element index e
    (1-based) when it has at least e entries, else by the tag when it has at
    least tag entries, else 0."""
    tags = np.asarray(tags, dtype=np.int64)
    ne = tags.size
    if isinstance(sigmaMap, dict):
        return np.array([float(sigmaMap.get(int(t), 0.0)) for t in tags])
    vec = np.asarray(sigmaMap, dtype=np.float64).ravel()
    if vec.size == 0:
        return np.zeros(ne)
    e1 = np.arange(1, ne + 1)
    return np.where(vec.size >= e1, vec[np.minimum(e1, vec.size) - 1],
                    np.where(vec.size >= tags, vec[np.clip(tags, 1, vec.size) - 1], 0.0))


def _element_volumes(mesh) -> np.ndarray:
    """|(v2 - v1) x (v3 - v1) . (v4 - v1)| / 6 per element (LossCalculator.m:106-114)."""
    P = np.asarray(mesh.P, dtype=np.float64)
    T = np.asarray(mesh.T, dtype=np.int64) - 1
    d1 = P[T[:, 1]] - P[T[:, 0]]
    d2 = P[T[:, 2]] - P[T[:, 0]]
    d3 = P[T[:, 3]] - P[T[:, 0]]
    return np.abs((np.cross(d1, d2) * d3).sum(axis=1)) / 6.0


class LossCalculator:
    """LossCalculator (src/post/LossCalculator.m:1-150) on the GPU engine:
    ohmic losses 0.5 w^2 Re(A^H M_sigma A) with M_sigma assembled by the
    engine, and Steinmetz / Bertotti iron losses from the element B of
    computeElementB and the element volumes."""

    def __init__(self, assembler: Assembler):
        self.Assembler = assembler
        self.PostProcessor = PostProcessor(assembler)

    def _quad_form(self, vals, colptr, rowind, a) -> float:
        """Re(a^H S a) for a symmetric real CSC S on the plan's pattern."""
        from . import _lib as L
        n = colptr.numel() - 1
        dev = vals.device
        parts = [a.real, a.imag] if a.is_complex() else [a]
        total = 0.0
        for v in parts:
            x = v.to(dev, torch.float64).contiguous()
            y = torch.empty(n, dtype=torch.float64, device=dev)
            L.check(L.lib.mpk_spmv_sym(L.ptr(colptr), L.ptr(rowind), L.ptr(vals), n, L.ptr(x),
                                       L.ptr(y), L.stream_ptr()))
            total += float(x @ y)
        return total

    def _m_sigma(self, space, sigmaMap):
        m = self.Assembler.assembleMassWeighted(space, sigmaMap, drop_zeros=False)
        return m.values, m.colptr.to(torch.int32).contiguous(), m.rowind.contiguous()

    @staticmethod
    def _col(A, k=None):
        a = A if torch.is_tensor(A) else torch.from_numpy(np.asarray(A))
        if a.dim() == 2:
            a = a[:, 0 if k is None else k]
        return a

    def computeOhmicLoss_Frequency(self, A_phasor, sigmaMap, freq, space) -> float:
        """0.5 w^2 Re(A^H M_sigma A) (:20-28)."""
        omega = 2.0 * np.pi * freq
        vals, cp, ri = self._m_sigma(space, sigmaMap)
        return 0.5 * omega ** 2 * self._quad_form(vals, cp, ri, self._col(A_phasor))

    def computeOhmicLoss_HBFEM(self, X_harmonics, sigmaMap, aft, space) -> float:
        """Sum over the non-DC harmonics of 0.5 (h w0)^2 A_h^H M_sigma A_h (:30-51)."""
        w0 = 2.0 * np.pi * aft.BaseFreq
        vals, cp, ri = self._m_sigma(space, sigmaMap)
        total = 0.0
        for k, h in enumerate(np.asarray(aft.Harmonics).ravel()):
            if h == 0:
                continue
            w = h * w0
            total += 0.5 * w ** 2 * self._quad_form(vals, cp, ri, self._col(X_harmonics, k))
        return total

    def computeElementVolumes(self) -> np.ndarray:
        return _element_volumes(self.Assembler.Mesh)

    def _b_mag(self, a) -> np.ndarray:
        B = self.PostProcessor.computeElementB(a)[:, :, 0]
        return torch.sqrt((B.abs() ** 2).sum(dim=0)).cpu().numpy()

    def computeIronLoss_Steinmetz(self, A_sol, freq, k_h, alpha, k_e, space) -> float:
        """sum(vol (k_h f B^alpha + k_e f^2 B^2)), B the element |B| of the
        first column (:53-70)."""
        B = self._b_mag(self._col(A_sol))
        p = k_h * freq * B ** alpha + k_e * freq ** 2 * B ** 2
        return float((p * self.computeElementVolumes()).sum())

    def computeIronLoss_HBFEM(self, A_sol, aft, k_h, alpha, k_e, space) -> float:
        """Steinmetz per non-DC harmonic at its frequency, summed (:72-104)."""
        vols = self.computeElementVolumes()
        total = 0.0
        for k, h in enumerate(np.asarray(aft.Harmonics).ravel()):
            if h == 0:
                continue
            f = h * aft.BaseFreq
            B = self._b_mag(self._col(A_sol, k))
            total += float(((k_h * f * B ** alpha + k_e * f ** 2 * B ** 2) * vols).sum())
        return total

    def computeBertottiLoss_TimeStep(self, dA_dt_col, matLib) -> dict:
        """Classic (Lam_sigma Lam_d^2 / 12 |dB/dt|^2) and excess
        (Loss_Kexc |dB/dt|^1.5) loss of the laminated materials (:116-150).
        The lamination data are MaterialData.extra entries."""
        dB = self._b_mag(self._col(dA_dt_col))
        vols = self.computeElementVolumes()
        tags = np.asarray(self.Assembler.Mesh.RegionTags)
        classic = excess = 0.0
        for tag in np.unique(tags):
            mat = dict(matLib).get(int(tag))
            ex = getattr(mat, "extra", {}) if mat is not None else {}
            if "Lam_d" not in ex or "Lam_sigma" not in ex:
                continue
            m = tags == tag
            classic += float((ex["Lam_sigma"] * ex["Lam_d"] ** 2 / 12.0 * dB[m] ** 2
                              * vols[m]).sum())
            kx = ex.get("Loss_Kexc")
            if kx is not None and kx > 0:
                excess += float((kx * dB[m] ** 1.5 * vols[m]).sum())
        return {"Classic": classic, "Excess": excess, "Total": classic + excess}


class ParameterExtractor:
    """ParameterExtractor.extractMatrices (src/post/ParameterExtractor.m:21-70):
    L = C^T K^-1 C over the windings' coupling vectors C, symmetrised, with K
    the Dirichlet-applied curl-curl matrix of nuMap, and R = diag(resistance).

    The reference hands K, which is singular on gradients, to MUMPS with the
    load C. The flux linkage C^T A is determined only by the part of C that
    drives no gradient field, C~ = C - M G Ln^-1 G^T C. The engine solves
    (K + eps M) A = C~ by auxiliary-space CG: K on gauged fields, with eps
    1e-12 of K's diagonal scale. The flux linkage is formed as C~^T A. In exact
    arithmetic that equals C^T A, since G^T M A = 0. The 1/eps conditioning
    leaves a rounding-level gradient part in A, which C~ (G^T C~ = 0) does not
    see and the raw C does (measured: 8e-7 of L with C, 3e-11 with C~)."""

    def __init__(self, assembler: Assembler):
        self.Assembler = assembler
        self.GaugeRegRel = 1e-12
        self.LinearRelTol = 1e-12
        self.LinearMaxIter = 50000

    def extractMatrices(self, space, nuMap, windings):
        import ctypes as C
        from . import _lib as L
        from .boundary import find_outer_boundary_dofs
        from .frequency import divergence_free_part
        from .transient import build_aux_space
        asm = self.Assembler
        wl = list(windings) if isinstance(windings, (list, tuple)) else [windings]
        nw = len(wl)
        Cm = asm.assembleWinding(space, wl)  # [N, nw]
        Rm = np.diag([float(w.Resistance) for w in wl])
        K = asm.assembleStiffness(space, nuMap, drop_zeros=False)
        M = asm.assembleMass(space, drop_zeros=False)
        fixed = find_outer_boundary_dofs(asm.Mesh, asm.DofHandler, space)
        n = K.shape[0]
        dev = K.values.device
        kd = K.to_scipy().diagonal()
        md = M.to_scipy().diagonal()
        free = ~fixed
        eps = self.GaugeRegRel * float(np.abs(kd[free]).mean() / np.abs(md[free]).mean())
        cvals = eps * M.values
        jsys = (K.values + cvals).contiguous()
        colptr = K.colptr.to(torch.int32).contiguous()
        rowind = K.rowind.contiguous()
        aux, keep = build_aux_space(asm, space.toString(), fixed, cvals, colptr, rowind)
        fixed_d = torch.from_numpy(fixed.astype(np.uint8)).to(dev)
        free_d = torch.from_numpy(free.astype(np.float64)).to(dev)
        A = torch.zeros((n, nw), dtype=torch.float64, device=dev)
        Ct = torch.zeros((n, nw), dtype=torch.float64, device=dev)
        self.Iterations = []
        for k in range(nw):
            ct = (divergence_free_part(asm, space, Cm[:, k].contiguous(), fixed) * free_d)
            ct = ct.to(torch.float64).contiguous()
            Ct[:, k] = ct
            x = torch.zeros(n, dtype=torch.float64, device=dev)
            it, rel = C.c_int(), C.c_double()
            L.check(L.lib.mpk_pcg_aux_solve(L.ptr(colptr), L.ptr(rowind), L.ptr(jsys), n,
                                            L.ptr(fixed_d), C.byref(aux), L.ptr(ct), L.ptr(x),
                                            L.ptr(keep["work"]), self.LinearRelTol,
                                            self.LinearMaxIter, 10, C.byref(it), C.byref(rel),
                                            L.stream_ptr()))
            L.check_krylov(it, rel, self.LinearRelTol,
                           f"ParameterExtractor.extractMatrices (winding {k + 1})")
            self.Iterations.append(int(it.value))
            A[:, k] = x
        Lm = (Ct.T @ A).cpu().numpy()
        return 0.5 * (Lm + Lm.T), Rm
