"""Host-side mirror of MagPackage's assembly layer on the MI355X engine.

Same class and method names as the reference (snake_case aliases), same
argument meaning and the same results (CSC matrices equal to what
``sparse()`` returns), computed by the HIP kernels of libmpk_assembly.so:

- ``Mesh``          src/Mesh.m  (P, T, RegionTags; ``generate_edges`` :114-138)
- ``FunctionSpace`` src/FunctionSpace.m
- ``FemConfig``     src/FemConfig.m:8-17
- ``DofHandler``    src/DofHandler.m:23-130 (Nedelec and Lagrange P1 spaces)
- ``Assembler``     src/Assembler.m:26-76, 182-268

Device arrays are ``torch`` tensors on ``cuda:0`` (torch is used for device
memory and streams only).  Matrices come back as :class:`CscMatrix` with
device-resident ``colptr`` / ``rowind`` / ``values``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from .material import pack_material_data


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


@dataclass
class FemConfig:
    AssemblyChunkSize: int = 2000
    DefaultQuadratureOrder: int = 2
    LinearTolerance: float = 1e-6
    DebugMode: bool = False

    @staticmethod
    def Default() -> "FemConfig":
        return FemConfig()


class FunctionSpace:
    def __init__(self, type_: str = "Nedelec", order: int = 1):
        self.Type = type_
        self.Order = order

    def toString(self) -> str:
        return f"{self.Type}_P{self.Order}"

    to_string = toString


class Mesh:
    """Tetrahedral mesh.  ``P``: (Nn, 3) float64, ``T``: (Ne, 4) 1-based ids."""

    def __init__(self, P, T, RegionTags=None):
        self.P = np.ascontiguousarray(P, dtype=np.float64)
        self.T = np.ascontiguousarray(T, dtype=np.int32)
        ne = self.T.shape[0]
        self.RegionTags = (np.ones(ne, dtype=np.int32) if RegionTags is None
                           else np.ascontiguousarray(RegionTags, dtype=np.int32))
        self.NumNodes = int(self.P.shape[0])
        self.NumElements = int(ne)
        dev = _dev()
        self.P_dev = torch.from_numpy(self.P.reshape(-1)).to(dev)
        self.T_dev = torch.from_numpy(self.T.reshape(-1)).to(dev)
        self.Tags_dev = torch.from_numpy(self.RegionTags).to(dev)
        self.Edges = None
        self.T2E = None
        self.T2E_Sign = None

    @staticmethod
    def from_kuhn(km) -> "Mesh":
        return Mesh(km.P, km.T, km.RegionTags)

    @staticmethod
    def load(filename: str, unit_tag: str = "m") -> "Mesh":
        """Mesh.load (src/Mesh.m:29-63): a COMSOL .mphtxt mesh, coordinates
        scaled to metres, or a Gmsh .msh mesh, read unscaled (a unit other
        than metres is warned about, as the reference does)."""
        from .mesh_io import read_mphtxt, read_msh
        if str(filename).lower().endswith(".mphtxt"):
            raw = read_mphtxt(str(filename), unit_tag)
        elif str(filename).lower().endswith(".msh"):
            raw = read_msh(str(filename))
            if str(unit_tag).lower() != "m":
                import warnings
                warnings.warn("read_msh does not scale coordinates; check the unit")
        else:
            raise ValueError(f"unsupported mesh format: {filename}")
        return Mesh(raw.P, raw.T, raw.RegionTags)

    def generateEdges(self) -> None:
        """Mesh.generateEdges (src/Mesh.m:114-138) on the GPU."""
        if self.Edges is not None:
            return
        ne, nn = self.NumElements, self.NumNodes
        dev = _dev()
        edges = torch.empty(max(12 * ne, 2), dtype=torch.int32, device=dev)
        t2e = torch.empty(max(6 * ne, 1), dtype=torch.int32, device=dev)
        sign = torch.empty(max(6 * ne, 1), dtype=torch.int8, device=dev)
        nedge = C.c_int64(0)
        L.check(L.lib.mpk_generate_edges(L.ptr(self.T_dev), ne, nn, L.ptr(edges), L.ptr(t2e),
                                         L.ptr(sign), C.byref(nedge), L.stream_ptr()))
        E = int(nedge.value)
        self.Edges = edges[: 2 * E].view(E, 2)
        self.T2E = t2e[: 6 * ne].view(ne, 6)
        self.T2E_Sign = sign[: 6 * ne].view(ne, 6)
        self.NumEdges = E

    generate_edges = generateEdges


class DofHandler:
    def __init__(self, mesh: Mesh):
        self.Mesh = mesh
        self.NumGlobalDofs = 0
        self.DofMaps = {}
        self.SpaceOffsets = {}
        self.SpaceLocalSizes = {}
        self.EntityMaps = {}

    def distributeDofs(self, space: FunctionSpace, targetRegionTags=None):
        """DofHandler.distributeDofs (src/DofHandler.m:23-122): Nedelec P1
        (:72-105) and Lagrange P1 (:52-70) spaces."""
        key = space.toString()
        if key in self.DofMaps:
            return self.DofMaps[key], self.SpaceOffsets[key]
        if space.Type.lower() == "lagrange" and space.Order == 1:
            return self._distribute_nodal(key, targetRegionTags)
        if space.Type.lower() != "nedelec" or space.Order != 1:
            raise ValueError(f"Unsupported space: {key} (the engine covers Nedelec P1)")
        mesh = self.Mesh
        mesh.generateEdges()
        ne, E = mesh.NumElements, mesh.NumEdges
        dev = _dev()
        active = None
        if targetRegionTags is not None and len(np.atleast_1d(targetRegionTags)) > 0:
            act = np.isin(mesh.RegionTags, np.atleast_1d(targetRegionTags)).astype(np.uint8)
            active = torch.from_numpy(act).to(dev)
        emap = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
        cell = torch.empty(max(6 * ne, 1), dtype=torch.int32, device=dev)
        nd = C.c_int64(0)
        offset = self.NumGlobalDofs
        L.check(L.lib.mpk_distribute_dofs(L.ptr(mesh.T2E), ne, E, L.ptr(active), offset,
                                          L.ptr(emap), L.ptr(cell), C.byref(nd), L.stream_ptr()))
        cell_dofs = cell[: 6 * ne].view(ne, 6)
        self.DofMaps[key] = cell_dofs
        self.SpaceOffsets[key] = offset
        self.SpaceLocalSizes[key] = int(nd.value)
        self.EntityMaps[key] = emap[:E]
        self.NumGlobalDofs += int(nd.value)
        return cell_dofs, offset

    distribute_dofs = distributeDofs

    def _distribute_nodal(self, key, targetRegionTags):
        """Lagrange P1: the nodes of the active elements, in increasing node id,
        numbered offset + 1, offset + 2, ...; CellDofs (Ne, 4) int32, 0 at
        inactive nodes (src/DofHandler.m:52-70)."""
        mesh = self.Mesh
        ne, nn = mesh.NumElements, mesh.NumNodes
        dev = _dev()
        T0 = mesh.T_dev.view(ne, 4).to(torch.int64) - 1
        if targetRegionTags is not None and len(np.atleast_1d(targetRegionTags)) > 0:
            act = torch.from_numpy(np.isin(mesh.RegionTags, np.atleast_1d(targetRegionTags)))
            T0a = T0[act.to(dev)]
        else:
            T0a = T0
        nodes = torch.unique(T0a)  # sorted, as MATLAB's unique
        offset = self.NumGlobalDofs
        emap = torch.zeros(max(nn, 1), dtype=torch.int32, device=dev)
        emap[nodes] = (offset + 1 + torch.arange(nodes.numel(), device=dev)).to(torch.int32)
        cell = emap[T0].contiguous()
        self.DofMaps[key] = cell
        self.SpaceOffsets[key] = offset
        self.SpaceLocalSizes[key] = int(nodes.numel())
        self.EntityMaps[key] = emap[:nn]
        self.NumGlobalDofs += int(nodes.numel())
        return cell, offset

    def packForKernel(self, space: FunctionSpace) -> dict:
        key = space.toString()
        return {"CellDofs": self.DofMaps[key], "Signs": self.Mesh.T2E_Sign}


class CscMatrix:
    """Device CSC matrix (0-based row indices), MATLAB's sparse storage."""

    def __init__(self, shape, colptr, rowind, values):
        self.shape = tuple(int(s) for s in shape)
        self.colptr = colptr
        self.rowind = rowind
        self.values = values

    @property
    def nnz(self) -> int:
        return int(self.rowind.numel())

    def to_scipy(self):
        import scipy.sparse as sp
        v = self.values.detach().cpu().numpy()
        return sp.csc_matrix((v, self.rowind.cpu().numpy().astype(np.int64),
                              self.colptr.cpu().numpy().astype(np.int64)), shape=self.shape)


class Plan:
    """Owner of an mpk_plan (pattern + incidence lists for one DOF map)."""

    def __init__(self, mesh: Mesh, cell_dofs: torch.Tensor, ndofs: int):
        self._p = C.c_void_p(None)
        L.check(L.lib.mpk_plan_create(L.ptr(mesh.T_dev), L.ptr(cell_dofs.reshape(-1)),
                                      mesh.NumElements, mesh.NumNodes, ndofs, L.stream_ptr(),
                                      C.byref(self._p)))
        nnz, nc, ne, ninc, mt = (C.c_int64() for _ in range(5))
        L.check(L.lib.mpk_plan_info(self._p, C.byref(nnz), C.byref(nc), C.byref(ne),
                                    C.byref(ninc), C.byref(mt)))
        self.nnz, self.n, self.ne = nnz.value, nc.value, ne.value
        self.n_incidences, self.max_tile_nnz = ninc.value, mt.value
        ring = C.c_int()
        L.check(L.lib.mpk_plan_ring_order(self._p, C.byref(ring)))
        self.ring_order = bool(ring.value)
        # ring record format: "compact" (8 B per incidence), "16" or None
        self.ring_format = {0: None, 1: "16", 2: "compact"}[int(ring.value)]
        cp, ri = C.c_void_p(), C.c_void_p()
        L.check(L.lib.mpk_plan_pattern(self._p, C.byref(cp), C.byref(ri)))
        self._colptr_addr, self._rowind_addr = cp.value, ri.value

    @property
    def handle(self):
        return self._p

    def pattern(self):
        """(colptr int64 [n+1], rowind int32 [nnz]) copies on the device."""
        dev = _dev()
        colptr = torch.empty(self.n + 1, dtype=torch.int32, device=dev)
        rowind = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        # copy the plan-owned pattern (device to device, stream-ordered)
        _memcpy_d2d(colptr.data_ptr(), self._colptr_addr, (self.n + 1) * 4)
        if self.nnz:
            _memcpy_d2d(rowind.data_ptr(), self._rowind_addr, self.nnz * 4)
        return colptr.to(torch.int64), rowind[: self.nnz]

    def drop_zeros(self, values: torch.Tensor, complex_vals: bool) -> CscMatrix:
        dev = _dev()
        colptr = torch.empty(self.n + 1, dtype=torch.int64, device=dev)
        rowind = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        vout = torch.empty_like(values)
        nnz = C.c_int64(0)
        L.check(L.lib.mpk_csc_drop_zeros(self._p, L.ptr(values), int(complex_vals), L.ptr(colptr),
                                         L.ptr(rowind), L.ptr(vout), C.byref(nnz),
                                         L.stream_ptr()))
        k = int(nnz.value)
        return CscMatrix((self.n, self.n), colptr, rowind[:k],
                         vout[: 2 * k] if complex_vals else vout[:k])

    def __del__(self):
        p = getattr(self, "_p", None)
        lib = getattr(L, "lib", None)  # None while the interpreter shuts down
        if p is not None and p.value and lib is not None:
            lib.mpk_plan_destroy(p)
            self._p = C.c_void_p(None)


def _memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    # torch has no raw-pointer copy: wrap both addresses as byte tensors.
    src_t = _wrap_device_bytes(src, nbytes)
    dst_t = _wrap_device_bytes(dst, nbytes)
    dst_t.copy_(src_t)


def _wrap_device_bytes(addr: int, nbytes: int) -> torch.Tensor:
    class _CAI:
        __cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1",
                                    "data": (addr, False), "version": 3, "strides": None}
    return torch.as_tensor(_CAI(), device=_dev())


def expand_material_map(map_or_array, tags: np.ndarray) -> np.ndarray:
    """Assembler.expandMaterialMap (src/Assembler.m:240-268)."""
    ne = len(tags)
    if isinstance(map_or_array, dict):
        vec = np.zeros(ne)
        for k, v in map_or_array.items():
            kv = float(k) if not isinstance(k, str) else float(k)
            if np.isscalar(v):
                vec[tags == kv] = v
        return vec
    a = np.asarray(map_or_array, dtype=np.float64)
    if a.ndim == 0 or a.size == 1:
        return np.full(ne, float(a.ravel()[0]))
    if a.ndim == 1 or min(a.shape) == 1:
        a = a.ravel()
        if a.size == ne:
            return a.copy()
        return a[tags - 1]
    raise ValueError("Assembler:InvalidInput Invalid material array format.")


def expand_source_map(map_or_array, tags: np.ndarray) -> np.ndarray:
    """Assembler.expandSourceMap (src/Assembler.m:308-322): (Ne, 3) current
    density, the memory of the reference's [3 x Ne] array.  Accepts a
    {tag: (jx, jy, jz)} dict, a 3-vector, or a (3, Ne) / (Ne, 3) array."""
    ne = len(tags)
    out = np.zeros((ne, 3))
    if isinstance(map_or_array, dict):
        for k, v in map_or_array.items():
            mask = tags == float(k)
            v = np.asarray(v, dtype=np.float64).ravel()
            if mask.any():
                out[mask] = v[:3]
        return out
    a = np.asarray(map_or_array, dtype=np.float64)
    if a.size == 3:
        out[:] = a.ravel()
        return out
    if a.shape == (3, ne):
        return np.ascontiguousarray(a.T)
    if a.shape == (ne, 3):
        return np.ascontiguousarray(a)
    raise ValueError("Source must be [3x1] or [3xN].")


class Assembler:
    def __init__(self, mesh: Mesh, dofHandler: DofHandler, config: FemConfig | None = None):
        self.Mesh = mesh
        self.DofHandler = dofHandler
        self.Config = config or FemConfig.Default()
        self._plans = {}

    # -- helpers ---------------------------------------------------------
    def plan(self, space: FunctionSpace) -> Plan:
        key = (space.toString(), self.DofHandler.NumGlobalDofs)
        if key not in self._plans:
            cell = self.DofHandler.DofMaps[space.toString()]
            self._plans[key] = Plan(self.Mesh, cell, self.DofHandler.NumGlobalDofs)
        return self._plans[key]

    def _elem_vec(self, map_or_array) -> torch.Tensor:
        v = expand_material_map(map_or_array, self.Mesh.RegionTags)
        return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(_dev())

    def stiffness_quadrature(self) -> L.Quadrature:
        return L.standard_quadrature(max(self.Config.DefaultQuadratureOrder, 2))

    def mass_quadrature(self, weighted: bool) -> L.Quadrature:
        o = self.Config.DefaultQuadratureOrder
        return L.standard_quadrature(o if weighted else max(o, 2))

    def _finish(self, plan: Plan, vals: torch.Tensor, complex_vals: bool, drop_zeros: bool):
        if drop_zeros:
            m = plan.drop_zeros(vals, complex_vals)
        else:
            colptr, rowind = plan.pattern()
            m = CscMatrix((plan.n, plan.n), colptr, rowind, vals)
        if complex_vals:
            m.values = torch.view_as_complex(m.values.view(-1, 2))
        return m

    # -- assembly (src/Assembler.m) ---------------------------------------
    def assembleStiffness(self, space, materialMap, precision: str = "fp64",
                          drop_zeros: bool = True, out: torch.Tensor | None = None):
        plan = self.plan(space)
        nu = self._elem_vec(materialMap)
        vals = out if out is not None else torch.empty(max(plan.nnz, 1), dtype=torch.float64,
                                                       device=_dev())
        q = self.stiffness_quadrature()
        L.check(L.lib.mpk_csc_curl_curl(plan.handle, L.ptr(self.Mesh.P_dev), L.ptr(nu),
                                        C.byref(q), L.PRECISIONS[precision], L.ptr(vals),
                                        L.stream_ptr()))
        return self._finish(plan, vals[: plan.nnz], False, drop_zeros)

    def assembleMass(self, space, precision: str = "fp64", drop_zeros: bool = True):
        plan = self.plan(space)
        vals = torch.empty(max(plan.nnz, 1), dtype=torch.float64, device=_dev())
        q = self.mass_quadrature(weighted=False)
        L.check(L.lib.mpk_csc_mass(plan.handle, L.ptr(self.Mesh.P_dev), None, C.byref(q),
                                   L.PRECISIONS[precision], L.ptr(vals), L.stream_ptr()))
        return self._finish(plan, vals[: plan.nnz], False, drop_zeros)

    def assembleMassWeighted(self, space, coeffMap, precision: str = "fp64",
                             drop_zeros: bool = True):
        plan = self.plan(space)
        coeff = self._elem_vec(coeffMap)
        vals = torch.empty(max(plan.nnz, 1), dtype=torch.float64, device=_dev())
        q = self.mass_quadrature(weighted=True)
        L.check(L.lib.mpk_csc_mass(plan.handle, L.ptr(self.Mesh.P_dev), L.ptr(coeff), C.byref(q),
                                   L.PRECISIONS[precision], L.ptr(vals), L.stream_ptr()))
        return self._finish(plan, vals[: plan.nnz], False, drop_zeros)

    def assembleFrequencyOperator(self, space, nuMap, sigmaMap, omega: float,
                                  precision: str = "fp64", drop_zeros: bool = True):
        """K + 1j*omega*M_sigma (FrequencySolver.m:127-131 without the eps_reg term)."""
        plan = self.plan(space)
        nu = self._elem_vec(nuMap)
        sg = self._elem_vec(sigmaMap)
        vals = torch.empty(max(2 * plan.nnz, 2), dtype=torch.float64, device=_dev())
        qk = self.stiffness_quadrature()
        qm = self.mass_quadrature(weighted=True)
        L.check(L.lib.mpk_csc_curl_curl_mass_complex(
            plan.handle, L.ptr(self.Mesh.P_dev), L.ptr(nu), L.ptr(sg), float(omega), C.byref(qk),
            C.byref(qm), L.PRECISIONS[precision], L.ptr(vals), L.stream_ptr()))
        return self._finish(plan, vals[: 2 * plan.nnz], True, drop_zeros)

    def material_tables(self, matLibData):
        """Device copy of the packed B-H tables + the ctypes struct pointing at them."""
        pk = pack_material_data(matLibData)
        dev = _dev()
        keep = {k: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
                for k, v in pk.items()}
        m = L.Material()
        m.linear_nu = L.ptr(keep["linNu"]); m.is_nonlinear = L.ptr(keep["isNon"])
        m.max_bsq = L.ptr(keep["maxB"]); m.breaks = L.ptr(keep["breaks"])
        m.coefs = L.ptr(keep["coefs"]); m.spline_start = L.ptr(keep["starts"])
        m.spline_count = L.ptr(keep["counts"]); m.coef_start = L.ptr(keep["coe_starts"])
        m.num_tags = keep["linNu"].numel()
        m.num_breaks = keep["breaks"].numel()
        m.num_coefs = keep["coefs"].numel()
        return m, keep

    def assembleJacobian(self, space, solutionA, matLibData, calcJ: bool = True,
                         precision: str = "fp64", drop_zeros: bool = True, tables=None):
        """[J, R] = assembleJacobian (src/Assembler.m:182-231)."""
        plan = self.plan(space)
        dev = _dev()
        if not torch.is_tensor(solutionA):
            solutionA = torch.from_numpy(np.ascontiguousarray(solutionA, dtype=np.float64))
        sol = solutionA.to(dev, dtype=torch.float64).contiguous()
        if sol.numel() < plan.n:
            sol = torch.cat([sol, torch.zeros(plan.n - sol.numel(), dtype=torch.float64, device=dev)])
        mat, keep = tables if tables is not None else self.material_tables(matLibData)
        tags = self.Mesh.Tags_dev
        ntag = int(mat.num_tags)
        tmin, tmax = int(self.Mesh.RegionTags.min()), int(self.Mesh.RegionTags.max())
        if tmin < 1 or tmax > ntag:
            raise L.MpkError(L.MPK_ERR_INDEX, f"region tags [{tmin}, {tmax}] outside material table")
        jv = torch.empty(max(plan.nnz, 1), dtype=torch.float64, device=dev) if calcJ else None
        res = torch.empty(max(plan.n, 1), dtype=torch.float64, device=dev)
        q = self.stiffness_quadrature()
        L.check(L.lib.mpk_csc_jacobian(plan.handle, L.ptr(self.Mesh.P_dev), L.ptr(tags),
                                       L.ptr(sol), C.byref(mat), C.byref(q), int(bool(calcJ)),
                                       L.PRECISIONS[precision], L.ptr(jv), L.ptr(res),
                                       L.stream_ptr()))
        del keep
        J = self._finish(plan, jv[: plan.nnz], False, drop_zeros) if calcJ else None
        return J, res[: plan.n]

    def assembleNewtonSet(self, space, solutionA, matLibData, sigmaMap=None,
                          precision: str = "fp64", outputs=("K", "M", "J", "R"), tables=None):
        """One pass over the mesh for the Newton quantities at solution A:
        secant stiffness K(nu(B_e)) (what assembleStiffness builds from the
        per-element nu the B-H curve gives at A), the sigma-mass (sigma = 1
        when sigmaMap is None), the Jacobian J and residual R of
        assembleJacobian (src/Assembler.m:182-231).  Returns a dict of raw
        structural-CSC value tensors (plan pattern, no zero dropping) for the
        requested outputs; FP64 and FP32_LOCAL only."""
        plan = self.plan(space)
        dev = _dev()
        if not torch.is_tensor(solutionA):
            solutionA = torch.from_numpy(np.ascontiguousarray(solutionA, dtype=np.float64))
        sol = solutionA.to(dev, dtype=torch.float64).contiguous()
        if sol.numel() < plan.n:
            sol = torch.cat([sol, torch.zeros(plan.n - sol.numel(), dtype=torch.float64, device=dev)])
        mat, keep = tables if tables is not None else self.material_tables(matLibData)
        ntag = int(mat.num_tags)
        tmin, tmax = int(self.Mesh.RegionTags.min()), int(self.Mesh.RegionTags.max())
        if tmin < 1 or tmax > ntag:
            raise L.MpkError(L.MPK_ERR_INDEX, f"region tags [{tmin}, {tmax}] outside material table")
        sig = None if sigmaMap is None else self._elem_vec(sigmaMap)
        out = {}
        for name in ("K", "M", "J"):
            if name in outputs:
                out[name] = torch.empty(max(plan.nnz, 1), dtype=torch.float64, device=dev)
        if "R" in outputs:
            out["R"] = torch.empty(max(plan.n, 1), dtype=torch.float64, device=dev)
        qk = self.stiffness_quadrature()
        qm = self.mass_quadrature(weighted=sigmaMap is not None)
        L.check(L.lib.mpk_csc_newton_set(
            plan.handle, L.ptr(self.Mesh.P_dev), L.ptr(self.Mesh.Tags_dev), L.ptr(sol),
            C.byref(mat), L.ptr(sig), C.byref(qk), C.byref(qm), L.PRECISIONS[precision],
            L.ptr(out.get("K")), L.ptr(out.get("M")), L.ptr(out.get("J")), L.ptr(out.get("R")),
            L.stream_ptr()))
        del keep
        return {k: (v[: plan.n] if k == "R" else v[: plan.nnz]) for k, v in out.items()}

    def assembleHBFEM(self, space, SolHarmonics, aft, matLibData, requestJacobian: bool = True):
        """The harmonic-balance assembly HBFEMSolver calls (src/solver/HBFEMSolver.m:78,96;
        the method is absent from the reference's Assembler, whose MATLAB kernel
        src/kernels/assemble_hbfem_kernel.m:1-240 defines it) on the engine's HBFEM
        kernel.  SolHarmonics [NumGlobalDofs x H] complex (host or device);
        aft: hbfem.AFT.  Returns (J_tri, R): J_tri a dict of device tensors I, J, V
        (1-based, the triplets with both indices > 0; None without
        requestJacobian) of the block-diagonal DC-reluctivity Jacobian, R the
        [NumGlobalDofs x H] complex128 residual (duplicates summed in triplet
        order)."""
        dev = _dev()
        key = space.toString()
        dofs = self.DofHandler.DofMaps[key]
        nd = int(self.DofHandler.NumGlobalDofs)
        sol = torch.as_tensor(SolHarmonics).to(dev)
        H = int(aft.NumHarmonics)
        if sol.dim() == 1:
            sol = sol[:, None]
        if sol.shape[0] != nd or sol.shape[1] != H:
            raise L.MpkError(L.MPK_ERR_ARGS, f"SolHarmonics must be [{nd} x {H}]")
        solc = sol.to(torch.complex128)
        sr = solc.real.t().contiguous()  # column-major [nd x H]
        si = solc.imag.t().contiguous()
        M = self.Mesh
        self.Mesh.generateEdges()
        ne = M.NumElements
        Td = M.T_dev.to(torch.float64)
        Dd = dofs.reshape(-1).to(torch.float64)
        Sd = M.T2E_Sign.reshape(-1).to(torch.float64)
        tags = M.Tags_dev.to(torch.float64)
        cm = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(dev).reshape(-1)
        dr, di = cm(aft.D_matrix.real), cm(aft.D_matrix.imag)
        pr, pi_ = cm(aft.P_matrix.real), cm(aft.P_matrix.imag)
        sc = torch.from_numpy(aft.jacobian_scalings()).to(dev)
        mat, keep = self.material_tables(matLibData)
        q = self.stiffness_quadrature()
        nk = 36 * H * ne if requestJacobian else 0
        out = lambda n: torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        Io, Jo, Vo = out(nk), out(nk), out(nk)
        Rr, Rc, Rvr, Rvi = (out(6 * H * ne) for _ in range(4))
        L.check(L.lib.mpk_hbfem_coo(
            L.ptr(M.P_dev), M.NumNodes, L.ptr(Td), ne, L.ptr(Dd), L.ptr(Sd), L.ptr(tags),
            L.ptr(sr), L.ptr(si), nd, H, L.ptr(dr), L.ptr(di), int(aft.NumTimeSteps), L.ptr(pr),
            L.ptr(pi_), C.byref(q), C.byref(mat), int(bool(requestJacobian)), L.ptr(sc),
            L.ptr(Io), L.ptr(Jo), L.ptr(Vo), L.ptr(Rr), L.ptr(Rc), L.ptr(Rvr), L.ptr(Rvi),
            L.stream_ptr()))
        del keep
        n6 = 6 * H * ne
        R = torch.zeros((nd, H), dtype=torch.complex128, device=dev)
        for part, vals in ((0, Rvr[:n6]), (1, Rvi[:n6])):
            colptr = torch.empty(H + 1, dtype=torch.int64, device=dev)
            rowind = torch.empty(max(n6, 1), dtype=torch.int32, device=dev)
            v = torch.empty(max(n6, 1), dtype=torch.float64, device=dev)
            nnz = C.c_int64(0)
            L.check(L.lib.mpk_sparse_coo_to_csc(L.ptr(Rr), L.ptr(Rc), L.ptr(vals), n6, nd, H, 0,
                                                L.ptr(colptr), L.ptr(rowind), L.ptr(v), n6,
                                                C.byref(nnz), L.stream_ptr()))
            k = int(nnz.value)
            cols = torch.repeat_interleave(torch.arange(H, device=dev), colptr.diff())
            dense = torch.zeros((nd, H), dtype=torch.float64, device=dev)
            dense[rowind[:k].long(), cols[:k]] = v[:k]
            if part == 0:
                R.real.copy_(dense)
            else:
                R.imag.copy_(dense)
        J_tri = None
        if requestJacobian:
            ok = (Io[:nk] > 0) & (Jo[:nk] > 0)
            J_tri = {"I": Io[:nk][ok], "J": Jo[:nk][ok], "V": Vo[:nk][ok]}
        return J_tri, R

    def source_quadrature(self) -> L.Quadrature:
        return L.standard_quadrature(max(self.Config.DefaultQuadratureOrder, 2))

    def assembleScalarLaplacian(self, space, coeffMap) -> CscMatrix:
        """K = assembleScalarLaplacian(space, coeffMap) (src/Assembler.m:121-132):
        assemble_scalar_laplacian_kernel_mex at the 1-point rule over the
        space's node DOFs, then sparse(I(valid), J(valid), V(valid), N, N),
        N = NumGlobalDofs, as a device CSC matrix."""
        from .fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
        if space.Type.lower() != "lagrange":
            raise ValueError("assembleScalarLaplacian needs a Lagrange space")
        mesh = self.Mesh
        cell = self.DofHandler.DofMaps[space.toString()]
        n = self.DofHandler.NumGlobalDofs
        ne, nn = mesh.NumElements, mesh.NumNodes
        dev = _dev()
        q_pts, q_w = get_quadrature_data("tet", 1)
        grad = np.ascontiguousarray(matlab_flat(lagrange_tet_p1(q_pts)[1]))
        q = L.make_quadrature(q_w)
        coeff = torch.from_numpy(expand_material_map(coeffMap, mesh.RegionTags)).to(dev)
        Td = mesh.T_dev.to(torch.float64)
        Dd = cell.reshape(-1).to(torch.float64)
        nt = 16 * ne
        Io = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
        Jo, Vo = torch.empty_like(Io), torch.empty_like(Io)
        L.check(L.lib.mpk_scalar_laplacian_coo(
            L.ptr(mesh.P_dev), nn, L.ptr(Td), ne, L.ptr(Dd), L.ptr(coeff), C.byref(q),
            grad.ctypes.data_as(C.c_void_p), L.ptr(Io), L.ptr(Jo), L.ptr(Vo), L.stream_ptr()))
        colptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
        rowind = torch.empty(max(nt, 1), dtype=torch.int32, device=dev)
        vals = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
        nnz = C.c_int64(0)
        L.check(L.lib.mpk_sparse_coo_to_csc(L.ptr(Io), L.ptr(Jo), L.ptr(Vo), nt, n, n, 1,
                                            L.ptr(colptr), L.ptr(rowind), L.ptr(vals), nt,
                                            C.byref(nnz), L.stream_ptr()))
        k = int(nnz.value)
        return CscMatrix((n, n), colptr, rowind[:k], vals[:k])

    def assembleCoupling(self, spaceRow, spaceCol, coeffMap) -> CscMatrix:
        """C = assembleCoupling(spaceRow, spaceCol, coeffMap) (src/Assembler.m:99-119):
        assemble_coupling_kernel_mex (Nedelec rows, Lagrange columns) at the
        config's quadrature order (at least 2), then sparse(I(valid), J(valid),
        V(valid), N, N) as a device CSC matrix."""
        from .fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
        if spaceRow.Type.lower() != "nedelec" or spaceCol.Type.lower() != "lagrange":
            raise ValueError("Unsupported space combination")
        mesh = self.Mesh
        dh = self.DofHandler
        n = dh.NumGlobalDofs
        ne, nn = mesh.NumElements, mesh.NumNodes
        dev = _dev()
        order = max(self.Config.DefaultQuadratureOrder, 2)
        q_pts, _ = get_quadrature_data("tet", order)
        grad = np.ascontiguousarray(matlab_flat(lagrange_tet_p1(q_pts)[1]))
        q = L.standard_quadrature(order)
        coeff = torch.from_numpy(expand_material_map(coeffMap, mesh.RegionTags)).to(dev)
        Td = mesh.T_dev.to(torch.float64)
        De = dh.DofMaps[spaceRow.toString()].reshape(-1).to(torch.float64)
        Sd = mesh.T2E_Sign.reshape(-1).to(torch.float64)
        Dn = dh.DofMaps[spaceCol.toString()].reshape(-1).to(torch.float64)
        nt = 24 * ne
        Io = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
        Jo, Vo = torch.empty_like(Io), torch.empty_like(Io)
        L.check(L.lib.mpk_coupling_coo(
            L.ptr(mesh.P_dev), nn, L.ptr(Td), ne, L.ptr(De), L.ptr(Sd), L.ptr(Dn), L.ptr(coeff),
            C.byref(q), grad.ctypes.data_as(C.c_void_p), L.ptr(Io), L.ptr(Jo), L.ptr(Vo),
            L.stream_ptr()))
        colptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
        rowind = torch.empty(max(nt, 1), dtype=torch.int32, device=dev)
        vals = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
        nnz = C.c_int64(0)
        L.check(L.lib.mpk_sparse_coo_to_csc(L.ptr(Io), L.ptr(Jo), L.ptr(Vo), nt, n, n, 1,
                                            L.ptr(colptr), L.ptr(rowind), L.ptr(vals), nt,
                                            C.byref(nnz), L.stream_ptr()))
        k = int(nnz.value)
        return CscMatrix((n, n), colptr, rowind[:k], vals[:k])

    def assembleSource(self, space, sourceMap, out: torch.Tensor | None = None):
        """F = assembleSource(space, sourceMap) (src/Assembler.m:78-97): the
        dense [NumGlobalDofs] source vector on the device."""
        plan = self.plan(space)
        js = sourceMap if torch.is_tensor(sourceMap) else torch.from_numpy(
            expand_source_map(sourceMap, self.Mesh.RegionTags)).to(_dev())
        js = js.to(dtype=torch.float64).contiguous()
        F = out if out is not None else torch.empty(max(plan.n, 1), dtype=torch.float64,
                                                    device=_dev())
        q = self.source_quadrature()
        L.check(L.lib.mpk_csc_source(plan.handle, L.ptr(self.Mesh.P_dev), L.ptr(js), C.byref(q),
                                     L.ptr(F), L.stream_ptr()))
        return F[: plan.n]

    def assembleWinding(self, space, windingObj) -> torch.Tensor:
        """C = assembleWinding(space, windingObj) (src/Assembler.m:134-180):
        assemble_winding_kernel_mex at the config's quadrature order with
        scale Turns / CrossSectionArea, then sparse(I(valid), 1, V(valid), N, 1),
        as a dense [N] device vector; a list of windings gives [N, k]."""
        if isinstance(windingObj, (list, tuple)):
            return torch.stack([self.assembleWinding(space, w) for w in windingObj], dim=1)
        mesh = self.Mesh
        cell = self.DofHandler.DofMaps[space.toString()]
        n = self.DofHandler.NumGlobalDofs
        ne, nn = mesh.NumElements, mesh.NumNodes
        dev = _dev()
        # MATLAB [k x Ne] column-major arrays are this package's (Ne, k) row-major ones
        Td = mesh.T_dev.to(torch.float64)
        Dd = cell.reshape(-1).to(torch.float64)
        Sd = mesh.T2E_Sign.reshape(-1).to(torch.float64)
        tags = mesh.Tags_dev.to(torch.float64)
        dirf = None
        if windingObj.DirectionField is not None and np.size(windingObj.DirectionField):
            df = np.asarray(windingObj.DirectionField, dtype=np.float64)
            if df.shape != (3, ne):
                raise ValueError("DirectionField must be [3 x NumElements]")
            dirf = torch.from_numpy(np.ascontiguousarray(df.T)).to(dev)
        uni = np.zeros(3)
        d = np.asarray(windingObj.Direction, dtype=np.float64).ravel()
        uni[: min(3, d.size)] = d[:3]
        scale = float(windingObj.Turns) / float(windingObj.CrossSectionArea)
        q = L.standard_quadrature(self.Config.DefaultQuadratureOrder)
        Io = torch.empty(max(6 * ne, 1), dtype=torch.float64, device=dev)
        Vo = torch.empty_like(Io)
        L.check(L.lib.mpk_winding_coo(
            L.ptr(mesh.P_dev), nn, L.ptr(Td), ne, L.ptr(Dd), L.ptr(Sd), L.ptr(tags),
            float(windingObj.RegionID), L.ptr(dirf), uni.ctypes.data_as(C.c_void_p), scale,
            C.byref(q), L.ptr(Io), L.ptr(Vo), L.stream_ptr()))
        nt = 6 * ne
        Jo = torch.ones_like(Io)
        colptr = torch.empty(2, dtype=torch.int64, device=dev)
        rowind = torch.empty(max(nt, 1), dtype=torch.int32, device=dev)
        vals = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
        nnz = C.c_int64(0)
        L.check(L.lib.mpk_sparse_coo_to_csc(L.ptr(Io), L.ptr(Jo), L.ptr(Vo), nt, n, 1, 1,
                                            L.ptr(colptr), L.ptr(rowind), L.ptr(vals), nt,
                                            C.byref(nnz), L.stream_ptr()))
        k = int(nnz.value)
        out = torch.zeros(max(n, 1), dtype=torch.float64, device=dev)
        out[rowind[:k].to(torch.int64)] = vals[:k]
        return out[:n]

    def computeElementB(self, solutionA, space=None):
        """PostProcessor.computeElementB (src/post/PostProcessor.m:64-126) for one
        solution column: (Ne, 3) device tensor of B at the elements."""
        key = space.toString() if space is not None else next(iter(self.DofHandler.DofMaps))
        cell = self.DofHandler.DofMaps[key]
        sol = solutionA if torch.is_tensor(solutionA) else torch.from_numpy(
            np.ascontiguousarray(solutionA, dtype=np.float64))
        sol = sol.to(_dev(), dtype=torch.float64).contiguous()
        ne = self.Mesh.NumElements
        B = torch.empty((max(ne, 1), 3), dtype=torch.float64, device=_dev())
        q = self.stiffness_quadrature()
        L.check(L.lib.mpk_element_b(L.ptr(self.Mesh.P_dev), self.Mesh.NumNodes,
                                    L.ptr(self.Mesh.T_dev), L.ptr(cell.reshape(-1)), ne,
                                    L.ptr(sol), sol.numel(), C.byref(q), L.ptr(B),
                                    L.stream_ptr()))
        return B[:ne]

    # snake_case aliases
    assemble_stiffness = assembleStiffness
    assemble_mass = assembleMass
    assemble_mass_weighted = assembleMassWeighted
    assemble_jacobian = assembleJacobian
    assemble_frequency_operator = assembleFrequencyOperator
    assemble_source = assembleSource
    assemble_winding = assembleWinding
    compute_element_b = computeElementB
