"""Field-circuit coupled transients on the GPU engine.

The reference's BDF1 and BDF2 solvers (src/solver/TransientBDF1Solver.m:1-358,
src/solver/TransientBDF2Solver.m:1-405) couple three things:
- the nonlinear A field, built with assembleJacobian (the MEX Jacobian kernel);
- a Coulomb gauge, a Lagrange multiplier P on the nodes with
  C_AP = int N_i . grad(phi_j);
- one circuit equation per winding:
  Scale * (R I + L dI/dt + W^T dA/dt - V(t)) = 0.

MUMPS factors the resulting monolithic, unsymmetric saddle-point Jacobian.
This module keeps the reference's time loop, Newton loop, residual norm, line
search and GeoBalance scale. Each Newton system is solved with the engine's
SPD solver instead, by the following exact reformulation.

Why the multiplier can be eliminated (all conductivities zero, as in the
reference's circuit tutorials):
- Let G be the edge-node incidence over free nodes, M the unit edge mass and
  Ln = G^T M G. Then C_AP = M G, and the curl operators vanish on gradients:
  G^T J = 0 for the Jacobian J as for K.
- The P rows hold G^T M A = 0 at every Newton iterate. The A rows then give
  Ln P = G^T W I. Eliminating P leaves J dA = (W - M G Ln^-1 G^T W) I + ...,
  that is W replaced by its M-orthogonal projection W~ onto discretely
  divergence-free loads.
- With W~ the reference's A-row residual R_fem - W I + M G P equals
  R_fem - W~ I. Its P rows vanish, and W^T A = W~^T A, so the residual norm,
  the Newton steps, the line search and the converged states are the
  reference's.
- The Newton matrix Jr = J + eps M (eps = GaugeRegRel x the stiffness-to-mass
  scale) is SPD. On right-hand sides with G^T r = 0 it keeps G^T M dA = 0
  exactly. eps enters only the Newton matrix, never the residual, so the
  converged states are unbiased; the step is inexact by O(eps) relative. Jr is
  solved by the auxiliary-space PCG.
- The circuit unknowns are eliminated by a Schur complement. Per Newton
  iteration, u = Jr^-1 R_A and v_k = Jr^-1 W~_k, and then
  (alpha W~^T V + diag(R + alpha L)) dI = -R_c / Scale + alpha W~^T u,
  dA = -u + V dI: the Newton step of the monolithic system. The row scale
  cancels; it enters the residual norm only, as in the reference.
- The EMF is formed with W~, equal to W on gauged states. Rounding-level
  gradient components of the iterates then cannot reach the circuit.

Regions with sigma > 0 make the multiplier a physical unknown, and that is not
built: solve() raises for them.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from .assembler import Assembler, expand_material_map
from .transient import TransientSolver


class Winding:
    """Winding (src/Winding.m:1-40): region, turns, resistance, cross-section
    area, a uniform direction and an optional [3 x Ne] direction field."""

    def __init__(self, name, id_, N, R, S, dir_=(0.0, 0.0, 0.0)):
        self.Name = name
        self.RegionID = id_
        self.Turns = float(N)
        self.Resistance = float(R)
        self.CrossSectionArea = float(S)
        d = np.asarray(dir_, dtype=np.float64).ravel()
        nd = float(np.linalg.norm(d))
        self.Direction = d / nd if nd > 0 else np.zeros(3)
        self.DirectionField = None

    def setDirectionField(self, field):
        self.DirectionField = np.asarray(field, dtype=np.float64)


class CurrentSource:
    """CurrentSource (src/CurrentSource.m:1-20): a 'Domain' current density
    J_Vector [Jx, Jy, Jz] on one region.  ``as_source_map`` gives the {tag: J}
    form that Assembler.assembleSource accepts."""

    def __init__(self, regionID, J_vec):
        J = np.asarray(J_vec, dtype=np.float64).ravel()
        if J.size != 3:
            raise ValueError("J_Vector must be a 3-element vector.")
        self.Type = "Domain"
        self.RegionID = regionID
        self.J_Vector = J

    def as_source_map(self) -> dict:
        return {int(self.RegionID): self.J_Vector.copy()}


class Circuit:
    """One entry of the reference's circuit struct array: R, L and
    V_source_func(t)."""

    def __init__(self, R: float, L: float, V_source_func):
        self.R = float(R)
        self.L = float(L)
        self.V_source_func = V_source_func


def _as_list(x):
    return list(x) if isinstance(x, (list, tuple)) else [x]


class TransientBDF2Solver(TransientSolver):
    """TransientBDF2Solver: properties and defaults of
    src/solver/TransientBDF2Solver.m:12-24; first step BDF1, then BDF2."""

    Order = 2

    def __init__(self, assembler: Assembler):
        super().__init__(assembler)
        self.MaxIterations = 30
        self.Tolerance = 1e-3
        self.RelTolerance = 1e-3
        self.UseLineSearch = True
        self.MaxLineSearchIters = 10
        self.MinRelaxation = 1e-4
        # engine's linear solve: J + eps M with eps = GaugeRegRel x median
        # diag(K_hard) / median diag(M) (K_hard: vacuum reluctivity in the
        # nonlinear regions; iron is ~1e4 times softer, so the shift is
        # ~1e-9 of the iron's stiffness); PCG to LinearRelTol
        self.GaugeRegRel = 1e-13
        self.LinearRelTol = 1e-11
        self.LinearMaxIter = 40000
        self.AutoRegularization = True
        self.NewtonIterations = []
        self.PcgIterations = []

    # ------------------------------------------------------------ set-up
    def _bound_nu(self, matLib, mode):
        """getBoundNuVec (TransientBDF1Solver.m:283-297)."""
        tags = self.Assembler.Mesh.RegionTags
        nu = np.zeros(len(tags))
        nu0 = 1.0 / (4 * math.pi * 1e-7)
        for t in np.unique(tags):
            m = matLib.get(int(t)) if isinstance(matLib, dict) else None
            if m is None:
                continue
            if m.Type == "Linear":
                v = m.Nu_Linear
            else:
                v = nu0 if mode == "hard" else nu0 / 1000.0
            nu[tags == t] = v
        return nu

    def _diag(self, vals):
        return vals[self.diag_pos].cpu().numpy()

    def calculateCircuitScale(self, matLib, space, dt):
        """GeoBalance row scale per winding (TransientBDF1Solver.m:232-281)."""
        asm = self.Assembler
        kh = self._diag(asm.assembleStiffness(space, self._bound_nu(matLib, "hard"),
                                              drop_zeros=False).values)
        ks = self._diag(asm.assembleStiffness(space, self._bound_nu(matLib, "soft"),
                                              drop_zeros=False).values)
        dm = np.abs(self._diag(self.m_sigma))
        hard = np.abs(kh) + dm / dt
        k_hard = float(np.median(hard[hard > 1e-12]))
        soft = np.abs(ks) + dm / dt
        nl_tags = [int(t) for t, m in matLib.items() if m.Type == "Nonlinear"]
        if not nl_tags:
            k_soft = k_hard
        else:
            cell = asm.DofHandler.DofMaps[space.toString()].cpu().numpy()
            mask = np.isin(asm.Mesh.RegionTags, nl_tags)
            dofs = np.unique(cell[mask].ravel())
            dofs = dofs[dofs > 0]
            k_soft = float(np.median(soft[dofs - 1]))
        k_target = math.sqrt(k_hard * k_soft)
        self._k_hard = k_hard
        scale = np.zeros(len(self._circuits))
        for k, (c, w) in enumerate(zip(self._circuits, self._W_host)):
            z = abs(c.R + c.L / dt)
            nzw = np.abs(w[w != 0])
            max_w = float(nzw.max()) if nzw.size else 0.0
            mc = max(z, max_w / dt)
            if mc < 1e-20:
                mc = 1.0
            scale[k] = k_target / mc
        return scale

    def _project_windings(self, W):
        """W~ = W - M G Ln^-1 G^T W over the free DOFs (direct nodal solve)."""
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        asm = self.Assembler
        mesh = asm.Mesh
        edges = mesh.Edges.cpu().numpy().astype(np.int64)
        emap = asm.DofHandler.EntityMaps[self._space_key].cpu().numpy().astype(np.int64)
        fixed = self.fixed.cpu().numpy().astype(bool)
        n = self.n
        act = emap > 0
        rows = emap[act] - 1
        E = edges[act] - 1
        dn = np.zeros(mesh.NumNodes, dtype=bool)
        dn[E[fixed[rows]].ravel()] = True
        free = np.nonzero(~dn)[0]
        fid = np.full(mesh.NumNodes, -1, dtype=np.int64)
        fid[free] = np.arange(free.size)
        a, b = fid[E[:, 0]], fid[E[:, 1]]
        ma, mb = a >= 0, b >= 0
        G = sp.csr_matrix((np.concatenate([-np.ones(int(ma.sum())), np.ones(int(mb.sum()))]),
                           (np.concatenate([rows[ma], rows[mb]]),
                            np.concatenate([a[ma], b[mb]]))), shape=(n, free.size))
        M = sp.csc_matrix((self.m_unit.cpu().numpy(), self.rowind.cpu().numpy(),
                           self.colptr.cpu().numpy()), shape=(n, n))
        Ln = (G.T @ M @ G).tocsc()
        lu = spla.splu(Ln)
        out = []
        for w in self._W_host:
            wf = np.where(fixed, 0.0, w)
            phi = lu.solve(G.T @ wf)
            wt = wf - M @ (G @ phi)
            wt[fixed] = 0.0
            out.append(torch.from_numpy(wt).to(self.dev))
        self._div_ratio = [float(np.linalg.norm(G.T @ np.where(fixed, 0.0, w)) /
                                 max(np.linalg.norm(w), 1e-300)) for w in self._W_host]
        return out

    # ---------------------------------------------------------- residuals
    def _residual(self, a, I, calc_j: bool):
        """(R_A masked, R_c scaled, norm): the reference's Res_vec with the
        multiplier eliminated (its P rows vanish at every iterate)."""
        self._jacobian(a, calc_j)
        dadt = self._alpha * a + self._beta_a
        didt = self._alpha * I + self._beta_i
        r = self.res.clone()
        for k, wt in enumerate(self._Wt):
            r -= wt * float(I[k])
        r *= self.free
        emf = np.array([float(torch.dot(w, dadt)) for w in self._Wt])
        Rv = np.array([c.R for c in self._circuits])
        Lv = np.array([c.L for c in self._circuits])
        rc = (Rv * I + Lv * didt + emf - self._V) * self._scale
        nrm = math.sqrt(float(torch.dot(r, r)) + float(np.dot(rc, rc)))
        return r, rc, nrm

    def _pcg(self, rhs):
        if self._aux_key != self._c_key:
            self._build_aux()
            self._aux_key = self._c_key
        aux, keep = self._aux
        dx = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        it, rel = C.c_int(), C.c_double()
        L.check(L.lib.mpk_pcg_aux_solve(L.ptr(self.colptr), L.ptr(self.rowind),
                                        L.ptr(self.jsys), self.n, L.ptr(self.fixed),
                                        C.byref(aux), L.ptr(rhs * self.free), L.ptr(dx),
                                        L.ptr(keep["work"]), self.LinearRelTol,
                                        self.LinearMaxIter, self.LinearCheckEvery,
                                        C.byref(it), C.byref(rel), self.sp))
        self._pcg_it += it.value
        # the solve stands in for the reference's direct one: it must converge
        L.check_krylov(it, rel, self.LinearRelTol, "circuit Newton step")
        return dx

    # -------------------------------------------------------------- solve
    def solve(self, space_A, space_P, matLib, sigmaMap, circuitProps, windingObj, timeSteps,
              fixedDofs_A, fixedDofs_P=None, init_State=None, monitorHandle=None,
              probePoint=None):
        """Returns (x_final, info): x_final = [A (num_A); I (numWindings)] as
        numpy, info = {times, FinalTime, CurrentHistory [steps x windings],
        ProbeB_History (with probePoint)} as the reference's info struct."""
        asm = self.Assembler
        mesh = asm.Mesh
        sig = expand_material_map(sigmaMap, mesh.RegionTags)
        if np.any(np.abs(sig) > 1e-12):
            raise NotImplementedError("conducting regions make the gauge multiplier a physical "
                                      "unknown; only sigma = 0 systems are solved")
        asm.DofHandler.distributeDofs(space_A)
        self._setup(space_A, matLib, sigmaMap, fixedDofs_A)
        self._circuits = _as_list(circuitProps)
        windings = _as_list(windingObj)
        nw = len(windings)
        if len(self._circuits) != nw:
            raise ValueError("one circuit per winding")
        self._W = [asm.assembleWinding(space_A, w) for w in windings]
        self._W_host = [w.cpu().numpy() for w in self._W]
        # unit mass (the gauge's M) and the regularised constant part eps*M
        self.m_unit = self.m_reg
        dt0 = float(np.asarray(timeSteps, dtype=np.float64).ravel()[0])
        self._scale = self.calculateCircuitScale(matLib, space_A, dt0)
        dmu = np.abs(self._diag(self.m_unit))
        self.reg_abs = self.GaugeRegRel * self._k_hard / float(np.median(dmu[dmu > 0]))
        self._combine(self.cvals, self.reg_abs, self.m_unit)
        self._c_key = ("gauge", self.reg_abs)
        self._Wt = self._project_windings(self._W)
        num_A = asm.DofHandler.SpaceLocalSizes[space_A.toString()]
        a = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        I = np.zeros(nw)
        if init_State is not None:
            st = np.asarray(init_State, dtype=np.float64).ravel()
            a[: min(st.size, num_A)] = torch.from_numpy(st[:num_A]).to(self.dev)
            if st.size > num_A:
                I[: st.size - num_A] = st[num_A: num_A + nw]
        a_prev, a_prev2 = a.clone(), a.clone()
        I_prev, I_prev2 = I.copy(), I.copy()
        post = None
        if probePoint is not None:
            from .post import PostProcessor
            post = PostProcessor(asm)
            self._probe_e = int(post.locate(np.asarray(probePoint, dtype=np.float64))[0])
        steps = np.asarray(timeSteps, dtype=np.float64).ravel()
        ns = steps.size
        cur_hist = np.zeros((ns, nw))
        b_hist = np.zeros(ns)
        times = np.zeros(ns)
        t = 0.0
        self.NewtonIterations, self.PcgIterations = [], []
        for k in range(ns):
            dt = float(steps[k])
            t += dt
            times[k] = t
            if k == 0 or self.Order == 1:
                self._alpha = 1.0 / dt
                self._beta_a = -a_prev / dt
                self._beta_i = -I_prev / dt
            else:
                self._alpha = 1.5 / dt
                self._beta_a = (-2.0 * a_prev + 0.5 * a_prev2) / dt
                self._beta_i = (-2.0 * I_prev + 0.5 * I_prev2) / dt
            self._V = np.array([c.V_source_func(t) for c in self._circuits], dtype=np.float64)
            self._pcg_it = 0
            init_norm = 0.0
            iters = 0
            for it in range(1, self.MaxIterations + 1):
                r, rc, nrm = self._residual(a, I, calc_j=True)
                if it == 1:
                    init_norm = nrm if nrm >= 1e-12 else 1.0
                rel = nrm / init_norm
                iters = it
                if nrm < self.Tolerance or (it > 1 and rel < self.RelTolerance):
                    break
                # Newton matrix J + eps M on the pattern; Schur complement for I
                self._combine(self.jsys, 1.0, self.jv, 1.0, self.cvals)
                u = self._pcg(r)
                V = [self._pcg(wt) for wt in self._Wt]
                Rv = np.array([c.R for c in self._circuits])
                Lv = np.array([c.L for c in self._circuits])
                S = np.zeros((nw, nw))
                rhs = -rc / self._scale
                for i in range(nw):
                    rhs[i] += self._alpha * float(torch.dot(self._Wt[i], u))
                    for j in range(nw):
                        S[i, j] = self._alpha * float(torch.dot(self._Wt[i], V[j]))
                    S[i, i] += Rv[i] + Lv[i] * self._alpha
                dI = np.linalg.solve(S, rhs)
                da = -u
                for j in range(nw):
                    da = da + V[j] * float(dI[j])
                step = 1.0
                if self.UseLineSearch:
                    for _ in range(self.MaxLineSearchIters + 1):
                        _, _, trial = self._residual(a + step * da, I + step * dI, calc_j=False)
                        if trial < nrm * 1.0001:
                            break
                        step *= 0.5
                a = a + step * da
                I = I + step * dI
            self.NewtonIterations.append(iters)
            self.PcgIterations.append(self._pcg_it)
            a_prev2, a_prev = a_prev, a.clone()
            I_prev2, I_prev = I_prev, I.copy()
            cur_hist[k] = I
            if post is not None:
                b_hist[k] = self._probe_b(a) if self._probe_e >= 0 else 0.0
            if monitorHandle is not None:
                monitorHandle(t, I.copy(), times[: k + 1].copy(), cur_hist[: k + 1].copy(),
                              b_hist[k], b_hist[: k + 1].copy())
        info = {"times": times, "FinalTime": t, "CurrentHistory": cur_hist.squeeze()
                if nw == 1 else cur_hist}
        if probePoint is not None:
            info["ProbeB_History"] = b_hist
        x = np.concatenate([a.cpu().numpy()[:num_A], I])
        return x, info

    def _probe_b(self, a) -> float:
        B = self.Assembler.computeElementB(a)
        return float(torch.linalg.vector_norm(B[self._probe_e]))


class TransientBDF1Solver(TransientBDF2Solver):
    """TransientBDF1Solver: backward Euler at every step
    (src/solver/TransientBDF1Solver.m:124-133), same properties."""

    Order = 1
