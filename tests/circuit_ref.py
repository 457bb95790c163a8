"""CPU restatement of TransientBDF2Solver.solve / TransientBDF1Solver.solve
(src/solver/TransientBDF2Solver.m:44-238, TransientBDF1Solver.m:35-335) on the
oracle's kernels, with the reference's monolithic saddle-point Newton system
solved by scipy's sparse LU in place of MUMPS: the checker of
magpackage_amd.circuit.

Unknowns [A (edge DOFs); P (one per node, DOF num_A + node id, as
DofHandler.distributeDofs numbers a whole-mesh Lagrange space); I (one per
winding)].  Per Newton iteration (assembleNewtonSystem, :300-345):
  J_sys = J_fem + alpha M_sigma + C_AP + C_AP' - [W col block]
          + [alpha Scale W' row block] + diag(Scale (R + L alpha)),
  Res   = R_fem + M_sigma dxdt + (C_AP + C_AP') x - W I,
  Res_c = Scale (R I + L dI/dt + W' dxdt - V),
Dirichlet rows/columns of the fixed A and P DOFs removed with a unit diagonal
(applyDirichlet), convergence on |Res(free)| and line search as :168-213.
The GeoBalance scale follows calculateCircuitScale (:232-281).  Test
infrastructure only.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from magpackage_amd.fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
from magpackage_amd.material import pack_material_data
from tests.helpers import mass_tables, stiffness_tables


def boundary_nodes(T):
    """Nodes of faces that belong to one tet (0-based)."""
    T0 = np.asarray(T, dtype=np.int64) - 1
    f = np.concatenate([T0[:, [0, 1, 2]], T0[:, [0, 1, 3]], T0[:, [0, 2, 3]], T0[:, [1, 2, 3]]])
    fs = np.sort(f, axis=1)
    u, cnt = np.unique(fs, axis=0, return_counts=True)
    return np.unique(u[cnt == 1].ravel())


def locate(P, T, point):
    """0-based tet holding `point` (first by index), -1 outside."""
    T0 = np.asarray(T, dtype=np.int64) - 1
    p0 = P[T0[:, 0]]
    J = np.stack([P[T0[:, 1]] - p0, P[T0[:, 2]] - p0, P[T0[:, 3]] - p0], axis=2)
    x = np.einsum("eij,ej->ei", np.linalg.inv(J), np.asarray(point)[None, :] - p0)
    lam = np.minimum(x.min(axis=1), 1.0 - x.sum(axis=1))
    hit = np.nonzero(lam >= -1e-12)[0]
    return int(hit[0]) if hit.size else -1


def circuit_bdf_cpu(mesh, lib, windings, circuits, time_steps, probe, order=2, tol=1e-3,
                    rel_tol=1e-3, max_iter=30, max_ls=10):
    """windings: list of dicts {region, turns, area, dir (Ne, 3)}; circuits:
    list of dicts {R, L, V(t)}.  Returns (current history [steps, nw],
    probe |B| history [steps], Newton iteration counts)."""
    mi = O.mesh_inputs(mesh)
    nA = mi["ndof"]
    P, T, D, S = mi["P"], mi["T"], mi["Dofs"], mi["Signs"]
    nn = mesh.P.shape[0]
    nfem = nA + nn
    nw = len(windings)
    ntot = nfem + nw
    ne = T.shape[0]
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    q_pts, qw2 = get_quadrature_data("tet", 2)
    grad = matlab_flat(lagrange_tet_p1(q_pts)[1])
    tags = mesh.RegionTags.astype(np.float64)
    DN = (nA + T).astype(np.float64)            # node DOFs num_A + node id
    Ic, Jc, Vc = O.coupling(P, T, D, S, DN, np.ones(ne), qw2, val, grad)
    C_AP = sp.csc_matrix((Vc, (Ic.astype(np.int64) - 1, Jc.astype(np.int64) - 1)),
                         shape=(ntot, ntot))
    W = np.zeros((ntot, nw))
    for k, w in enumerate(windings):
        Iw, Vw = O.winding(P, T, D, S, tags, w["region"], w["dir"], np.zeros(3),
                           w["turns"] / w["area"], qm, val)
        ok = Iw > 0
        np.add.at(W[:, k], Iw[ok].astype(np.int64) - 1, Vw[ok])
    # Dirichlet: outer-boundary edges and nodes
    fixedA = O.boundary_edges(mesh.T, mi["edges"])
    emap = np.zeros(mi["edges"].shape[0], dtype=np.int64)
    fixed = np.zeros(ntot, dtype=bool)
    cell = mi["cell"].astype(np.int64)
    t2e = mi["t2e"].astype(np.int64)
    emap[t2e.ravel() - 1] = cell.ravel()
    fdofs = emap[fixedA]
    fixed[fdofs[fdofs > 0] - 1] = True
    fixed[nA + boundary_nodes(mesh.T)] = True
    fixed[nfem:] = False

    def fem(x, calc_j):
        Ij, Jj, Vj, Ri, Rv = O.jacobian(P, T, D, S, tags, x[:nA], qw, curl, packed, calcJ=calc_j)
        R = np.zeros(ntot)
        np.add.at(R, Ri.astype(np.int64) - 1, Rv)
        Jm = (sp.csc_matrix((Vj, (Ij.astype(np.int64) - 1, Jj.astype(np.int64) - 1)),
                            shape=(ntot, ntot)) if calc_j else None)
        return Jm, R

    # GeoBalance (calculateCircuitScale)
    def nu_bound(mode):
        nu = np.zeros(ne)
        for tg, m in lib.items():
            if m.Type == "Linear":
                v = m.Nu_Linear
            else:
                v = 1 / (4 * math.pi * 1e-7) * (1.0 if mode == "hard" else 1e-3)
            nu[mesh.RegionTags == tg] = v
        return nu

    def kdiag(nu):
        Ik, Jk, Vk = O.curl_curl(P, T, D, S, nu, qw, curl)
        K = sp.csc_matrix((Vk, (Ik.astype(np.int64) - 1, Jk.astype(np.int64) - 1)),
                          shape=(nfem, nfem))
        return np.abs(K.diagonal())

    dt0 = float(time_steps[0])
    hard = kdiag(nu_bound("hard"))
    k_hard = np.median(hard[hard > 1e-12])
    soft = kdiag(nu_bound("soft"))
    nl = [tg for tg, m in lib.items() if m.Type == "Nonlinear"]
    if nl:
        dofs = np.unique(cell[np.isin(mesh.RegionTags, nl)].ravel())
        dofs = dofs[dofs > 0]
        k_soft = np.median(soft[dofs - 1])
    else:
        k_soft = k_hard
    k_target = math.sqrt(k_hard * k_soft)
    scale = np.zeros(nw)
    for k, c in enumerate(circuits):
        z = abs(c["R"] + c["L"] / dt0)
        mw = np.abs(W[:, k][W[:, k] != 0]).max()
        scale[k] = k_target / max(z, mw / dt0)
    Rv_ = np.array([c["R"] for c in circuits])
    Lv_ = np.array([c["L"] for c in circuits])
    CC = (C_AP + C_AP.T).tocsc()

    def residual(x, ctx, calc_j):
        Jf, Rf = fem(x, calc_j)
        dxdt = ctx["alpha"] * x + ctx["beta"]
        Res = Rf + CC @ x - W @ x[nfem:]
        emf = W.T @ dxdt
        I = x[nfem:]
        Res[nfem:] = (Rv_ * I + Lv_ * dxdt[nfem:] + emf - ctx["V"]) * scale
        return Jf, Res

    x = np.zeros(ntot)
    x_prev, x_prev2 = x.copy(), x.copy()
    e_probe = locate(P, mesh.T, probe)
    cur, bh, iters = [], [], []
    t = 0.0
    for k, dt in enumerate(time_steps):
        t += dt
        if k == 0 or order == 1:
            ctx = {"alpha": 1.0 / dt, "beta": -x_prev / dt}
        else:
            ctx = {"alpha": 1.5 / dt, "beta": (-2.0 * x_prev + 0.5 * x_prev2) / dt}
        ctx["V"] = np.array([c["V"](t) for c in circuits])
        init = 1.0
        n_it = 0
        for it in range(1, max_iter + 1):
            Jf, Res = residual(x, ctx, True)
            chk = Res.copy()
            chk[fixed] = 0.0
            nrm = np.linalg.norm(chk)
            if it == 1:
                init = nrm if nrm >= 1e-12 else 1.0
            n_it = it
            if nrm < tol or (it > 1 and nrm / init < rel_tol):
                break
            wr, wc = np.nonzero(W)
            wv = W[wr, wc]
            Wcol = sp.csc_matrix((-wv, (wr, nfem + wc)), shape=(ntot, ntot))
            Wrow = sp.csc_matrix((ctx["alpha"] * wv * scale[wc], (nfem + wc, wr)),
                                 shape=(ntot, ntot))
            dg = np.zeros(ntot)
            dg[nfem:] = (Rv_ + Lv_ * ctx["alpha"]) * scale
            Jsys = (Jf + CC + Wcol + Wrow + sp.diags(dg)).tocsc()
            free = (~fixed).astype(np.float64)
            Jbc = (sp.diags(free) @ Jsys @ sp.diags(free) + sp.diags(fixed.astype(float))).tocsc()
            rb = Res.copy()
            rb[fixed] = 0.0
            dx = spla.splu(Jbc, permc_spec="MMD_AT_PLUS_A").solve(-rb)
            a = 1.0
            for _ in range(max_ls + 1):
                _, Rt = residual(x + a * dx, ctx, False)
                Rt[fixed] = 0.0
                if np.linalg.norm(Rt) < nrm * 1.0001:
                    break
                a *= 0.5
            x = x + a * dx
        iters.append(n_it)
        x_prev2, x_prev = x_prev, x.copy()
        cur.append(x[nfem:].copy())
        if e_probe >= 0:
            B = O.element_b(P, T, D, S, x[:nA], curl)
            bh.append(float(np.linalg.norm(B[e_probe])))
        else:
            bh.append(0.0)
    return np.array(cur), np.array(bh), iters, {"scale": scale, "W": W[:nA], "nA": nA}


def inductance_matrix(km, nu, windings):
    """L_ij = W_i^T A_j for K A_j = W_j with the gauge G^T M A = 0, solved as
    (K + gamma M G Ln^-1 G^T M) A = W~ on the free edges (SPD, equal to K on
    gauged fields), W~ the M-orthogonal divergence-free part of W.  A
    gauge-free route to the inductances, independent of the saddle point.
    windings: dicts {region, turns, area, dir (Ne, 3)}."""
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    P, T, D, S = mi["P"], mi["T"], mi["Dofs"], mi["Signs"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    K = O.sparse_scipy(*O.curl_curl(P, T, D, S, nu, qw, curl), n, n, drop_zeros=False)
    M = O.sparse_scipy(*O.mass(P, T, D, S, None, qm, val), n, n, drop_zeros=False)
    W = np.zeros((n, len(windings)))
    for k, w in enumerate(windings):
        Iw, Vw = O.winding(P, T, D, S, km.RegionTags.astype(np.float64), w["region"], w["dir"],
                           np.zeros(3), w["turns"] / w["area"], qm, val)
        ok = Iw > 0
        np.add.at(W[:, k], Iw[ok].astype(np.int64) - 1, Vw[ok])
    edges = mi["edges"]
    emap = np.zeros(edges.shape[0], dtype=np.int64)
    emap[mi["t2e"].ravel() - 1] = mi["cell"].ravel()
    fixed = np.zeros(n, dtype=bool)
    fd = emap[O.boundary_edges(km.T, edges)]
    fixed[fd[fd > 0] - 1] = True
    free_nodes = np.setdiff1d(np.arange(km.P.shape[0]), boundary_nodes(km.T))
    fid = np.full(km.P.shape[0], -1)
    fid[free_nodes] = np.arange(free_nodes.size)
    rows = emap - 1
    a, b = fid[edges[:, 0] - 1], fid[edges[:, 1] - 1]
    ma, mb = (a >= 0) & (rows >= 0), (b >= 0) & (rows >= 0)
    G = sp.csr_matrix((np.r_[-np.ones(ma.sum()), np.ones(mb.sum())],
                       (np.r_[rows[ma], rows[mb]], np.r_[a[ma], b[mb]])),
                      shape=(n, free_nodes.size))
    f = ~fixed
    Kf, Mf, Gf, Wf = K[f][:, f], M[f][:, f], G[f], W[f]
    Ln = (Gf.T @ Mf @ Gf).toarray()
    MG = (Mf @ Gf).toarray()
    Wt = Wf - MG @ np.linalg.solve(Ln, Gf.T @ Wf)
    # the penalty vanishes on gauged fields for any weight; this one matches K's scale
    gamma = Kf.diagonal().mean() / Mf.diagonal().mean()
    A = np.linalg.solve(Kf.toarray() + gamma * MG @ np.linalg.solve(Ln, MG.T), Wt)
    # gauged: G^T M A = 0 up to rounding of its own terms
    assert np.all(np.abs(MG.T @ A) <= 1e-9 * (np.abs(MG.T) @ np.abs(A)))
    return Wf.T @ A
