"""CPU restatement of TransientSolver.solve, pure-A branch
(src/solver/TransientSolver.m:38-267, 274-311), on the oracle's kernels and
scipy's direct solver: the checker of magpackage_amd.transient.

Steps per time step: F from the source kernel + sparse (src/Assembler.m:78-93),
b = F + M_sigma x_prev / dt; Newton on R = R_nu + M_sigma (x - x_prev)/dt - F
(+ reg M_reg x) with J = J_nu + M_sigma/dt (+ reg M_reg); reg fixed at the first
iteration of the first step from mean |diag(J_nu + M_sigma/dt)| (:150-158);
applyDirichlet (src/BoundaryCondition.m:81-112); convergence and line search
as :173-237.  Test infrastructure only.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from magpackage_amd.material import pack_material_data
from tests.helpers import mass_tables, stiffness_tables


def apply_dirichlet(K, F, fixed):
    free = (~fixed).astype(np.float64)
    Df = sp.diags(free)
    Ks = (Df @ K @ Df + sp.diags(fixed.astype(np.float64))).tocsc()
    Fs = F.copy()
    Fs[fixed] = 0.0
    return Ks, Fs


def transient_cpu(mesh, lib, sigma_e, source, time_steps, fixed, tol=1e-5, rel_tol=1e-6,
                  max_iter=50, reg_rel=1e-5, max_ls=10, load=None):
    """source(t) -> (Ne, 3) current density, or load(t) -> the assembled
    source vector [n].  Returns (solutions, newton iteration counts)."""
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    P, T, D, S = mi["P"], mi["T"], mi["Dofs"], mi["Signs"]
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    tags = mesh.RegionTags.astype(np.float64)
    Msig = O.sparse_scipy(*O.mass(P, T, D, S, sigma_e, qm, val), n, n, drop_zeros=False)
    Mreg = O.sparse_scipy(*O.mass(P, T, D, S, None, qm, val), n, n, drop_zeros=False)

    def assemble_source(t):
        if load is not None:
            return np.asarray(load(t), dtype=np.float64)
        I, V = O.source(P, T, D, S, source(t), qm, val)
        F = np.zeros(n)
        np.add.at(F, I.astype(np.int64) - 1, V)
        return F

    def nu_part(x, calc_j):
        I, J, V, Ri, Rv = O.jacobian(P, T, D, S, tags, x, qw, curl, packed, calcJ=calc_j)
        R = np.zeros(n)
        np.add.at(R, Ri.astype(np.int64) - 1, Rv)
        Jm = O.sparse_scipy(I, J, V, n, n, drop_zeros=False) if calc_j else None
        return Jm, R

    x_prev = np.zeros(n)
    reg = 0.0
    sols, iters = [], []
    t = 0.0
    for dt in time_steps:
        t += dt
        b = assemble_source(t) + Msig @ x_prev / dt
        x = x_prev.copy()
        initial = 1.0
        for it in range(1, max_iter + 1):
            Jn, Rn = nu_part(x, True)
            Jsys = (Jn + Msig / dt).tocsc()
            if reg == 0.0:
                d = np.abs(Jsys.diagonal())
                d = d[d > 0]
                reg = (d.mean() if d.size else 1.0) * reg_rel
            Cm = (Msig / dt + reg * Mreg).tocsc()
            Jsys = (Jn + Cm).tocsc()
            R = Rn + Cm @ x - b
            Jbc, Rbc = apply_dirichlet(Jsys, R, fixed)
            norm = np.linalg.norm(Rbc)
            if it == 1:
                initial = norm if norm >= 1e-20 else 1.0
            if norm < tol or norm / initial < rel_tol:
                break
            dx = spla.spsolve(Jbc, -Rbc)
            alpha = 1.0
            for k in range(max_ls + 1):
                a_try = alpha * 0.5 ** k
                xt = x + a_try * dx
                _, Rt = nu_part(xt, False)
                Rt = Rt + Cm @ xt - b
                Rt[fixed] = 0.0
                rt = np.linalg.norm(Rt)
                if rt < norm or (it == 1 and k < 3):
                    alpha = a_try
                    break
                if rt > norm * 10:
                    alpha *= 0.1
            x = x + alpha * dx
        sols.append(x.copy())
        iters.append(it)
        x_prev = x
    return sols, iters


def transient_av_cpu(mesh, lib, sigma_e, source, time_steps, fixed_A, v_fixed_nodes=None,
                     tol=1e-5, rel_tol=1e-6, max_iter=50, reg_rel=1e-5, max_ls=10,
                     line_search=True, load=None):
    """TransientSolver.solve with the scalar-potential space
    (src/solver/TransientSolver.m:35-267, evaluateSystemState :287-311).
    Unknowns [A (1..nA); V (nodes of the conducting elements, increasing
    id)]. Per Newton iteration:
      R = R_nu + M_sigma (x - x_prev)/dt - F + C V + [C^T (A - A_prev)/dt + L V]
          (+ reg M_reg x),
      J = J_nu + M_sigma/dt + C + C^T/dt + L (+ reg M_reg),
    with C = assembleCoupling(A, V, sigma) and L = assembleScalarLaplacian(V,
    sigma); reg = RegScaleRel * mean of the positive |diag J| at the first
    iteration of the first step; fixed DOFs: fixed_A, the given grounded
    nodes, else the first V DOF. The line search adds the regularisation
    before zeroing the fixed rows (:222-226). With linear materials only
    (checkLinearity) there is no line search and a step stops after an update
    from a residual below 1e-9 (:207, :240). Returns (solutions [N], Newton
    iteration counts, nA)."""
    from magpackage_amd.fem_tables import get_quadrature_data, lagrange_tet_p1, matlab_flat
    mi = O.mesh_inputs(mesh)
    nA = mi["ndof"]
    P, T, D, S = mi["P"], mi["T"], mi["Dofs"], mi["Signs"]
    packed = pack_material_data(lib)
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    tags = mesh.RegionTags.astype(np.float64)
    sig = np.asarray(sigma_e, dtype=np.float64)
    T0 = np.asarray(mesh.T, dtype=np.int64) - 1
    nodes = np.unique(T0[sig > 1e-12].ravel())
    nmap = np.zeros(mesh.P.shape[0], dtype=np.int64)
    nmap[nodes] = nA + 1 + np.arange(nodes.size)
    N = nA + nodes.size

    def csc(I, J, V):
        ok = (I > 0) & (J > 0)
        return sp.csc_matrix((V[ok], (I[ok].astype(np.int64) - 1, J[ok].astype(np.int64) - 1)),
                             shape=(N, N))

    Msig = csc(*O.mass(P, T, D, S, sig, qm, val))
    Mreg = csc(*O.mass(P, T, D, S, None, qm, val))
    DN = nmap[T0].astype(np.float64)
    q1, w1 = get_quadrature_data("tet", 1)
    Lvv = csc(*O.scalar_laplacian(P, T, DN, sig, w1, matlab_flat(lagrange_tet_p1(q1)[1])))
    q2, w2 = get_quadrature_data("tet", 2)
    Cav = csc(*O.coupling(P, T, D, S, DN, sig, w2, val, matlab_flat(lagrange_tet_p1(q2)[1])))
    fixed = np.zeros(N, dtype=bool)
    fixed[:nA] = np.asarray(fixed_A, dtype=bool)[:nA]
    if v_fixed_nodes is None or len(v_fixed_nodes) == 0:
        fixed[nA] = True
    else:
        d = nmap[np.asarray(v_fixed_nodes, dtype=np.int64)]
        fixed[d[d > 0] - 1] = True

    def assemble_source(t):
        F = np.zeros(N)
        if load is not None:
            F[:nA] = np.asarray(load(t), dtype=np.float64)[:nA]
            return F
        I, V = O.source(P, T, D, S, source(t), qm, val)
        np.add.at(F, I.astype(np.int64) - 1, V)
        return F

    def state(x, x_prev, dt, F, calc_j):
        I, J, V, Ri, Rv = O.jacobian(P, T, D, S, tags, x[:nA], qw, curl, packed, calcJ=calc_j)
        R = np.zeros(N)
        np.add.at(R, Ri.astype(np.int64) - 1, Rv)
        R = R + Msig @ (x - x_prev) / dt - F + Cav @ x + (Cav.T @ (x - x_prev)) / dt + Lvv @ x
        return (csc(I, J, V) if calc_j else None), R

    linear = all(getattr(m, "Type", "Linear") == "Linear" for m in lib.values())
    x_prev = np.zeros(N)
    reg = 0.0
    sols, iters = [], []
    t = 0.0
    for dt in time_steps:
        t += dt
        F = assemble_source(t)
        Jc = (Msig / dt + Cav + Cav.T / dt + Lvv).tocsc()
        x = x_prev.copy()
        initial = 1.0
        for it in range(1, max_iter + 1):
            Jn, R = state(x, x_prev, dt, F, True)
            Jsys = (Jn + Jc).tocsc()
            if reg == 0.0:
                d = np.abs(Jsys.diagonal())
                d = d[d > 0]
                reg = (d.mean() if d.size else 1.0) * reg_rel
            Jsys = (Jsys + reg * Mreg).tocsc()
            R = R + reg * (Mreg @ x)
            Jbc, Rbc = apply_dirichlet(Jsys, R, fixed)
            norm = np.linalg.norm(Rbc)
            if it == 1:
                initial = norm if norm >= 1e-20 else 1.0
            if norm < tol or norm / initial < rel_tol:
                break
            dx = spla.splu(Jbc, permc_spec="MMD_AT_PLUS_A").solve(-Rbc)
            alpha = 1.0
            if line_search and not linear:
                for k in range(max_ls + 1):
                    a_try = alpha * 0.5 ** k
                    xt = x + a_try * dx
                    _, Rt = state(xt, x_prev, dt, F, False)
                    Rt = Rt + reg * (Mreg @ xt)
                    Rt[fixed] = 0.0
                    rt = np.linalg.norm(Rt)
                    if rt < norm or (it == 1 and k < 3):
                        alpha = a_try
                        break
                    if rt > norm * 10:
                        alpha *= 0.1
            x = x + alpha * dx
            if linear and norm < 1e-9:
                break
        sols.append(x.copy())
        iters.append(it)
        x_prev = x
    return sols, iters, nA
