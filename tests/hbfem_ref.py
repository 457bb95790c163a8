"""CPU restatement of HBFEMSolver.solve and HBFEMCoupledSolver.solve
(src/solver/HBFEMSolver.m:29-183, src/solver/HBFEMCoupledSolver.m:37-352) on the
oracle's HBFEM and winding kernels, with the reference's monolithic Newton
system solved by scipy's sparse LU in place of MUMPS: the checker of the
harmonic-balance solvers in magpackage_amd.hbfem.

The residual is sparse(R_rows, R_cols, R_val) of the HBFEM kernel [nd x H]. The
Jacobian is sparse(I, J, V) over nd*H, plus the harmonic-repeated K_add block.
The circuit blocks follow HBFEMCoupledSolver.m:186-228. Dirichlet handling
follows applyDirichlet. The line search, the forced-step branch and the stop
test follow the reference. Test infrastructure only.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from magpackage_amd.material import pack_material_data
from tests.helpers import mass_tables, stiffness_tables


def _kernel(mi, tags, X, aft, packed, request_j):
    qw, curl, _ = stiffness_tables()
    n = mi["ndof"]
    H = aft.NumHarmonics
    out = O.hbfem(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, X.real, X.imag,
                  aft.D_matrix.real, aft.D_matrix.imag, aft.P_matrix.real, aft.P_matrix.imag,
                  qw, curl, packed, request_j, aft.jacobian_scalings())
    I, J, V, Rr, Rc, Rvr, Rvi = out
    R = np.zeros((n, H), dtype=complex)
    ok = (Rr > 0) & (Rc > 0)
    np.add.at(R, (Rr[ok].astype(int) - 1, Rc[ok].astype(int) - 1), Rvr[ok] + 1j * Rvi[ok])
    Jm = None
    if request_j:
        ok = (I > 0) & (J > 0)
        Jm = sp.csc_matrix((V[ok], (I[ok].astype(int) - 1, J[ok].astype(int) - 1)),
                           shape=(n * H, n * H))
    return Jm, R


def _dirichlet(A, b, fixed):
    f = (~fixed).astype(float)
    Ab = (sp.diags(f) @ A @ sp.diags(f) + sp.diags(fixed.astype(float))).tocsc()
    bb = b.copy()
    bb[fixed] = 0.0
    return Ab, bb


def hbfem_coupled_cpu(mesh, lib, aft, winding, Rc, V_harm, fixed, kadd, tol=1e-4,
                      max_iter=50, max_bt=20, bt_factor=0.5, min_step=1e-6):
    """winding: dict(region, turns, area, dir (Ne, 3)).  kadd: scipy [nd x nd]
    or None.  Returns (A [nd x H], I [H], iterations, residual history)."""
    mi = O.mesh_inputs(mesh)
    n = mi["ndof"]
    H = aft.NumHarmonics
    w = 2 * np.pi * aft.BaseFreq * aft.Harmonics
    packed = pack_material_data(lib)
    tags = mesh.RegionTags.astype(np.float64)
    qm, val, _ = mass_tables(2)
    Iw, Vw = O.winding(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], tags, winding["region"],
                       winding["dir"], np.zeros(3), winding["turns"] / winding["area"], qm, val)
    Cv = np.zeros(n)
    ok = Iw > 0
    np.add.at(Cv, Iw[ok].astype(int) - 1, Vw[ok])
    Kg = sp.block_diag([kadd] * H).tocsc() if kadd is not None else None
    fixed_sys = np.concatenate([np.tile(fixed, H), np.zeros(H, dtype=bool)])
    free = ~fixed_sys
    A = np.zeros((n, H), dtype=complex)
    I = np.zeros(H, dtype=complex)
    Vh = np.asarray(V_harm, dtype=complex)

    def residuals(Ac, Ic, rj):
        Jm, R = _kernel(mi, tags, Ac, aft, packed, rj)
        rf = (R - Cv[:, None] * Ic[None, :]).ravel(order="F")
        if Kg is not None:
            rf = rf + Kg @ Ac.ravel(order="F")
        rcir = Rc * Ic + 1j * w * (Cv @ Ac) - Vh
        return Jm, rf, rcir

    _, rf, rcir = residuals(A, I, False)
    mag = np.mean(np.abs(rf))
    mag = mag if mag >= 1e-5 else 1.0
    base = abs(Rc) if abs(Rc) >= 1e-10 else 1.0
    rs = min(max(mag / base, 1e3), 1e9)
    cur = np.linalg.norm(np.concatenate([rf, rcir * rs])[free])
    norm_base = cur if cur >= 1e-5 else 1.0
    hist = [cur]
    it = 0
    for it in range(1, max_iter + 1):
        Jf, rf, rcir = residuals(A, I, True)
        if Kg is not None:
            Jf = Jf + Kg
        ci = np.nonzero(Cv)[0]
        cv = Cv[ci]
        rows, cols, vals = [], [], []
        for k in range(H):
            off = k * n
            rows += [ci + off]
            cols += [np.full(ci.size, n * H + k)]
            vals += [-cv + 0j]
            rows += [np.full(ci.size, n * H + k)]
            cols += [ci + off]
            vals += [1j * w[k] * cv * rs]
        rows += [n * H + np.arange(H)]
        cols += [n * H + np.arange(H)]
        vals += [np.full(H, Rc * rs, dtype=complex)]
        Cpl = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                            shape=(n * H + H, n * H + H))
        Jsys = (sp.block_diag([Jf, sp.csc_matrix((H, H))]) + Cpl).tocsc()
        rhs = -np.concatenate([rf, rcir * rs])
        Jb, rb = _dirichlet(Jsys, rhs, fixed_sys)
        dx = spla.splu(Jb.astype(complex), permc_spec="MMD_AT_PLUS_A").solve(rb)
        dA = dx[: n * H].reshape((n, H), order="F")
        dI = dx[n * H:]
        alpha, accepted = 1.0, False
        for _ in range(max_bt + 1):
            At, It = A + alpha * dA, I + alpha * dI
            _, rft, rct = residuals(At, It, False)
            new = np.linalg.norm(np.concatenate([rft, rct * rs])[free])
            if new < cur or alpha < min_step:
                A, I, cur, accepted = At, It, new, True
                break
            alpha *= bt_factor
        if not accepted:
            A, I = A + 1e-3 * dA, I + 1e-3 * dI
        hist.append(cur)
        if cur / norm_base < tol:
            break
    return A, I, it, hist
