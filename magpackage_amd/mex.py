"""MEX-signature drop-ins: the reference's kernel entry points on the GPU.

Same names, argument lists, argument meaning, output layout and error
identifier as the MEX files that ``Assembler.m`` calls:

- ``assemble_curl_curl_kernel_mex(P, T, CellDofs, Signs, Nu, Qw, CurlRef)``
  src/mex/assemble_curl_curl_kernel_mex.cpp:18-121 (7 inputs)
- ``assemble_mass_kernel_mex(P, T, CellDofs, Signs, Coeffs, Qw, ValRef)``
  src/mex/assemble_mass_kernel_mex.cpp:16-127 (7 inputs, Coeffs may be [])
- ``assemble_jacobian_kernel_mex(P, T, CellDofs, Signs, Tags, SolA, Qw,
  CurlRef, linNu, isNon, maxBSq, breaks, coefs, starts, counts, coeStarts,
  calcJ)`` src/mex/assemble_jacobian_kernel_mex.cpp:7-153 (17 inputs)
- ``assemble_source_kernel_mex(P, T, Dofs, Signs, Js, Qw, ValRef)``
  src/mex/assemble_source_kernel_mex.cpp:16-120 (7 inputs)
- ``assemble_scalar_laplacian_kernel_mex(P, T, Dofs, Coeffs, Qw, GradRef)``
  src/mex/assemble_scalar_laplacian_kernel_mex.cpp:13-105 (6 inputs)
- ``assemble_coupling_kernel_mex(P, T, DofsE, Signs, DofsN, Coeffs, Qw, NedRef,
  GradRef)`` src/mex/assemble_coupling_kernel_mex.cpp:13-131 (9 inputs)
- ``assemble_winding_kernel_mex(P, T, Dofs, Signs, RegionTags, CoilID,
  DirField, UniformDir, Scale, Qw, RefVal)``
  src/mex/assemble_winding_kernel_mex.cpp:13-126 (11 inputs, DirField may be [])
- ``assemble_hbfem_kernel_mex(P, T, Dofs, Signs, Tags, SolH_Real, SolH_Imag,
  Dmat_Real, Dmat_Imag, Pmat_Real, Pmat_Imag, Qw, CurlRef, linNu, isNon,
  maxBSq, breaks, coefs, starts, counts, RequestJacobian, Scalings[,
  coeStarts])`` src/mex/assemble_hbfem_kernel_mex.cpp:25-260 (22 inputs; the
  MEX reads an uninitialised coefficient-offset table, here it is the optional
  23rd input or derived as packMaterialData builds it)
- ``sparse_gpu(I, J, V, m, n)``: ``valid = I>0 & J>0; sparse(...)``
  (src/Assembler.m:37-38) on the GPU.

Inputs are host arrays in MATLAB shapes (``P`` [3 x Nn], ``T`` [4 x Ne],
...), read in column-major order exactly like ``mxGetPr``; outputs are host
column vectors.  Work runs on ``cuda:0`` through libmpk_assembly.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _colmajor(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F"))


def _ncols(x) -> int:
    a = np.asarray(x)
    return int(a.shape[1]) if a.ndim >= 2 else (1 if a.size else 0)


def _dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(a).to(torch.device("cuda", torch.cuda.current_device()))


def _arity(args, n):
    if len(args) != n:
        raise L.MpkError(L.MPK_ERR_ARGS, f"Expected {n} inputs.")


def _out(n):
    return torch.empty(max(n, 1), dtype=torch.float64,
                       device=torch.device("cuda", torch.cuda.current_device()))


def _quad(Qw, table, curl: bool) -> L.Quadrature:
    qw = _colmajor(Qw)
    tab = _colmajor(table)
    if tab.size < 18 * qw.size:
        raise L.MpkError(L.MPK_ERR_ARGS, "reference table smaller than 6 x 3 x numel(Qw)")
    return L.make_quadrature(qw, None if curl else tab[: 18 * qw.size],
                             tab[: 18 * qw.size] if curl else None)


def assemble_curl_curl_kernel_mex(*args):
    _arity(args, 7)
    P, T, Dofs, Signs, Nu, Qw, CurlRef = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    q = _quad(Qw, CurlRef, True)
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Signs, Nu)]
    Io, Jo, Vo = _out(36 * ne), _out(36 * ne), _out(36 * ne)
    L.check(L.lib.mpk_curl_curl_coo(L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]),
                                    L.ptr(d[4]), C.byref(q), L.ptr(Io), L.ptr(Jo), L.ptr(Vo),
                                    L.stream_ptr()))
    return tuple(t[: 36 * ne].cpu().numpy() for t in (Io, Jo, Vo))


def assemble_mass_kernel_mex(*args):
    _arity(args, 7)
    P, T, Dofs, Signs, Coeffs, Qw, ValRef = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    q = _quad(Qw, ValRef, False)
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Signs)]
    c = _colmajor(Coeffs) if Coeffs is not None else np.zeros(0)
    cd = _dev(c) if c.size else None
    Io, Jo, Vo = _out(36 * ne), _out(36 * ne), _out(36 * ne)
    L.check(L.lib.mpk_mass_coo(L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]),
                               L.ptr(cd), int(c.size), C.byref(q), L.ptr(Io), L.ptr(Jo),
                               L.ptr(Vo), L.stream_ptr()))
    return tuple(t[: 36 * ne].cpu().numpy() for t in (Io, Jo, Vo))


def assemble_jacobian_kernel_mex(*args):
    _arity(args, 17)
    (P, T, Dofs, Signs, Tags, SolA, Qw, CurlRef, lin, isn, maxb, brk, coe, starts, counts,
     coe_starts, calcJ) = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    calc = bool(np.asarray(calcJ).ravel()[0])
    q = _quad(Qw, CurlRef, True)
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Signs, Tags)]
    sol = _colmajor(SolA)
    sd = _dev(sol) if sol.size else None
    tabs = [_dev(_colmajor(x)) if np.asarray(x).size else None
            for x in (lin, isn, maxb, brk, coe, starts, counts, coe_starts)]
    m = L.Material()
    (m.linear_nu, m.is_nonlinear, m.max_bsq, m.breaks, m.coefs, m.spline_start, m.spline_count,
     m.coef_start) = [L.ptr(t) for t in tabs]
    m.num_tags = int(np.asarray(lin).size)
    m.num_breaks = int(np.asarray(brk).size)
    m.num_coefs = int(np.asarray(coe).size)
    n36 = 36 * ne if calc else 0
    Io, Jo, Vo = _out(n36), _out(n36), _out(n36)
    Ri, Rv = _out(6 * ne), _out(6 * ne)
    L.check(L.lib.mpk_jacobian_coo(L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]),
                                   L.ptr(d[4]), L.ptr(sd), int(sol.size), C.byref(q), C.byref(m),
                                   int(calc), L.ptr(Io), L.ptr(Jo), L.ptr(Vo), L.ptr(Ri),
                                   L.ptr(Rv), L.stream_ptr()))
    return (Io[:n36].cpu().numpy(), Jo[:n36].cpu().numpy(), Vo[:n36].cpu().numpy(),
            Ri[: 6 * ne].cpu().numpy(), Rv[: 6 * ne].cpu().numpy())


def assemble_source_kernel_mex(*args):
    """[I, V] = assemble_source_kernel_mex(P, T, Dofs, Signs, Js, Qw, ValRef)
    (src/mex/assemble_source_kernel_mex.cpp:16-120); Js is [3 x Ne]."""
    _arity(args, 7)
    P, T, Dofs, Signs, Js, Qw, ValRef = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    js = np.asarray(Js, dtype=np.float64)
    if js.ndim != 2 or js.shape[0] != 3 or js.shape[1] != ne:
        raise L.MpkError(L.MPK_ERR_ARGS, "Js must be [3 x NumElems]")
    q = _quad(Qw, ValRef, False)
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Signs, js)]
    Io, Vo = _out(6 * ne), _out(6 * ne)
    L.check(L.lib.mpk_source_coo(L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]),
                                 L.ptr(d[4]), C.byref(q), L.ptr(Io), L.ptr(Vo), L.stream_ptr()))
    return tuple(t[: 6 * ne].cpu().numpy() for t in (Io, Vo))


def _grad_table(Qw, GradRef) -> np.ndarray:
    nq = _colmajor(Qw).size
    g = _colmajor(GradRef)
    if g.size < 12 * nq:
        raise L.MpkError(L.MPK_ERR_ARGS, "GradRef smaller than 4 x 3 x numel(Qw)")
    return np.ascontiguousarray(g[: 12 * nq])


def assemble_scalar_laplacian_kernel_mex(*args):
    """[I, J, V] = assemble_scalar_laplacian_kernel_mex(P, T, Dofs, Coeffs, Qw, GradRef)
    (src/mex/assemble_scalar_laplacian_kernel_mex.cpp:13-105); Dofs [4 x Ne]."""
    _arity(args, 6)
    P, T, Dofs, Coeffs, Qw, GradRef = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    q = L.make_quadrature(_colmajor(Qw), None, None)
    g = _grad_table(Qw, GradRef)
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Coeffs)]
    Io, Jo, Vo = _out(16 * ne), _out(16 * ne), _out(16 * ne)
    L.check(L.lib.mpk_scalar_laplacian_coo(
        L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]), C.byref(q),
        g.ctypes.data_as(C.c_void_p), L.ptr(Io), L.ptr(Jo), L.ptr(Vo), L.stream_ptr()))
    return tuple(t[: 16 * ne].cpu().numpy() for t in (Io, Jo, Vo))


def assemble_coupling_kernel_mex(*args):
    """[I, J, V] = assemble_coupling_kernel_mex(P, T, DofsE, Signs, DofsN, Coeffs, Qw,
    NedRef, GradRef) (src/mex/assemble_coupling_kernel_mex.cpp:13-131)."""
    _arity(args, 9)
    P, T, DofsE, Signs, DofsN, Coeffs, Qw, NedRef, GradRef = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    q = _quad(Qw, NedRef, False)
    g = _grad_table(Qw, GradRef)
    d = [_dev(_colmajor(x)) for x in (P, T, DofsE, Signs, DofsN, Coeffs)]
    Io, Jo, Vo = _out(24 * ne), _out(24 * ne), _out(24 * ne)
    L.check(L.lib.mpk_coupling_coo(
        L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]), L.ptr(d[4]), L.ptr(d[5]),
        C.byref(q), g.ctypes.data_as(C.c_void_p), L.ptr(Io), L.ptr(Jo), L.ptr(Vo),
        L.stream_ptr()))
    return tuple(t[: 24 * ne].cpu().numpy() for t in (Io, Jo, Vo))


def assemble_winding_kernel_mex(*args):
    """[I, V] = assemble_winding_kernel_mex(P, T, Dofs, Signs, RegionTags, CoilID,
    DirField, UniformDir, Scale, Qw, RefVal)
    (src/mex/assemble_winding_kernel_mex.cpp:13-126); DirField [3 x Ne] or []."""
    _arity(args, 11)
    P, T, Dofs, Signs, Tags, CoilID, DirField, UniDir, Scale, Qw, RefVal = args
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    q = _quad(Qw, RefVal, False)
    uni = _colmajor(UniDir)
    if uni.size < 3:
        raise L.MpkError(L.MPK_ERR_ARGS, "UniformDir must hold 3 values")
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Signs, Tags)]
    dirf = None
    if np.size(DirField):
        df = np.asarray(DirField, dtype=np.float64)
        if df.ndim != 2 or df.shape[0] != 3 or df.shape[1] != ne:
            raise L.MpkError(L.MPK_ERR_ARGS, "DirField must be [3 x NumElems] or []")
        dirf = _dev(_colmajor(df))
    Io, Vo = _out(6 * ne), _out(6 * ne)
    L.check(L.lib.mpk_winding_coo(
        L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]), L.ptr(d[4]),
        float(np.asarray(CoilID).ravel()[0]), L.ptr(dirf), uni[:3].ctypes.data_as(C.c_void_p),
        float(np.asarray(Scale).ravel()[0]), C.byref(q), L.ptr(Io), L.ptr(Vo), L.stream_ptr()))
    return tuple(t[: 6 * ne].cpu().numpy() for t in (Io, Vo))


def assemble_hbfem_kernel_mex(*args):
    """[I, J, V, R_rows, R_cols, R_val_Real, R_val_Imag] = assemble_hbfem_kernel_mex(...)
    (src/mex/assemble_hbfem_kernel_mex.cpp:25-260).  Imaginary parts may be []."""
    if len(args) < 22:
        raise L.MpkError(L.MPK_ERR_ARGS, "Need 22 inputs (Split Complex).")
    (P, T, Dofs, Signs, Tags, SolR, SolI, DmR, DmI, PmR, PmI, Qw, CurlRef, linNu, isNon,
     maxB, breaks, coefs, starts, counts, reqJ, Scalings) = args[:22]
    ne = _ncols(T)
    nn = _colmajor(P).size // 3
    sol = np.asarray(SolR, dtype=np.float64)
    if sol.ndim == 1:
        sol = sol[:, None]
    nd, H = sol.shape
    nt = _colmajor(DmR).size // H  # Dmat_Real is [Nt x H]
    cnt = _colmajor(counts)
    if len(args) >= 23:
        coe = _colmajor(args[22])
    else:
        # packMaterialData (src/Assembler.m:270-304): coefficient blocks of the
        # nonlinear tags in ascending tag order, 4 per spline piece
        coe = np.concatenate([[0.0], np.cumsum(4.0 * cnt)[:-1]]) if cnt.size else cnt
    q = _quad(Qw, CurlRef, True)
    d = [_dev(_colmajor(x)) for x in (P, T, Dofs, Signs, Tags)]
    opt = lambda x: _dev(_colmajor(x)) if np.size(x) else None
    sr, si = _dev(_colmajor(sol)), opt(SolI)
    dr, di, pr, pi_ = opt(DmR), opt(DmI), opt(PmR), opt(PmI)
    mats = [_dev(_colmajor(x)) for x in (linNu, isNon, maxB, breaks, coefs, starts, cnt, coe)]
    mat = L.Material(linear_nu=L.ptr(mats[0]), is_nonlinear=L.ptr(mats[1]), max_bsq=L.ptr(mats[2]),
                     breaks=L.ptr(mats[3]), coefs=L.ptr(mats[4]), spline_start=L.ptr(mats[5]),
                     spline_count=L.ptr(mats[6]), coef_start=L.ptr(mats[7]),
                     num_tags=int(_colmajor(linNu).size), num_breaks=int(_colmajor(breaks).size),
                     num_coefs=int(_colmajor(coefs).size))
    request_j = bool(np.asarray(reqJ).ravel()[0])
    sc = _dev(_colmajor(Scalings))
    nk = 36 * H * ne if request_j else 0
    Io, Jo, Vo = _out(nk), _out(nk), _out(nk)
    R = [_out(6 * H * ne) for _ in range(4)]
    L.check(L.lib.mpk_hbfem_coo(
        L.ptr(d[0]), nn, L.ptr(d[1]), ne, L.ptr(d[2]), L.ptr(d[3]), L.ptr(d[4]), L.ptr(sr),
        L.ptr(si), nd, H, L.ptr(dr), L.ptr(di), nt, L.ptr(pr), L.ptr(pi_), C.byref(q),
        C.byref(mat), int(request_j), L.ptr(sc), L.ptr(Io), L.ptr(Jo), L.ptr(Vo), L.ptr(R[0]),
        L.ptr(R[1]), L.ptr(R[2]), L.ptr(R[3]), L.stream_ptr()))
    return (Io[:nk].cpu().numpy(), Jo[:nk].cpu().numpy(), Vo[:nk].cpu().numpy(),
            *(t[: 6 * H * ne].cpu().numpy() for t in R))


def sparse_gpu(I, J, V, m, n, drop_zeros: bool = True):
    """Returns (colptr int64 [n+1], rowind int32 0-based [nnz], vals [nnz]) on the host."""
    I, J, V = _colmajor(I), _colmajor(J), _colmajor(V)
    nt = I.size
    dev = torch.device("cuda", torch.cuda.current_device())
    Id, Jd, Vd = (_dev(x) if nt else None for x in (I, J, V))
    colptr = torch.empty(int(n) + 1, dtype=torch.int64, device=dev)
    rowind = torch.empty(max(nt, 1), dtype=torch.int32, device=dev)
    vals = torch.empty(max(nt, 1), dtype=torch.float64, device=dev)
    nnz = C.c_int64(0)
    L.check(L.lib.mpk_sparse_coo_to_csc(L.ptr(Id), L.ptr(Jd), L.ptr(Vd), nt, int(m), int(n),
                                        int(bool(drop_zeros)), L.ptr(colptr), L.ptr(rowind),
                                        L.ptr(vals), nt, C.byref(nnz), L.stream_ptr()))
    k = int(nnz.value)
    return colptr.cpu().numpy(), rowind[:k].cpu().numpy(), vals[:k].cpu().numpy()
