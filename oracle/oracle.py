"""ctypes wrapper of the CPU oracle (oracle/libmpk_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  Functions take
numpy arrays in the reference's memory layouts (MATLAB column-major, which
is C-order ``(N, k)`` for a ``[k x N]`` MATLAB matrix) and return numpy.
See mpk_oracle.c for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmpk_oracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "mpk_oracle.c").stat().st_mtime:
        build()
    return C.CDLL(str(LIB_PATH))


_lib = _load()
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i64 = C.c_int64


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


_lib.orc_generate_edges.restype = _i64
_lib.orc_distribute_dofs.restype = _i64
_lib.orc_sparse.restype = _i64
_lib.orc_sparse_par.restype = _i64
_lib.orc_csc_add_jw.restype = _i64


def set_threads(n: int) -> None:
    os.environ["OMP_NUM_THREADS"] = str(n)


def generate_edges(T):
    """Mesh.generateEdges: T (Ne,4) 1-based -> (Edges (E,2), T2E (Ne,6), Sign (Ne,6))."""
    T, Tp = _i(T)
    ne = T.shape[0]
    edges = np.zeros((6 * ne, 2), dtype=np.int32)
    t2e = np.zeros((ne, 6), dtype=np.int32)
    sign = np.zeros((ne, 6), dtype=np.int8)
    n = _lib.orc_generate_edges(Tp, ne, edges.ctypes.data_as(_ip), t2e.ctypes.data_as(_ip),
                                sign.ctypes.data_as(C.POINTER(C.c_int8)))
    return edges[:n].copy(), t2e, sign


def distribute_dofs(t2e, nedge, active=None, offset=0):
    t2e, tp = _i(t2e)
    ne = t2e.shape[0]
    emap = np.zeros(nedge, dtype=np.int32)
    cell = np.zeros((ne, 6), dtype=np.int32)
    act = None
    if active is not None:
        act = np.ascontiguousarray(active, dtype=np.uint8)
    n = _lib.orc_distribute_dofs(tp, ne, nedge,
                                 act.ctypes.data_as(C.POINTER(C.c_uint8)) if act is not None else None,
                                 offset, emap.ctypes.data_as(_ip), cell.ctypes.data_as(_ip))
    return emap, cell, int(n)


def curl_curl(P, T, Dofs, Signs, Nu, Qw, CurlRef):
    """assemble_curl_curl_kernel_mex: returns (I, J, V) of length 36*Ne."""
    P, Pp = _d(P); T, Tp = _d(T); Dofs, Dp = _d(Dofs); Signs, Sp = _d(Signs)
    Nu, Np = _d(Nu); Qw, Qp = _d(Qw); CurlRef, Cp = _d(CurlRef)
    ne = T.size // 4
    I = np.zeros(36 * ne); J = np.zeros(36 * ne); V = np.zeros(36 * ne)
    _lib.orc_curl_curl_kernel(Pp, Tp, Dp, Sp, Np, Qp, _i64(Qw.size), Cp, _i64(ne),
                              I.ctypes.data_as(_dp), J.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return I, J, V


def mass(P, T, Dofs, Signs, Coeffs, Qw, ValRef):
    """assemble_mass_kernel_mex (Coeffs may be None / [] -> sigma = 1)."""
    P, Pp = _d(P); T, Tp = _d(T); Dofs, Dp = _d(Dofs); Signs, Sp = _d(Signs)
    Qw, Qp = _d(Qw); ValRef, Vp = _d(ValRef)
    ne = T.size // 4
    if Coeffs is None:
        cp, nc = None, 0
    else:
        Coeffs, cp = _d(Coeffs)
        nc = Coeffs.size
    I = np.zeros(36 * ne); J = np.zeros(36 * ne); V = np.zeros(36 * ne)
    _lib.orc_mass_kernel(Pp, Tp, Dp, Sp, cp, _i64(nc), Qp, _i64(Qw.size), Vp, _i64(ne),
                         I.ctypes.data_as(_dp), J.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return I, J, V


def jacobian(P, T, Dofs, Signs, Tags, SolA, Qw, CurlRef, packed, calcJ=True):
    """assemble_jacobian_kernel_mex: returns (I, J, V, R_idx, R_val)."""
    P, Pp = _d(P); T, Tp = _d(T); Dofs, Dp = _d(Dofs); Signs, Sp = _d(Signs)
    Tags, Tgp = _d(Tags); SolA, Ap = _d(SolA); Qw, Qp = _d(Qw); CurlRef, Cp = _d(CurlRef)
    arrs = [_d(packed[k]) for k in ("linNu", "isNon", "maxB", "breaks", "coefs", "starts",
                                    "counts", "coe_starts")]
    ne = T.size // 4
    n36 = 36 * ne if calcJ else 0
    I = np.zeros(max(n36, 1)); J = np.zeros(max(n36, 1)); V = np.zeros(max(n36, 1))
    Ri = np.zeros(6 * ne); Rv = np.zeros(6 * ne)
    _lib.orc_jacobian_kernel(Pp, Tp, Dp, Sp, Tgp, Ap, Qp, _i64(Qw.size), Cp,
                             *[a[1] for a in arrs], C.c_int(1 if calcJ else 0), _i64(ne),
                             I.ctypes.data_as(_dp), J.ctypes.data_as(_dp), V.ctypes.data_as(_dp),
                             Ri.ctypes.data_as(_dp), Rv.ctypes.data_as(_dp))
    return I[:n36], J[:n36], V[:n36], Ri, Rv


def source(P, T, Dofs, Signs, Js, Qw, ValRef):
    """[I, V] = assemble_source_kernel_mex(P, T, Dofs, Signs, Js, Qw, ValRef); Js (Ne, 3)."""
    Pp = _d(P); Tp = _d(T); Dp = _d(Dofs); Sp = _d(Signs); Jp = _d(Js)
    Qp = _d(Qw); Rp = _d(ValRef)
    ne = Tp[0].size // 4
    I = np.zeros(6 * ne); V = np.zeros(6 * ne)
    _lib.orc_source_kernel(Pp[1], Tp[1], Dp[1], Sp[1], Jp[1], Qp[1], _i64(Qp[0].size), Rp[1],
                           _i64(ne), I.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return I, V


def scalar_laplacian(P, T, Dofs, Coeffs, Qw, GradRef):
    """[I, J, V] = assemble_scalar_laplacian_kernel_mex(P, T, Dofs, Coeffs, Qw, GradRef);
    Dofs (Ne, 4) node DOFs, GradRef flat [4 x 3 x Nq]."""
    Pp = _d(P); Tp = _d(T); Dp = _d(Dofs); Cp = _d(Coeffs); Qp = _d(Qw); Gp = _d(GradRef)
    ne = Tp[0].size // 4
    I = np.zeros(16 * ne); J = np.zeros(16 * ne); V = np.zeros(16 * ne)
    _lib.orc_scalar_laplacian_kernel(Pp[1], Tp[1], Dp[1], Cp[1], Qp[1], _i64(Qp[0].size), Gp[1],
                                     _i64(ne), I.ctypes.data_as(_dp), J.ctypes.data_as(_dp),
                                     V.ctypes.data_as(_dp))
    return I, J, V


def coupling(P, T, DofsE, Signs, DofsN, Coeffs, Qw, NedRef, GradRef):
    """[I, J, V] = assemble_coupling_kernel_mex(P, T, DofsE, Signs, DofsN, Coeffs, Qw,
    NedRef, GradRef): 24 triplets per element."""
    Pp = _d(P); Tp = _d(T); Ep = _d(DofsE); Sp = _d(Signs); Np = _d(DofsN); Cp = _d(Coeffs)
    Qp = _d(Qw); Rp = _d(NedRef); Gp = _d(GradRef)
    ne = Tp[0].size // 4
    I = np.zeros(24 * ne); J = np.zeros(24 * ne); V = np.zeros(24 * ne)
    _lib.orc_coupling_kernel(Pp[1], Tp[1], Ep[1], Sp[1], Np[1], Cp[1], Qp[1], _i64(Qp[0].size),
                             Rp[1], Gp[1], _i64(ne), I.ctypes.data_as(_dp),
                             J.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return I, J, V


def winding(P, T, Dofs, Signs, Tags, CoilID, DirField, UniDir, Scale, Qw, RefVal):
    """[I, V] = assemble_winding_kernel_mex(P, T, Dofs, Signs, RegionTags, CoilID,
    DirField (Ne, 3) or None, UniformDir (3,), Scale, Qw, RefVal)."""
    Pp = _d(P); Tp = _d(T); Dp = _d(Dofs); Sp = _d(Signs); Gp = _d(Tags); Up = _d(UniDir)
    Qp = _d(Qw); Rp = _d(RefVal)
    Fp = _d(DirField) if DirField is not None and np.size(DirField) else (None, None)
    ne = Tp[0].size // 4
    I = np.zeros(6 * ne); V = np.zeros(6 * ne)
    _lib.orc_winding_kernel(Pp[1], Tp[1], Dp[1], Sp[1], Gp[1], C.c_double(float(CoilID)),
                            Fp[1], Up[1], C.c_double(float(Scale)), Qp[1], _i64(Qp[0].size),
                            Rp[1], _i64(ne), I.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return I, V


def hbfem(P, T, Dofs, Signs, Tags, SolR, SolI, DmatR, DmatI, PmatR, PmatI, Qw, CurlRef, packed,
          requestJ, Scalings):
    """assemble_hbfem_kernel_mex restated: Sol [numDofs x H], Dmat [Nt x H],
    Pmat [H x Nt] (real and imaginary parts; an imaginary part may be None).
    Returns (I, J, V, R_rows, R_cols, R_val_real, R_val_imag)."""
    SolR = np.asarray(SolR, dtype=np.float64)
    nd, H = SolR.shape
    Nt = np.asarray(DmatR).shape[0]
    f = lambda a: _d(np.asarray(a, dtype=np.float64).ravel(order="F")) if a is not None else (None, None)
    Pp = _d(P); Tp = _d(T); Dp = _d(Dofs); Sp = _d(Signs); Gp = _d(Tags)
    sr, si, dr, di, pr, pi_ = f(SolR), f(SolI), f(DmatR), f(DmatI), f(PmatR), f(PmatI)
    Qp = _d(Qw); Cp = _d(CurlRef); Scp = _d(Scalings)
    arrs = [_d(packed[k]) for k in ("linNu", "isNon", "maxB", "breaks", "coefs", "starts",
                                    "counts", "coe_starts")]
    ne = Tp[0].size // 4
    nk = 36 * H * ne if requestJ else 0
    I = np.zeros(max(nk, 1)); J = np.zeros(max(nk, 1)); V = np.zeros(max(nk, 1))
    Rr, Rc, Rvr, Rvi = (np.zeros(6 * H * ne) for _ in range(4))
    _lib.orc_hbfem_kernel(Pp[1], Tp[1], Dp[1], Sp[1], Gp[1], sr[1], si[1], _i64(nd), _i64(H),
                          dr[1], di[1], _i64(Nt), pr[1], pi_[1], Qp[1], _i64(Qp[0].size), Cp[1],
                          *[a[1] for a in arrs], C.c_int(1 if requestJ else 0), Scp[1], _i64(ne),
                          I.ctypes.data_as(_dp), J.ctypes.data_as(_dp), V.ctypes.data_as(_dp),
                          Rr.ctypes.data_as(_dp), Rc.ctypes.data_as(_dp),
                          Rvr.ctypes.data_as(_dp), Rvi.ctypes.data_as(_dp))
    return I[:nk], J[:nk], V[:nk], Rr, Rc, Rvr, Rvi


def element_b(P, T, Dofs, Signs, sol, CurlRef):
    """PostProcessor.computeElementB for one solution column: B (Ne, 3)."""
    Pp = _d(P); Tp = _d(T); Dp = _d(Dofs); Sp = _d(Signs); sp = _d(sol); Cp = _d(CurlRef)
    ne = Tp[0].size // 4
    B = np.zeros((ne, 3))
    _lib.orc_element_b(Pp[1], Tp[1], Dp[1], Sp[1], sp[1], Cp[1], _i64(ne), B.ctypes.data_as(_dp))
    return B


def eval_nu(s, tags, packed):
    s, sp = _d(s)
    tags, tp = _i(tags)
    arrs = [_d(packed[k]) for k in ("linNu", "isNon", "maxB", "breaks", "coefs", "starts",
                                    "counts", "coe_starts")]
    nu = np.zeros(s.size); dnu = np.zeros(s.size)
    _lib.orc_eval_nu_many(sp, tp, _i64(s.size), *[a[1] for a in arrs],
                          nu.ctypes.data_as(_dp), dnu.ctypes.data_as(_dp))
    return nu, dnu


def sparse(I, J, V, m, n, drop_zeros=True):
    """sparse(I(valid), J(valid), V(valid), m, n): returns (colptr, rowind0, vals)."""
    I, Ip = _d(I); J, Jp = _d(J); V, Vp = _d(V)
    nt = I.size
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowind = np.zeros(max(nt, 1), dtype=np.int64)
    vals = np.zeros(max(nt, 1))
    nnz = _lib.orc_sparse(Ip, Jp, Vp, _i64(nt), _i64(m), _i64(n), C.c_int(1 if drop_zeros else 0),
                          colptr.ctypes.data_as(C.POINTER(C.c_int64)),
                          rowind.ctypes.data_as(C.POINTER(C.c_int64)), vals.ctypes.data_as(_dp))
    if nnz < 0:
        raise IndexError("sparse(): index exceeds matrix dimensions")
    return colptr, rowind[:nnz].copy(), vals[:nnz].copy()


def sparse_par(I, J, V, m, n, drop_zeros=True):
    """sparse() on OpenMP threads (orc_sparse_par): the same (colptr, rowind0,
    vals) as sparse(), bit for bit; used to time the host path."""
    I, Ip = _d(I); J, Jp = _d(J); V, Vp = _d(V)
    nt = I.size
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowind = np.empty(max(nt, 1), dtype=np.int64)
    vals = np.empty(max(nt, 1))
    nnz = _lib.orc_sparse_par(Ip, Jp, Vp, _i64(nt), _i64(m), _i64(n),
                              C.c_int(1 if drop_zeros else 0),
                              colptr.ctypes.data_as(C.POINTER(C.c_int64)),
                              rowind.ctypes.data_as(C.POINTER(C.c_int64)), vals.ctypes.data_as(_dp))
    if nnz == -1:
        raise IndexError("sparse(): index exceeds matrix dimensions")
    if nnz < 0:
        raise MemoryError("sparse(): host allocation failed")
    return colptr, rowind[:nnz], vals[:nnz]


def csc_add_jw(k, mm, w, n):
    """K + 1j*w*M of two CSC triples (colptr, rowind0, vals) with n columns:
    returns (colptr, rowind0, complex vals) over the union of the patterns."""
    kp, kr, kv = (np.ascontiguousarray(k[0], dtype=np.int64), np.ascontiguousarray(k[1], dtype=np.int64),
                  np.ascontiguousarray(k[2], dtype=np.float64))
    mp, mr, mv = (np.ascontiguousarray(mm[0], dtype=np.int64), np.ascontiguousarray(mm[1], dtype=np.int64),
                  np.ascontiguousarray(mm[2], dtype=np.float64))
    cap = max(kr.size + mr.size, 1)
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowind = np.empty(cap, dtype=np.int64)
    re = np.empty(cap)
    im = np.empty(cap)
    p64 = C.POINTER(C.c_int64)
    nnz = _lib.orc_csc_add_jw(_i64(n), kp.ctypes.data_as(p64), kr.ctypes.data_as(p64),
                              kv.ctypes.data_as(_dp), mp.ctypes.data_as(p64), mr.ctypes.data_as(p64),
                              mv.ctypes.data_as(_dp), C.c_double(w), colptr.ctypes.data_as(p64),
                              rowind.ctypes.data_as(p64), re.ctypes.data_as(_dp), im.ctypes.data_as(_dp))
    if nnz < 0:
        raise MemoryError("csc_add_jw: host allocation failed")
    out = np.empty(nnz, dtype=np.complex128)
    out.real = re[:nnz]
    out.imag = im[:nnz]
    return colptr, rowind[:nnz], out


def sparse_scipy(I, J, V, m, n, drop_zeros=True):
    import scipy.sparse as sp
    c, r, v = sparse(I, J, V, m, n, drop_zeros)
    return sp.csc_matrix((v, r, c), shape=(m, n))


def boundary_edges(T, edges):
    T, Tp = _i(T)
    edges, Ep = _i(edges)
    out = np.zeros(edges.shape[0], dtype=np.uint8)
    _lib.orc_boundary_edges(Tp, _i64(T.shape[0]), Ep, _i64(edges.shape[0]),
                            out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.astype(bool)


# ---- composite helpers mirroring Assembler.m on the oracle ----------------

def mesh_inputs(mesh, offset=0, active=None):
    """Reference packing of a mesh: (P, T(double), CellDofs(double), Signs(double), nedge, ndof)."""
    edges, t2e, sign = generate_edges(mesh.T)
    emap, cell, nd = distribute_dofs(t2e, edges.shape[0], active, offset)
    return dict(P=mesh.P, T=mesh.T.astype(np.float64), Dofs=cell.astype(np.float64),
                Signs=sign.astype(np.float64), edges=edges, t2e=t2e, ndof=nd + offset,
                cell=cell, sign=sign)
