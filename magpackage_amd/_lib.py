"""ctypes binding of libmpk_assembly.so (the C ABI in include/mpk_assembly.h and
include/mpk_mlp.h).

The library is required: importing this module raises ``ImportError`` when
the built ``libmpk_assembly.so`` is missing, and there is no CPU fallback.
Device arrays are passed as ``torch`` tensors (their ``data_ptr``), host
tables as ``ctypes`` structs.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libmpk_assembly.so"
if not _LIB_PATH.exists():
    raise ImportError(
        f"{_LIB_PATH} is missing: build it with `python -m magpackage_amd.build` "
        "(the engine has no CPU fallback)")
lib = C.CDLL(str(_LIB_PATH))

MPK_OK = 0
MPK_ERR_ARGS = 1
MPK_ERR_INDEX = 2
MPK_ERR_HIP = 3
MPK_ERR_LIMIT = 4

PREC_FP64_EXACT = 0
PREC_FP64 = 1
PREC_FP32_LOCAL = 2
PRECISIONS = {"fp64_exact": PREC_FP64_EXACT, "fp64": PREC_FP64, "fp32_local": PREC_FP32_LOCAL}


class MpkError(RuntimeError):
    """Error raised by the engine; ``code`` is the mpk status, ``ident`` mirrors
    the reference's MEX error identifier for argument errors."""

    def __init__(self, code: int, msg: str):
        self.code = code
        self.ident = "MagPackage:Args" if code == MPK_ERR_ARGS else "MagPackage:Engine"
        super().__init__(f"[{self.ident}] (mpk status {code}) {msg}")


class Quadrature(C.Structure):
    _fields_ = [("nq", C.c_int32), ("qw", C.c_double * 8), ("q_pts", C.c_double * 24),
                ("ref_val", C.c_double * 144), ("ref_curl", C.c_double * 144)]


class Material(C.Structure):
    _fields_ = [("linear_nu", C.c_void_p), ("is_nonlinear", C.c_void_p), ("max_bsq", C.c_void_p),
                ("breaks", C.c_void_p), ("coefs", C.c_void_p), ("spline_start", C.c_void_p),
                ("spline_count", C.c_void_p), ("coef_start", C.c_void_p),
                ("num_tags", C.c_int64), ("num_breaks", C.c_int64), ("num_coefs", C.c_int64)]


class BatchItem(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("nu", C.c_void_p), ("sigma", C.c_void_p),
                ("vals_c128", C.c_void_p)]


class AuxSpace(C.Structure):
    _fields_ = [("num_nodes", C.c_int64), ("edge_nodes", C.c_void_p), ("node_ptr", C.c_void_p),
                ("node_refs", C.c_void_p), ("ln_ptr", C.c_void_p), ("ln_idx", C.c_void_p),
                ("ln_vals", C.c_void_p), ("sweeps", C.c_int32), ("omega", C.c_double)]


_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double
_pq = C.POINTER(Quadrature)
_pm = C.POINTER(Material)
_pi64 = C.POINTER(C.c_int64)
_plan_p = C.c_void_p

_SIGS = {
    "mpk_abi_version": ([], _i32),
    "mpk_last_error": ([], C.c_char_p),
    "mpk_device_count": ([], _i32),
    "mpk_curl_curl_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, _pq, _vp, _vp, _vp, _vp], _i32),
    "mpk_mass_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _pq, _vp, _vp, _vp, _vp], _i32),
    "mpk_jacobian_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _pq, _pm, _i32,
                          _vp, _vp, _vp, _vp, _vp, _vp], _i32),
    "mpk_source_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, _pq, _vp, _vp, _vp], _i32),
    "mpk_scalar_laplacian_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _pq, _vp, _vp, _vp, _vp,
                                  _vp], _i32),
    "mpk_coupling_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _pq, _vp, _vp, _vp, _vp,
                          _vp], _i32),
    "mpk_hbfem_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _vp, _vp,
                       _i32, _vp, _vp, _pq, _pm, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                       _vp], _i32),
    "mpk_winding_coo": ([_vp, _i64, _vp, _i64, _vp, _vp, _vp, C.c_double, _vp, _vp, C.c_double,
                         _pq, _vp, _vp, _vp], _i32),
    "mpk_element_b": ([_vp, _i64, _vp, _vp, _i64, _vp, _i64, _pq, _vp, _vp], _i32),
    "mpk_csc_source": ([_plan_p, _vp, _vp, _pq, _vp, _vp], _i32),
    "mpk_eval_nu": ([_vp, _vp, _i64, _pm, _vp, _vp, _vp], _i32),
    "mpk_sparse_coo_to_csc": ([_vp, _vp, _vp, _i64, _i64, _i64, _i32, _vp, _vp, _vp, _i64,
                               _pi64, _vp], _i32),
    "mpk_generate_edges": ([_vp, _i64, _i64, _vp, _vp, _vp, _pi64, _vp], _i32),
    "mpk_distribute_dofs": ([_vp, _i64, _i64, _vp, _i64, _vp, _vp, _pi64, _vp], _i32),
    "mpk_plan_create": ([_vp, _vp, _i64, _i64, _i64, _vp, C.POINTER(_plan_p)], _i32),
    "mpk_plan_destroy": ([_plan_p], _i32),
    "mpk_plan_info": ([_plan_p, _pi64, _pi64, _pi64, _pi64, _pi64], _i32),
    "mpk_plan_pattern": ([_plan_p, C.POINTER(_vp), C.POINTER(_vp)], _i32),
    "mpk_plan_ring_order": ([_plan_p, C.POINTER(C.c_int)], _i32),
    "mpk_plan_set_coords": ([_plan_p, _vp, _vp], _i32),
    "mpk_csc_curl_curl": ([_plan_p, _vp, _vp, _pq, _i32, _vp, _vp], _i32),
    "mpk_csc_mass": ([_plan_p, _vp, _vp, _pq, _i32, _vp, _vp], _i32),
    "mpk_csc_curl_curl_mass_complex": ([_plan_p, _vp, _vp, _vp, _d, _pq, _pq, _i32, _vp, _vp],
                                       _i32),
    "mpk_csc_jacobian": ([_plan_p, _vp, _vp, _vp, _pm, _pq, _i32, _i32, _vp, _vp, _vp], _i32),
    "mpk_csc_newton_set": ([_plan_p, _vp, _vp, _vp, _pm, _vp, _pq, _pq, _i32, _vp, _vp, _vp,
                            _vp, _vp], _i32),
    "mpk_csc_newton_set_noalloc": ([_plan_p, _vp, _vp, _vp, _pm, _vp, _pq, _pq, _i32, _vp, _vp,
                                    _vp, _vp, _vp], _i32),
    "mpk_csc_drop_zeros": ([_plan_p, _vp, _i32, _vp, _vp, _vp, _pi64, _vp], _i32),
    "mpk_spmv_sym": ([_vp, _vp, _vp, _i64, _vp, _vp, _vp], _i32),
    "mpk_csc_combine": ([_i64, _d, _vp, _d, _vp, _d, _vp, _d, _vp, _vp, _vp], _i32),
    "mpk_pcg_work_size": ([_i64, _pi64], _i32),
    "mpk_pcg_aux_work_size": ([_i64, _i64, _pi64], _i32),
    "mpk_batch_create": ([C.POINTER(BatchItem), _i64, _vp, _i32, _pq, _pq, _i32, _vp,
                          C.POINTER(_vp)], _i32),
    "mpk_batch_schedule": ([_vp, _i32, _i32], _i32),
    "mpk_batch_info": ([_vp, _pi64, _pi64, _pi64], _i32),
    "mpk_batch_run": ([_vp, _vp], _i32),
    "mpk_batch_destroy": ([_vp], _i32),
    "mpk_pcg_aux_solve": ([_vp, _vp, _vp, _i64, _vp, C.POINTER(AuxSpace), _vp, _vp, _vp, _d, _i32,
                           _i32, C.POINTER(C.c_int), C.POINTER(C.c_double), _vp], _i32),
    "mpk_cocg_aux_work_size": ([_i64, _i64, _pi64], _i32),
    "mpk_cocg_aux_solve": ([_vp, _vp, _vp, _vp, _i64, _vp, C.POINTER(AuxSpace), _vp, _vp, _vp,
                            _d, _i32, _i32, C.POINTER(C.c_int), C.POINTER(C.c_double), _vp],
                           _i32),
    "mpk_mlp_sync_bf16": ([_vp, _vp, _vp], _i32),
    "mpk_mlp_forward": ([_vp, _vp, _vp, _i64, _vp, _vp], _i32),
    "mpk_mlp_work_size": ([_i64, _pi64], _i32),
    "mpk_mlp_loss_grad": ([_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp], _i32),
    "mpk_mlp_adam": ([_vp, _vp, _vp, _vp, _i64, C.c_float, C.c_float, C.c_float, C.c_float,
                      _vp, _vp], _i32),
    "mpk_pcg_solve": ([_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _d, _i32, _i32,
                       C.POINTER(C.c_int), C.POINTER(C.c_double), _vp], _i32),
}
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_SIGS)


def check(rc: int) -> None:
    if rc != MPK_OK:
        msg = lib.mpk_last_error()
        raise MpkError(rc, msg.decode() if msg else "")


class KrylovError(RuntimeError):
    """A device Krylov solve (mpk_pcg_solve, mpk_pcg_aux_solve,
    mpk_cocg_aux_solve) stopped at its iteration limit or on a non-finite
    residual.  The C ABI reports both through its iteration count and
    relative residual, not a status code; the reference's direct solves
    (LinearSolver.m:23-24 checks NaN/Inf) have no such outcome, so a caller
    standing in for one raises this instead of returning the unconverged x."""

    def __init__(self, what: str, iterations: int, rel: float, tol: float):
        self.iterations, self.rel_residual, self.tolerance = iterations, rel, tol
        super().__init__(f"[MagPackage:Solver] {what} did not converge: relative residual "
                         f"{rel:.3g} after {iterations} iterations (tolerance {tol:.3g})")


def check_krylov(iters, rel, tol: float, what: str = "Krylov solve",
                 finite_only: bool = False) -> tuple[int, float]:
    """Raises KrylovError when the solve's relative residual is not finite or,
    unless ``finite_only`` (inexact Newton steps, whose outer residual judges
    them), above ``tol``.  ``iters`` / ``rel`` are ctypes values or numbers.
    Returns (iterations, rel)."""
    import math
    it = int(getattr(iters, "value", iters))
    r = float(getattr(rel, "value", rel))
    if not math.isfinite(r) or (not finite_only and r > tol):
        raise KrylovError(what, it, r, tol)
    return it, r


def ptr(t) -> int | None:
    """Device/host address of a torch tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None) -> int | None:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def make_quadrature(q_w: np.ndarray, ref_val_flat=None, ref_curl_flat=None, q_pts=None) -> Quadrature:
    """Packs the reference tables (MATLAB column-major) into an mpk_quadrature."""
    q = Quadrature()
    nq = int(np.asarray(q_w).size)
    if not 1 <= nq <= 8:
        raise MpkError(MPK_ERR_ARGS, f"nq={nq} outside [1, 8]")
    q.nq = nq
    for i, w in enumerate(np.asarray(q_w, dtype=np.float64).ravel()):
        q.qw[i] = float(w)
    if q_pts is not None:
        pts = np.asarray(q_pts, dtype=np.float64).ravel(order="F")
        for i, v in enumerate(pts):
            q.q_pts[i] = float(v)
    if ref_val_flat is not None:
        for i, v in enumerate(np.asarray(ref_val_flat, dtype=np.float64).ravel()):
            q.ref_val[i] = float(v)
    if ref_curl_flat is not None:
        for i, v in enumerate(np.asarray(ref_curl_flat, dtype=np.float64).ravel()):
            q.ref_curl[i] = float(v)
    return q


def standard_quadrature(order: int) -> Quadrature:
    """Tables for ``get_quadrature_data('tet', order)`` + ``nedelec_tet_p1``."""
    from .fem_tables import get_quadrature_data, matlab_flat, nedelec_tet_p1
    q_pts, q_w = get_quadrature_data("tet", order)
    val, curl = nedelec_tet_p1(q_pts)
    return make_quadrature(q_w, matlab_flat(val), matlab_flat(curl), q_pts)
