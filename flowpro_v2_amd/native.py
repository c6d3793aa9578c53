"""ctypes binding of libflowpro_hip.so (include/flowpro_hip.h).

The library must be built in-tree (``python flowpro_v2_amd/build.py``); there is
no fallback: importing the binding without the library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflowpro_hip.so")

FP_OK, FP_NOT_CONVERGED, FP_SINGULAR_BLOCK, FP_NONFINITE, FP_UNSUPPORTED = 0, 1, 2, 3, 4
FP_EINVAL, FP_EHIP, FP_ENOMEM = -1, -2, -3
RECOVERABLE = {FP_NOT_CONVERGED: "GMRES did not converge", FP_SINGULAR_BLOCK: "singular diagonal block",
               FP_NONFINITE: "non-finite solution"}

# every symbol declared in include/flowpro_hip.h
EXPORTS = ["fp_neighbor_matrix", "fp_rcm_order", "fp_create", "fp_destroy", "fp_last_error", "fp_info",
           "fp_element_order", "fp_set_state", "fp_get_state", "fp_next_level", "fp_restore_level", "fp_time_step",
           "fp_assemble", "fp_linear_solve", "fp_update", "fp_export_residual", "fp_export_jacobian",
           "fp_export_solution", "fp_jv", "fp_jtv", "fp_time_kernel", "fp_stream", "fp_linear_solve_jfnk",
           "fp_functional", "fp_adjoint_solve", "fp_wall_distance", "fp_neighbor_matrix_tet", "fp_rcm_order_tet",
           "fp_limiter", "fp_export_artificial_viscosity", "fp_export_averages", "fp_neighbor_matrix_quad",
           "fp_rcm_order_quad", "fp_neighbor_matrix_mixed", "fp_autotune_report", "fp_time_step_noalloc"]


class FpCase(C.Structure):
    _fields_ = [
        ("n_points", C.c_int), ("pxy", C.POINTER(C.c_double)),
        ("n_elems", C.c_int), ("tp", C.POINTER(C.c_int)),
        ("n_bnd", C.c_int), ("bnd", C.POINTER(C.c_int)),
        ("wall_distance", C.POINTER(C.c_double)),
        ("order", C.c_int),
        ("n_vol", C.c_int), ("vol_xi", C.POINTER(C.c_double)), ("vol_w", C.POINTER(C.c_double)),
        ("n_face", C.c_int), ("face_xi", C.POINTER(C.c_double)), ("face_w", C.POINTER(C.c_double)),
        ("model", C.c_int), ("viscous", C.c_int),
        ("kappa", C.c_double), ("p_out", C.c_double), ("reynolds", C.c_double), ("prandtl", C.c_double),
        ("has_attack", C.c_int), ("attack", C.c_double),
        ("scalar_a0", C.c_double), ("scalar_a1", C.c_double), ("scalar_k", C.c_double),
        ("beta0", C.c_double), ("penalty", C.c_double),
        ("ordering", C.c_int), ("preconditioner", C.c_int), ("precond_bits", C.c_int),
        ("gmres_restart", C.c_int), ("gmres_cycles", C.c_int), ("gmres_tol", C.c_double),
        ("positivity_limiter", C.c_int),
        ("sa_vt_in", C.c_double), ("sa_sigma", C.c_double), ("sa_cb1", C.c_double), ("sa_cb2", C.c_double),
        ("sa_ka", C.c_double), ("sa_cw2", C.c_double), ("sa_cw3", C.c_double), ("sa_cv1", C.c_double),
        ("sa_prt", C.c_double),
        ("dimension", C.c_int),
        ("damp_tol", C.c_double), ("damp_const", C.c_double), ("damp_inner_tol", C.c_double),
        ("n_damp_inner_coef", C.c_int), ("damp_inner_coef", C.c_double * 5),
        ("numerical_flux", C.c_int),
        ("elem_type", C.c_int),
        ("elem_types", C.POINTER(C.c_int)), ("elem_orders", C.POINTER(C.c_int)),
        ("gauss_rules_dir", C.c_char_p),
        ("cv", C.c_double), ("p_ref", C.c_double), ("rho_ref", C.c_double),
    ] + [(f"kw_{n}", C.c_double) for n in ("k_in", "om_in", "suth", "a1", "sk1", "som1", "alpha1", "beta1", "betast",
                                            "sk2", "som2", "alpha2", "beta2", "prt")]


_lib = None


def lib() -> C.CDLL:
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python flowpro_v2_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, lp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_longlong)
    sig = {
        "fp_neighbor_matrix": (C.c_int, [C.c_int, ip, C.c_int, ip, ip]),
        "fp_rcm_order": (C.c_int, [C.c_int, ip, ip]),
        "fp_create": (C.c_int, [C.POINTER(FpCase), C.POINTER(vp)]),
        "fp_destroy": (C.c_int, [vp]),
        "fp_last_error": (C.c_char_p, [vp]),
        "fp_info": (C.c_int, [vp, lp]),
        "fp_element_order": (C.c_int, [vp, ip]),
        "fp_set_state": (C.c_int, [vp, dp, dp, dp]),
        "fp_get_state": (C.c_int, [vp, dp]),
        "fp_next_level": (C.c_int, [vp]),
        "fp_restore_level": (C.c_int, [vp]),
        "fp_time_step": (C.c_int, [vp, dp]), "fp_time_step_noalloc": (C.c_int, [vp, dp]),
        "fp_assemble": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp]),
        "fp_linear_solve": (C.c_int, [vp, ip, dp]),
        "fp_update": (C.c_int, [vp, ip]),
        "fp_export_residual": (C.c_int, [vp, dp, dp]),
        "fp_export_jacobian": (C.c_int, [vp, dp, dp]),
        "fp_export_solution": (C.c_int, [vp, dp]),
        "fp_jv": (C.c_int, [vp, C.c_double, dp, dp]),
        "fp_jtv": (C.c_int, [vp, C.c_double, dp, dp]),
        "fp_time_kernel": (C.c_int, [vp, C.c_int, C.c_int, dp]),
        "fp_stream": (vp, [vp]),
        "fp_linear_solve_jfnk": (C.c_int, [vp, ip, dp]),
        "fp_functional": (C.c_int, [vp, C.c_int, dp, dp]),
        "fp_adjoint_solve": (C.c_int, [vp, C.c_double, dp, dp, ip, dp]),
        "fp_wall_distance": (C.c_int, [C.c_int, dp, C.c_int, ip, ip, dp]),
        "fp_neighbor_matrix_tet": (C.c_int, [C.c_int, ip, C.c_int, ip, ip]),
        "fp_rcm_order_tet": (C.c_int, [C.c_int, ip, ip]),
        "fp_neighbor_matrix_quad": (C.c_int, [C.c_int, ip, C.c_int, ip, ip]),
        "fp_rcm_order_quad": (C.c_int, [C.c_int, ip, ip]),
        "fp_neighbor_matrix_mixed": (C.c_int, [C.c_int, ip, ip, C.c_int, ip, ip]),
        "fp_autotune_report": (C.c_int, [vp, ip, ip, dp, C.c_int]),
        "fp_limiter": (C.c_int, [vp, C.c_int]),
        "fp_export_artificial_viscosity": (C.c_int, [vp, dp, dp]),
        "fp_export_averages": (C.c_int, [vp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class FatalError(RuntimeError):
    """A fatal (< 0) return code: bad arguments or a HIP error."""


def neighbor_matrix(TP: np.ndarray, bnd: np.ndarray) -> np.ndarray:
    TP = np.ascontiguousarray(TP, dtype=np.int32)
    bnd = np.ascontiguousarray(bnd, dtype=np.int32)
    TT = np.zeros_like(TP)
    rc = lib().fp_neighbor_matrix(TP.shape[0], iptr(TP), bnd.shape[0], iptr(bnd), iptr(TT))
    if rc:
        raise FatalError(f"fp_neighbor_matrix: {rc}")
    return TT


def neighbor_matrix_tet(TP: np.ndarray, bnd: np.ndarray) -> np.ndarray:
    """createNeighborMatrix on tetrahedra (fp_neighbor_matrix_tet): TT[n_elems][4]."""
    TP = np.ascontiguousarray(TP, dtype=np.int32)
    bnd = np.ascontiguousarray(bnd, dtype=np.int32).reshape(-1, 4)
    TT = np.zeros_like(TP)
    rc = lib().fp_neighbor_matrix_tet(TP.shape[0], iptr(TP), bnd.shape[0], iptr(bnd), iptr(TT))
    if rc:
        raise FatalError(f"fp_neighbor_matrix_tet: {rc}")
    return TT


def rcm_order_tet(TT: np.ndarray) -> np.ndarray:
    TT = np.ascontiguousarray(TT, dtype=np.int32)
    order = np.zeros(TT.shape[0], np.int32)
    rc = lib().fp_rcm_order_tet(TT.shape[0], iptr(TT), iptr(order))
    if rc:
        raise FatalError(f"fp_rcm_order_tet: {rc}")
    return order


def neighbor_matrix_quad(TP: np.ndarray, bnd: np.ndarray) -> np.ndarray:
    """createNeighborMatrix on quadrilaterals (fp_neighbor_matrix_quad): TT[n_elems][4]."""
    TP = np.ascontiguousarray(TP, dtype=np.int32)
    bnd = np.ascontiguousarray(bnd, dtype=np.int32).reshape(-1, 3)
    TT = np.zeros_like(TP)
    rc = lib().fp_neighbor_matrix_quad(TP.shape[0], iptr(TP), bnd.shape[0], iptr(bnd), iptr(TT))
    if rc:
        raise FatalError(f"fp_neighbor_matrix_quad: {rc}")
    return TT


def rcm_order_quad(TT: np.ndarray) -> np.ndarray:
    TT = np.ascontiguousarray(TT, dtype=np.int32)
    order = np.zeros(TT.shape[0], np.int32)
    rc = lib().fp_rcm_order_quad(TT.shape[0], iptr(TT), iptr(order))
    if rc:
        raise FatalError(f"fp_rcm_order_quad: {rc}")
    return order


def wall_distance(PXY: np.ndarray, TP: np.ndarray, TT: np.ndarray) -> np.ndarray:
    """Vertex distance to the wall faces along mesh edges (fp_wall_distance)."""
    PXY = np.ascontiguousarray(PXY, dtype=np.float64)
    TP = np.ascontiguousarray(TP, dtype=np.int32)
    TT = np.ascontiguousarray(TT, dtype=np.int32)
    out = np.zeros(PXY.shape[0])
    rc = lib().fp_wall_distance(PXY.shape[0], dptr(PXY), TP.shape[0], iptr(TP), iptr(TT), dptr(out))
    if rc:
        raise FatalError(f"fp_wall_distance: {rc}")
    return out


def rcm_order(TT: np.ndarray) -> np.ndarray:
    TT = np.ascontiguousarray(TT, dtype=np.int32)
    order = np.zeros(TT.shape[0], np.int32)
    rc = lib().fp_rcm_order(TT.shape[0], iptr(TT), iptr(order))
    if rc:
        raise FatalError(f"fp_rcm_order: {rc}")
    return order


def neighbor_matrix_mixed(TP: np.ndarray, types: np.ndarray, bnd: np.ndarray) -> np.ndarray:
    """createNeighborMatrix on a mesh of triangles and squares (fp_neighbor_matrix_mixed):
    TP[n_elems][4] (the fourth vertex unused on a triangle), TT[n_elems][4]."""
    TP = np.ascontiguousarray(TP, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    bnd = np.ascontiguousarray(bnd, dtype=np.int32)
    TT = np.zeros((TP.shape[0], 4), np.int32)
    rc = lib().fp_neighbor_matrix_mixed(TP.shape[0], iptr(TP), iptr(types), bnd.shape[0], iptr(bnd), iptr(TT))
    if rc:
        raise FatalError(f"fp_neighbor_matrix_mixed: {rc}")
    return TT
