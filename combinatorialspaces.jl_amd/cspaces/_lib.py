"""ctypes binding of libdec_hip.so (the C ABI declared in include/dec_hip.h)."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int8, c_int32, c_int64, c_uint64, c_void_p

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# DEC_HIP_LIB names another build of the library (the Julia extension reads
# the same variable)
LIB_PATH = os.environ.get("DEC_HIP_LIB") or os.path.join(_PKG, "libdec_hip.so")

DEC_OK = 0
DEC_F64, DEC_F32 = 0, 1
DEC_CIRCUMCENTER, DEC_BARYCENTER = 0, 1
DEC_GEOMETRIC_HODGE, DEC_DIAGONAL_HODGE = 0, 1
(DEC_OP_WEDGE, DEC_OP_WEDGE_DD, DEC_OP_DIFFERENTIAL, DEC_OP_BOUNDARY, DEC_OP_DUAL_DERIVATIVE,
 DEC_OP_HODGE, DEC_OP_INV_HODGE, DEC_OP_LAPLACIAN0, DEC_OP_AVG01, DEC_OP_WEDGE_PD, DEC_OP_FLAT_SHARP,
 DEC_OP_WEDGE_DP, DEC_OP_INTERIOR_PRODUCT_DD, DEC_OP_LIE_DD, DEC_OP_DUAL_LAPLACIAN,
 DEC_OP_P2_D2_INTERPOLATION, DEC_OP_P3_D3_INTERPOLATION, DEC_OP_D0_P0_INTERPOLATION,
 DEC_OP_P0_D0_INTERPOLATION) = range(1, 20)
DEC_SMOOTH_CG, DEC_SMOOTH_GS, DEC_SMOOTH_JACOBI, DEC_SMOOTH_BLOCK_GS = 0, 1, 2, 3
DEC_MG_DIRECT, DEC_MG_GALERKIN = 0, 1
DEC_SOLVE_PLAIN_F64, DEC_SOLVE_MIXED_IR, DEC_SOLVE_PCG_F64, DEC_SOLVE_PCG_MIXED = 0, 1, 2, 3

(DEC_OK, DEC_ERR_ARG, DEC_ERR_UNSUPPORTED, DEC_ERR_OOM, DEC_ERR_HIP, DEC_ERR_NOCONV,
 DEC_ERR_NONORIENTABLE) = range(7)
STATUS_NAMES = {0: "DEC_OK", 1: "DEC_ERR_ARG", 2: "DEC_ERR_UNSUPPORTED", 3: "DEC_ERR_OOM",
                4: "DEC_ERR_HIP", 5: "DEC_ERR_NOCONV", 6: "DEC_ERR_NONORIENTABLE"}

# name -> (restype, argtypes); every symbol include/dec_hip.h declares.
P = c_void_p
PI32, PI8, PF64, PI64 = POINTER(c_int32), POINTER(c_int8), POINTER(c_double), POINTER(c_int64)
SIGNATURES = {
    "dec_last_error": (c_char_p, []),
    "dec_version": (c_char_p, []),
    "dec_fill_uniform": (c_int, [P, c_int64, c_uint64, c_int, P]),
    "dec_mesh_create": (c_int, [c_int32, c_int32, c_int32, P, P, P, P, P, P, P, P, c_int32, POINTER(P)]),
    "dec_mesh_from_triangles": (c_int, [c_int32, c_int32, P, P, c_int32, c_int32, POINTER(P)]),
    "dec_mesh_icosphere": (c_int, [c_int32, POINTER(P)]),
    "dec_mesh_triangulated_grid": (c_int, [c_double, c_double, c_double, c_double, c_int32, POINTER(P)]),
    "dec_mesh_subdivide": (c_int, [P, c_int32, POINTER(P)]),
    "dec_mesh_counts": (c_int, [P, PI32, PI32, PI32]),
    "dec_mesh_export": (c_int, [P, P, P, P, P, P, P, P, P]),
    "dec_mesh_dualize": (c_int, [P, c_int32]),
    "dec_mesh_export_dual": (c_int, [P, P, P, P, P, P]),
    "dec_mesh_destroy": (c_int, [P]),
    "dec_mesh3_create": (c_int, [c_int32] * 4 + [P] * 13 + [c_int32, POINTER(P)]),
    "dec_mesh3_from_tets": (c_int, [c_int32, c_int32, P, P, c_int32, c_int32, POINTER(P)]),
    "dec_mesh_dim": (c_int, [P, PI32]),
    "dec_mesh3_counts": (c_int, [P, PI32, PI32, PI32, PI32]),
    "dec_mesh3_export": (c_int, [P, P, P, P, P, P]),
    "dec_mesh3_export_dual": (c_int, [P, P, P, P, P, P, P, P]),
    "dec_op_create": (c_int, [P, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(P)]),
    "dec_op_shape": (c_int, [P, PI64, PI64, PI64]),
    "dec_op_apply": (c_int, [P, P, P, P, P]),
    "dec_op_apply_many": (c_int, [c_int32, P, P, P, P, P]),
    "dec_op_solve_info": (c_int, [P, PI32, PF64, P]),
    "dec_op_export_csr": (c_int, [P, PI64, P, P, P]),
    "dec_op_export_tables": (c_int, [P, P, P]),
    "dec_op_destroy": (c_int, [P]),
    "dec_mg_create": (c_int, [P, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(P)]),
    "dec_mg_create_mode": (c_int, [P, c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(P)]),
    "dec_mg_create_from_csr": (c_int, [c_int32, P] + [P] * 9 + [c_int32, c_int32, POINTER(P)]),
    "dec_mg_nlevels": (c_int, [P, PI32]),
    "dec_mg_level_size": (c_int, [P, c_int32, PI32]),
    "dec_mg_finest_mesh": (c_int, [P, POINTER(P)]),
    "dec_mg_export_level": (c_int, [P, c_int32, c_int32, PI64, P, P, P]),
    "dec_mg_export_colors": (c_int, [P, c_int32, P, PI32]),
    "dec_mg_export_blocks": (c_int, [P, c_int32, P, PI32]),
    "dec_mg_set_smoother": (c_int, [P, c_int32, c_int32, c_double, c_int32]),
    "dec_mg_cycles": (c_int, [P, c_int32, c_int32, P, P, P]),
    "dec_mg_full_multigrid": (c_int, [P, c_int32, c_int32, P, P, P]),
    "dec_mg_solve": (c_int, [P, c_int32, c_double, c_int32, P, P, PI32, PF64, P]),
    "dec_mg_destroy": (c_int, [P]),
    "dec_nn_create": (c_int, [P, c_uint64, POINTER(P)]),
    "dec_nn_nparams": (c_int, [P, PI32]),
    "dec_nn_get_params": (c_int, [P, P]),
    "dec_nn_set_params": (c_int, [P, P]),
    "dec_nn_forward": (c_int, [P, c_int32, c_int32, P, P, c_int32, P]),
    "dec_nn_smooth": (c_int, [P, c_int32, P, P, P]),
    "dec_nn_loss_grad": (c_int, [P, c_int32, c_int32, P, P, c_int32, PF64, P, P]),
    "dec_nn_train": (c_int, [P, c_int32, c_int32, P, P, c_int32, c_int32, ctypes.c_float, c_uint64, P, P]),
    "dec_nn_destroy": (c_int, [P]),
    "dec_mg_set_learned": (c_int, [P, P, c_int32, c_int32, c_int32, c_int32]),
    "dec_mg_learned_stats": (c_int, [P, PI32, PI32]),
    "dec_batch_create": (c_int, [c_int32, P, c_int32, POINTER(P)]),
    "dec_batch_layout": (c_int, [P, PI64, P, P]),
    "dec_batch_tune": (c_int, [P, P, c_char_p, PI32, PI32, PI32]),
    "dec_batch_heat_steps": (c_int, [P, P, c_int32, c_int32, P]),
    "dec_batch_heat_steps_noalloc": (c_int, [P, P, c_int32, P]),
    "dec_batch_destroy": (c_int, [P]),
}


class DecError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libdec_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != DEC_OK:
        msg = lib().dec_last_error()
        raise DecError(status, msg.decode() if msg else "")
    return status


def ptr(a):
    """Host numpy array -> void*."""
    return None if a is None else a.ctypes.data_as(c_void_p)
