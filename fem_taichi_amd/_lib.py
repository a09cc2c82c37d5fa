"""ctypes binding of ``libfem355.so`` (the C ABI declared in include/fem355.h).

The library is built in-tree (``__graft_entry__.build()``) and loaded from
this package directory.  There is no fallback: when the shared library is
missing or no GPU is visible, every compute entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfem355.so")

FEM_OK = 0
FEM_E_INVALID = -1
FEM_E_HIP = -2
FEM_E_LIMIT = -3
FEM_E_NUMERIC = -4

DTYPE_F64 = 0
DTYPE_F32 = 1
DTYPE_F32F64 = 2
LAYOUT_BSR = 0
LAYOUT_CSR = 1

_ERRNAMES = {FEM_E_INVALID: "FEM_E_INVALID", FEM_E_HIP: "FEM_E_HIP", FEM_E_LIMIT: "FEM_E_LIMIT",
             FEM_E_NUMERIC: "FEM_E_NUMERIC"}


class FemError(RuntimeError):
    """A failed fem355 call: ``code`` is the FEM_E_* status."""

    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class FemMesh(ctypes.Structure):
    """Mirror of ``struct fem_mesh`` (include/fem355.h)."""

    _fields_ = [
        ("num_node", ctypes.c_int64), ("num_elem", ctypes.c_int64), ("nnzb", ctypes.c_int64),
        ("coords", ctypes.c_void_p), ("elem", ctypes.c_void_p), ("sm_node", ctypes.c_void_p),
        ("re_node", ctypes.c_void_p), ("co_node", ctypes.c_void_p), ("co_elem", ctypes.c_void_p),
        ("ne_ptr", ctypes.c_void_p), ("ne_list", ctypes.c_void_p),
        ("num_groups", ctypes.c_int64), ("max_group_elems", ctypes.c_int32),
        ("max_group_rows", ctypes.c_int32), ("max_group_slots", ctypes.c_int32),
        ("max_group_tasks", ctypes.c_int32),
        ("grp_row_ptr", ctypes.c_void_p), ("grp_rows", ctypes.c_void_p),
        ("grp_elem_ptr", ctypes.c_void_p), ("grp_elems", ctypes.c_void_p),
        ("grp_enodes", ctypes.c_void_p),
        ("grp_task_ptr", ctypes.c_void_p), ("tasks", ctypes.c_void_p),
        ("grp_row_rec", ctypes.c_void_p), ("grp_inc", ctypes.c_void_p),
    ]


class GridTransfer(ctypes.Structure):
    """Mirror of ``struct fem_grid_transfer`` (include/fem355.h)."""

    _fields_ = [("nf", ctypes.c_int32 * 3), ("nc", ctypes.c_int32 * 3), ("left", ctypes.c_void_p),
                ("wright", ctypes.c_void_p), ("span", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int
_D = ctypes.c_double
_MP = ctypes.POINTER(FemMesh)
_TP = ctypes.POINTER(GridTransfer)

# name -> (restype, argtypes); every name here is declared in include/fem355.h
SIGNATURES = {
    "fem_last_error": (ctypes.c_char_p, []),
    "fem_abi_version": (_I32, []),
    "fem_device_count": (_I32, [ctypes.POINTER(ctypes.c_int)]),
    "fem_pattern_work_bytes": (_I64, [_I64, _I64]),
    "fem_pattern_count": (_I32, [_P, _I64, _I64, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_pattern_fill": (_I32, [_P, _I64, _P, _P, _P, _P, _P]),
    "fem_csr_index_K": (_I32, [_MP, _P, _P, _P]),
    "fem_csr_index_M": (_I32, [_MP, _P, _P, _P]),
    "fem_csr_values_M": (_I32, [_MP, _P, _P, _P]),
    "fem_hex8_assemble": (_I32, [_MP, _P, _D, _D, _D, _P, _I32, _I32, _P, _P]),
    "fem_hex8_assemble_lds_bytes": (_I64, [ctypes.c_int32] * 2 + [_I32, _I32, _I32]),
    "fem_hex8_assemble_kn": (_I32, [_MP, _P, _D, _D, _P, _P, _P, _I32, _P]),
    "fem_hex8_kn_lds_bytes": (_I64, [ctypes.c_int32, ctypes.c_int32]),
    "fem_hex8_assemble_ks": (_I32, [_MP, _P, _D, _D, _P, _P, _I32, _P]),
    "fem_spin_softening": (_I32, [_MP, _P, ctypes.POINTER(ctypes.c_double), _P, _I32, _P]),
    "fem_apply_penalty": (_I32, [_MP, _P, _P, _I64, _P, _I32, _I32, _D, _P]),
    "fem_body_force_work_bytes": (_I64, [_I64]),
    "fem_body_force": (_I32, [_MP, _D, ctypes.POINTER(ctypes.c_double), _P, _P, _P, _I64, _P]),
    "fem_stress_work_bytes": (_I64, [_I64]),
    "fem_stress_recover": (_I32, [_MP, _P, _D, _D, _P, _P, _P, _I64, _P]),
    "fem_bsr_spmv": (_I32, [_MP, _P, _I32, _P, _P, _P]),
    "fem_bsr_spmm": (_I32, [_MP, _P, _P, _P, _I64, _P]),
    "fem_mass_spmm": (_I32, [_MP, _P, _P, _P, _I64, _P]),
    "fem_bsr_spmm16_grid": (_I32, [_MP, _I64, _I64, _I64, _P, _P, _P, _P]),
    "fem_spmm16_grid_partials": (_I64, [_I64, _I64, _I64]),
    "fem_bsr_spmm16_grid_pq": (_I32, [_MP, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _P]),
    "fem_mass_spmm16_grid": (_I32, [_MP, _I64, _I64, _I64, _P, _P, _P, _P]),
    "fem_mg_smooth16_grid": (_I32, [_MP, _I64, _I64, _I64, _P, _P, ctypes.c_float, _P, _P, _P, _I32, _P]),
    "fem_bj_setup": (_I32, [_MP, _P, _I32, _P, _P, _P]),
    "fem_pcg_work_bytes": (_I64, [_I64]),
    "fem_pcg_start": (_I32, [_MP, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_pcg_iterate": (_I32, [_MP, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_mpcg_work_bytes": (_I64, [_I64]),
    "fem_mpcg_start": (_I32, [_MP, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_mpcg_iterate": (_I32, [_MP, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_meshconv_fwd_bf16": (_I32, [_P, _I32, _P, _P, _I32, _I32, _I32, _I32, _P, _P, _P, _I32, _P]),
    "fem_meshconv_flip_bf16": (_I32, [_P, _I32, _I32, _P, _I32, _I32, _P]),
    "fem_meshconv_wgrad_work_bytes": (_I64, [_I32, _I32, _I32, _I32, _I32, _I32]),
    "fem_meshconv_wgrad_bf16": (_I32, [_P, _I32, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _I64, _P]),
    "fem_colsum_work_bytes": (_I64, [_I64, _I32, _I32]),
    "fem_colsum": (_I32, [_P, _I32, _I64, _I32, _I32, _P, _P, _I64, _P]),
    "fem_relu_bwd_bf16": (_I32, [_P, _P, _I64, _P, _P]),
    "fem_to_bf16": (_I32, [_P, _I64, _I32, _I32, _P, _P]),
    "fem_meshconv_fwd_c4_bf16": (_I32, [_P, _P, _P, _I32, _I32, _I32, _I32, _P, _P, _I32, _P]),
    "fem_meshconv_pack_c4_bf16": (_I32, [_P, _I32, _I32, _P, _P]),
    "fem_meshconv_features_c4": (_I32, [_P, _I64, _P, _P, _P, _P]),
    "fem_meshconv_correction": (_I32, [_P, _I32, _I64, _P, _D, _P, _P, _P]),
    "fem_pcg_replace_residual": (_I32, [_MP, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_mg_smooth": (_I32, [_MP, _P, _I32, _P, ctypes.c_float, _P, _P, _P, _I32, _P]),
    "fem_mg_restrict": (_I32, [_TP, _P, _P, _I64, _P]),
    "fem_mg_prolong_add": (_I32, [_TP, _P, _P, _I64, _P]),
    "fem_dense_gemv_f32": (_I32, [_P, _I64, _P, _P, _P]),
    "fem_mg_smooth_multi": (_I32, [_MP, _P, _P, ctypes.c_float, _P, _P, _P, _I64, _I32, _P]),
    "fem_dense_gemm_nb_f32": (_I32, [_P, _I64, _P, _P, _I64, _P]),
    "fem_gemm_tn_work_bytes": (_I64, [_I64, _I64]),
    "fem_gemm_tn_f64": (_I32, [_P, _I64, _P, _I64, _I64, _I64, _I64, _P, _P, _I64, _P]),
    "fem_gemm_tn_f64_noalloc": (_I32, [_P, _I64, _P, _I64, _I64, _I64, _I64, _P, _P]),
    "fem_gemm_nn_f64": (_I32, [_P, _I64, _P, _I64, _P, _I64, _I64, _I64, _I64, _D, _D, _P]),
    "fem_chol_upper_inv_f64": (_I32, [_P, _I64, _P, _P, _P]),
    "fem_fcg_work_bytes": (_I64, [_I64]),
    "fem_fcg_direction": (_I32, [_I64, _I64, _P, _I32, _P, _P, _P, _I32, _P, _I64, _P]),
    "fem_fcg_update": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_fcg_residual": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P]),
    "fem_fcg_direction_pr": (_I32, [_I64, _I64, _P, _I32, _P, _P, _P, _P, _I32, _P, _I64, _P]),
    "fem_fcg_update_split": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _I64, _P]),
}

_lib = None


def lib():
    """Load (once) and return the ctypes handle; raises if the .so is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"fem355 HIP library not built: {LIB_PATH} is missing "
                              "(run __graft_entry__.build())")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def last_error() -> str:
    return lib().fem_last_error().decode(errors="replace")


def check(rc: int):
    if rc != FEM_OK:
        raise FemError(rc, last_error())


def call(name, *args):
    """Call an int-returning entry point and raise FemError on failure."""
    check(getattr(lib(), name)(*args))
