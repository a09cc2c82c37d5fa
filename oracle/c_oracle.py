"""ctypes wrapper of oracle/libfem_oracle.so (the C restatement of the
reference's CPU assembly; TEST INFRASTRUCTURE ONLY -- the CPU baseline)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libfem_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"{_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(_PATH)
        P = ctypes.c_void_p
        h.ref_assemble_K_csr.restype = None
        h.ref_assemble_K_csr.argtypes = [ctypes.c_int64, ctypes.c_int64, P, P, P, ctypes.c_double,
                                         ctypes.c_double, P, P, P, P, P]
        h.ref_assemble_M_csr.restype = None
        h.ref_assemble_M_csr.argtypes = [ctypes.c_int64, P, P, ctypes.c_double, P, P, P, P]
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def assemble_K_csr(coords, elem, young, young_scalar, nu, sm, co, re, with_cols=True):
    """ValK (and ColK) of InstallMatrixK (FEM_model_csr.py:150-165), unpenalised."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.int32)
    sm = np.ascontiguousarray(sm, dtype=np.int32)
    co = np.ascontiguousarray(co, dtype=np.int32)
    re = np.ascontiguousarray(re, dtype=np.int32)
    yv = None if young is None else np.ascontiguousarray(young, dtype=np.float64)
    nnzb = int(sm[-1])
    val = np.zeros(9 * nnzb)
    col = np.zeros(9 * nnzb, dtype=np.int32) if with_cols else None
    lib().ref_assemble_K_csr(coords.shape[0], elem.shape[0], _p(coords), _p(elem), _p(yv),
                             float(young_scalar), float(nu), _p(sm), _p(co), _p(re), _p(val), _p(col))
    return val, col


def assemble_M_csr(coords, elem, rho, sm, co, re):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.int32)
    sm = np.ascontiguousarray(sm, dtype=np.int32)
    co = np.ascontiguousarray(co, dtype=np.int32)
    re = np.ascontiguousarray(re, dtype=np.int32)
    val = np.zeros(3 * int(sm[-1]))
    lib().ref_assemble_M_csr(elem.shape[0], _p(coords), _p(elem), float(rho), _p(sm), _p(co), _p(re),
                             _p(val))
    return val
