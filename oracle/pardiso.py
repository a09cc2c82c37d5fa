"""Direct sparse solve for the CPU baseline (TEST INFRASTRUCTURE ONLY).

The reference factors K with Intel MKL Pardiso through pyMKL / pymatsolver
(SOLVE_static.py:31-33, SOLVE_eigen.py:32-33), which are absent here.  This
module calls MKL Pardiso's public C interface (``pardiso_64`` in libmkl_rt,
64-bit integers, so the call does not depend on the library's interface
layer) through ctypes: real symmetric positive definite (mtype 2), the upper
triangle in 1-based CSR, default iparm.  When libmkl_rt does not load it
falls back to SciPy's SuperLU (``scipy.sparse.linalg.splu``); ``kind`` says
which one ran.
"""
from __future__ import annotations

import ctypes
import glob
import os

import numpy as np
import scipy.sparse as sp

_MKL = None
_MKL_TRIED = False


def _mkl():
    global _MKL, _MKL_TRIED
    if not _MKL_TRIED:
        _MKL_TRIED = True
        cands = [os.environ.get("FEM_MKL_RT", "")] + sorted(glob.glob("/opt/conda/lib/libmkl_rt.so*")) + [
            "libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1"]
        for c in cands:
            if not c:
                continue
            try:
                h = ctypes.CDLL(c, mode=ctypes.RTLD_GLOBAL)
                f = h.pardiso_64
            except (OSError, AttributeError):
                continue
            P = ctypes.c_void_p
            f.restype = None
            f.argtypes = [P] * 16
            _MKL = h
            break
    return _MKL


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DirectSolver:
    """Factor an SPD scipy matrix once; ``solve(b)`` for (n,) or (n, k) b."""

    def __init__(self, K, prefer="pardiso"):
        K = sp.csr_matrix(K)
        self.n = K.shape[0]
        self._pt = None
        mkl = _mkl() if prefer == "pardiso" else None
        if mkl is not None:
            U = sp.triu(K, format="csr")
            U.sort_indices()
            self._a = np.ascontiguousarray(U.data, dtype=np.float64)
            self._ia = np.ascontiguousarray(U.indptr, dtype=np.int64) + 1
            self._ja = np.ascontiguousarray(U.indices, dtype=np.int64) + 1
            self._pt = np.zeros(64, dtype=np.int64)
            self._iparm = np.zeros(64, dtype=np.int64)        # iparm[0] = 0: MKL defaults
            self._perm = np.zeros(self.n, dtype=np.int64)
            self._mkl = mkl
            self._call(12, np.zeros(1), np.zeros(1), 1)        # analysis + numerical factorisation
            self.kind = "mkl_pardiso"
        else:
            from scipy.sparse.linalg import splu
            self._lu = splu(sp.csc_matrix(K))
            self.kind = "scipy_superlu"

    def _call(self, phase, b, x, nrhs):
        i32 = lambda v: np.array([v], dtype=np.int64)   # noqa: E731
        maxfct, mnum, mtype, ph, n, nr, msg, err = i32(1), i32(1), i32(2), i32(phase), i32(self.n), i32(nrhs), \
            i32(0), i32(0)
        self._mkl.pardiso_64(_ptr(self._pt), _ptr(maxfct), _ptr(mnum), _ptr(mtype), _ptr(ph), _ptr(n), _ptr(self._a),
                          _ptr(self._ia), _ptr(self._ja), _ptr(self._perm), _ptr(nr), _ptr(self._iparm), _ptr(msg),
                          _ptr(b), _ptr(x), _ptr(err))
        if int(err[0]) != 0:
            raise RuntimeError(f"MKL Pardiso phase {phase} returned error {int(err[0])}")

    def solve(self, b):
        b = np.asarray(b, dtype=np.float64)
        if self._pt is None:
            return self._lu.solve(b)
        one = b.ndim == 1
        B = b.reshape(self.n, -1)
        k = B.shape[1]
        Bf = np.asfortranarray(B)                              # Pardiso takes column-major blocks
        X = np.zeros_like(Bf)
        self._call(33, Bf, X, k)
        return X.reshape(-1) if one else np.ascontiguousarray(X)

    def release(self):
        if self._pt is not None:
            self._call(-1, np.zeros(1), np.zeros(1), 1)
            self._pt = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
