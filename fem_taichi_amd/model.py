"""Drop-in ``Model`` (the reference's FEM_model_csr.Model / FEM_model_bsr.Model
linear part) backed by the fem355 HIP kernels.

Public methods keep the reference's names and return types:
``Get_MatrixKL()`` / ``Get_MatrixK()`` -> scipy CSR (FEM_model_csr.py:118-127),
``Get_MatrixM()`` -> scipy CSR (:129-137), ``Get_VectorF()`` -> ndarray (3N,)
(:141-146).  Device-resident forms for the GPU solvers are exposed as
``K_bsr(dtype)``, ``M_blocks()`` and ``F_device()``.

Deliberate deviations from the reference, each a fix of a reference quirk:
repeated calls do not accumulate (the reference's fields are never reset,
so a second ``Get_MatrixKL`` doubles K); nodal loads keep their float value
unless ``int_loads=True`` (the reference casts them to int,
FEM_model_csr.py:78); ``Materials['EX']`` may be a length-E array of
per-element moduli (the reference has one scalar).
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .runtime import device, stream_ptr
from .runtime import to_host as _to_host

PENALTY = 1e20   # FEM_model_csr.py:241


def stage_bounds(Bounds):
    """FEM_model_csr.py:64-73: (4, nb) [ID-1, c0, c1, c2]; with no bounds the
    reference stages one zero column, which fixes node Nind[0]."""
    tmp = []
    for ids, c in Bounds:
        ids = np.asarray(ids, dtype=np.float64)[None, :]
        cc = np.repeat(np.asarray(c, dtype=np.float64)[:, None], ids.shape[-1], axis=-1)
        tmp.append(np.concatenate((ids, cc), axis=0))
    return np.concatenate(tmp, axis=-1) if tmp else np.zeros((4, 1))


def stage_forces(Forces, int_loads=False):
    """FEM_model_csr.py:75-84 (with the int cast of :78 only when int_loads)."""
    tmp = []
    for ids, v in Forces:
        ids = np.asarray(ids, dtype=np.float64)[None, :]
        vv = np.asarray(v, dtype=np.float64)
        if int_loads:
            vv = np.trunc(vv)
        tmp.append(np.concatenate((ids, np.repeat(vv[:, None], ids.shape[-1], axis=-1)), axis=0))
    return np.concatenate(tmp, axis=-1) if tmp else np.zeros((4, 1))


class Model:
    """Linear hex8 model: stiffness K, consistent mass M and load F on the GPU."""

    def __init__(self, name, Physics, Materials, Omegas=(0, 0, 100), Largedisp=False, int_loads=False):
        self.name = name
        self.DE, self.EX, self.PO = Materials['DE'], Materials['EX'], Materials['PO']
        self.OM = np.array(Omegas, dtype=np.float64)
        from .physics import Physics as _Physics
        Physics = _Physics.adopt(Physics)      # also takes the reference's Read object
        self.Physics = Physics
        self.Largedisp = Largedisp
        self.dev = Physics.dev
        self.Num_Node = Physics.Num_Node
        self.Num_Elem = Physics.Num_Elem
        self.Num_MaxN = Physics.Num_MaxN
        self.Num_Free = self.Num_Node * 3
        self.nnzb = int(Physics.SmNode[-1])
        self.nnz_K, self.nnz_M = self.nnzb * 9, self.nnzb * 3
        d = device()
        ex = np.asarray(self.EX, dtype=np.float64)
        if ex.ndim == 0:
            self.young_scalar, self.young = float(ex), None
        else:
            if ex.shape != (self.Num_Elem,):
                raise ValueError(f"Materials['EX'] must be a scalar or have shape ({self.Num_Elem},)")
            self.young_scalar, self.young = 0.0, torch.as_tensor(ex).to(d)
        bounds = stage_bounds(Physics.Bounds)
        self.bounds_temp = bounds
        nind = Physics.Nind
        bid = bounds[0].astype(np.int64)
        if bid.size and (bid.min() < 0 or bid.max() >= len(nind)):
            raise ValueError("a bound node ID is outside the Node table")
        self.bound_node = torch.as_tensor(nind[bid].astype(np.int32)).to(d)
        self.bound_flag = torch.as_tensor(np.ascontiguousarray(bounds[1:4].T)).to(d)
        forces = stage_forces(Physics.Forces, int_loads)
        self.forces_temp = forces
        load = np.zeros((self.Num_Node, 3))
        fid = forces[0].astype(np.int64)
        if fid.size and (fid.min() < 0 or fid.max() >= len(nind)):
            raise ValueError("a loaded node ID is outside the Node table")
        np.add.at(load, nind[fid], forces[1:4].T)
        self.nodal_load = torch.as_tensor(load.reshape(-1)).to(d)
        self._cache = {}
        self._csr_k_dev = None

    # ---------------------------------------------------------------- device forms
    def _young_ptr(self):
        return self.young.data_ptr() if self.young is not None else None

    def assemble_K(self, dtype="f64", layout="bsr", penalty=True, out=None):
        """K values (9*nnzb) on the device; BSR (nnzb,3,3) or the reference's CSR order."""
        tdt = torch.float64 if dtype == "f64" else torch.float32
        cdt = _lib.DTYPE_F64 if dtype == "f64" else _lib.DTYPE_F32
        lay = _lib.LAYOUT_BSR if layout == "bsr" else _lib.LAYOUT_CSR
        vals = out if out is not None else torch.empty(9 * self.nnzb, dtype=tdt, device=device())
        st = stream_ptr()
        _lib.call("fem_hex8_assemble", self.dev.ref, self._young_ptr(), self.young_scalar, float(self.PO),
                  float(self.DE), vals.data_ptr(), cdt, lay, None, st)
        if penalty:
            self.apply_penalty(vals, dtype, layout)
        return vals

    def apply_penalty(self, vals, dtype="f64", layout="bsr"):
        """ModifyK (FEM_model_csr.py:233-242): SET constrained diagonals to 1e20."""
        cdt = _lib.DTYPE_F64 if dtype == "f64" else _lib.DTYPE_F32
        lay = _lib.LAYOUT_BSR if layout == "bsr" else _lib.LAYOUT_CSR
        _lib.call("fem_apply_penalty", self.dev.ref, self.bound_node.data_ptr(), self.bound_flag.data_ptr(),
                  int(self.bound_node.numel()), vals.data_ptr(), cdt, lay, PENALTY, stream_ptr())

    def M_blocks(self):
        """Mass coupling scalar of every node-pair block (nnzb,), device f64."""
        if "M" not in self._cache:
            mb = torch.empty(self.nnzb, dtype=torch.float64, device=device())
            _lib.call("fem_hex8_assemble", self.dev.ref, None, 0.0, float(self.PO), float(self.DE), None,
                      _lib.DTYPE_F64, _lib.LAYOUT_BSR, mb.data_ptr(), stream_ptr())
            self._cache["M"] = mb
        return self._cache["M"]

    def K_bsr(self, dtype="f64"):
        key = "K" + dtype
        if key not in self._cache:
            self._cache[key] = self.assemble_K(dtype, "bsr")
        return self._cache[key]

    def F_device(self):
        """The load vector (3N,) f64 on the device."""
        if "F" not in self._cache:
            f = torch.empty(3 * self.Num_Node, dtype=torch.float64, device=device())
            om = (ctypes.c_double * 3)(*[float(w) for w in self.OM])
            _lib.call("fem_body_force", self.Physics.fields_dev.ref, float(self.DE), om,
                      self.nodal_load.data_ptr(), f.data_ptr(), None, 0, stream_ptr())
            self._cache["F"] = f
        return self._cache["F"]

    # ---------------------------------------------------------------- reference API
    # ---- large displacement (FEM_model_bsr.py:134-145, 476-598) ----------------

    def LDinit(self, displace):
        """FEM_model_bsr.py:588-598: the displacement (3N,) the large-displacement
        matrices are formed at (kept on the device)."""
        if torch.is_tensor(displace):
            d = displace.detach().to(device=device(), dtype=torch.float64).reshape(-1)
        else:
            d = torch.as_tensor(np.asarray(displace, dtype=np.float64).reshape(-1)).to(device())
        if d.numel() != 3 * self.Num_Node:
            raise ValueError(f"displacement must have {3 * self.Num_Node} entries")
        self.U_ld = d.contiguous().clone()

    def assemble_KN(self, layout="bsr", penalty=True):
        """(KT, KM) values (9*nnzb each) on the device at the displacement of
        LDinit (zero if never set): the tangent and the matrix of the residual
        KM U - F (UnitK, FEM_model_bsr.py:476-505), penalty on both (:248-259)."""
        if getattr(self, "U_ld", None) is None:
            self.U_ld = torch.zeros(3 * self.Num_Node, dtype=torch.float64, device=device())
        kdev = self.Physics.kn_dev
        lay = _lib.LAYOUT_BSR if layout == "bsr" else _lib.LAYOUT_CSR
        kt = torch.empty(9 * self.nnzb, dtype=torch.float64, device=device())
        km = torch.empty_like(kt)
        _lib.call("fem_hex8_assemble_kn", kdev.ref, self._young_ptr(), self.young_scalar, float(self.PO),
                  self.U_ld.data_ptr(), kt.data_ptr(), km.data_ptr(), lay, stream_ptr())
        if penalty:
            self.apply_penalty(kt, "f64", layout)
            self.apply_penalty(km, "f64", layout)
        return kt, km

    def Get_MatrixKN(self):
        """FEM_model_bsr.py:134-145: (KT, KM) as scipy CSR (3N x 3N)."""
        kt, km = self.assemble_KN("csr")
        n = self.Num_Free
        out = []
        for v in (kt, km):
            row, col = self._csr_index_K()          # each matrix its own index arrays
            out.append(sp.csr_matrix((_to_host(v), col, row), shape=(n, n), dtype=np.float64))
        return tuple(out)

    # ---- prestressed stiffness for modal analysis (no counterpart in the
    #      reference; Readme.md:189 lists stress stiffening and spin softening
    #      as missing) --------------------------------------------------------

    def assemble_KS(self, displace, omega=None, layout="bsr", penalty=True):
        """K_L + K_sigma(U) - spin softening, values (9*nnzb) on the device:
        the linear stiffness, the stress-stiffening matrix of the linear
        stress of the displacement U (fem_hex8_assemble_ks), and, when
        ``omega`` (rad/s, 3 components) is given, the spin softening
        -m_jk (|w|^2 I - w w^T) of the rotating frame (fem_spin_softening);
        then the penalty."""
        if torch.is_tensor(displace):
            u = displace.detach().to(device=device(), dtype=torch.float64).reshape(-1).contiguous()
        else:
            u = torch.as_tensor(np.asarray(displace, dtype=np.float64).reshape(-1)).to(device())
        if u.numel() != 3 * self.Num_Node:
            raise ValueError(f"displacement must have {3 * self.Num_Node} entries")
        lay = _lib.LAYOUT_BSR if layout == "bsr" else _lib.LAYOUT_CSR
        k = torch.empty(9 * self.nnzb, dtype=torch.float64, device=device())
        st = stream_ptr()
        _lib.call("fem_hex8_assemble_ks", self.Physics.kn_dev.ref, self._young_ptr(), self.young_scalar,
                  float(self.PO), u.data_ptr(), k.data_ptr(), lay, st)
        if omega is not None:
            w = (ctypes.c_double * 3)(*[float(v) for v in omega])
            _lib.call("fem_spin_softening", self.dev.ref, self.M_blocks().data_ptr(), w, k.data_ptr(), lay, st)
        if penalty:
            self.apply_penalty(k, "f64", layout)
        return k

    def Get_MatrixKS(self, displace, omega=None):
        """assemble_KS as scipy CSR (3N x 3N), the reference's CSR layout."""
        k = self.assemble_KS(displace, omega, "csr")
        row, col = self._csr_index_K()
        n = self.Num_Free
        return sp.csr_matrix((_to_host(k), col, row), shape=(n, n), dtype=np.float64)

    def _csr_index_K(self):
        """RowK / ColK of the reference's CSR layout (FEM_model_csr.py:163, 168-173) as fresh host
        arrays.  They depend on the pattern only, so they are built on the device once per Model and
        kept there; every returned matrix gets its own host copy (as the reference's to_numpy() does),
        so a caller that edits one matrix's structure in place (eliminate_zeros, sort) leaves every
        other matrix alone."""
        if self._csr_k_dev is None:
            d = device()
            row = torch.empty(3 * self.Num_Node + 1, dtype=torch.int32, device=d)
            col = torch.empty(max(self.nnz_K, 1), dtype=torch.int32, device=d)
            _lib.call("fem_csr_index_K", self.dev.ref, row.data_ptr(), col.data_ptr(), stream_ptr())
            self._csr_k_dev = (row, col[:self.nnz_K])
        row, col = self._csr_k_dev
        return _to_host(row), _to_host(col)

    def Get_MatrixKL(self):
        """FEM_model_csr.py:118-127: penalised stiffness as scipy CSR (3N x 3N)."""
        vals = self.assemble_K("f64", "csr")
        row, col = self._csr_index_K()
        n = self.Num_Free
        return sp.csr_matrix((_to_host(vals), col, row), shape=(n, n), dtype=np.float64)

    Get_MatrixK = Get_MatrixKL

    def Get_MatrixM(self):
        """FEM_model_csr.py:129-137: consistent mass as scipy CSR (3N x 3N)."""
        mb = self.M_blocks()
        d = device()
        row = torch.empty(3 * self.Num_Node + 1, dtype=torch.int32, device=d)
        col = torch.empty(max(self.nnz_M, 1), dtype=torch.int32, device=d)
        val = torch.empty(max(self.nnz_M, 1), dtype=torch.float64, device=d)
        st = stream_ptr()
        _lib.call("fem_csr_index_M", self.dev.ref, row.data_ptr(), col.data_ptr(), st)
        _lib.call("fem_csr_values_M", self.dev.ref, mb.data_ptr(), val.data_ptr(), st)
        n = self.Num_Free
        return sp.csr_matrix((_to_host(val[:self.nnz_M]), _to_host(col[:self.nnz_M]), _to_host(row)),
                             shape=(n, n), dtype=np.float64)

    def Get_VectorF(self):
        """FEM_model_csr.py:141-146: load vector in DOF order 3*i + p."""
        return self.F_device().cpu().numpy()
