"""Drop-in ``Post`` (the reference's POST_process.Post stress part) on the GPU.

``Get_Stress()`` returns the (N, 10) array of POST_process.py:72-99:
[sxx, syy, szz, txy, tyz, txz, principal 1..3 (descending), von Mises term]
computed by fem_stress_recover (nodal averaging over CoElem, the
reference's Gauss-to-node extrapolation map as written, Jacobi eigenvalues
in FP64 and the reference's von Mises expression).

``Output_field_BIN`` / ``Output_field_ASC`` keep the reference's writers'
names, arguments and output files (Results/<name>/<file>.plt or .dat,
POST_process.py:32-69, 101-135); the binary file is written by
``tecplot.write_plt`` from the published format, not through TecIO.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import _lib
from .runtime import device, stream_ptr


class Post:
    def __init__(self, name, Physics, Materials, Solutions):
        self.name = name
        self.path = None           # Results/<name>/, created by the writers
        self.DE, self.EX, self.PO = Materials['DE'], Materials['EX'], Materials['PO']
        from .physics import Physics as _Physics
        Physics = _Physics.adopt(Physics)      # also takes the reference's Read object
        self.Physics = Physics
        self.Num_Node = Physics.Node.shape[-1]
        self.Num_Elem = Physics.Elem.shape[-1]
        if isinstance(Solutions, torch.Tensor):
            sol = Solutions.detach().double().cpu().numpy()
        else:
            sol = np.asarray(Solutions, dtype=np.float64)
        # the reference keeps Solutions.squeeze().T, i.e. (3, N)   (POST_process.py:28)
        self.Solutions = sol.squeeze().T if sol.ndim > 1 else sol.reshape(-1, 3).T
        self._u = torch.as_tensor(np.ascontiguousarray(self.Solutions.T).reshape(-1)).to(device())

    def Get_Stress(self):
        N, E = self.Num_Node, self.Num_Elem
        ex = np.asarray(self.EX, dtype=np.float64)
        young = None if ex.ndim == 0 else torch.as_tensor(ex).to(device())
        out = torch.empty(N * 10, dtype=torch.float64, device=device())
        _lib.call("fem_stress_recover", self.Physics.fields_dev.ref, young.data_ptr() if young is not None else None,
                  float(ex) if ex.ndim == 0 else 0.0, float(self.PO), self._u.data_ptr(), out.data_ptr(),
                  None, 0, stream_ptr())
        return out.cpu().numpy().reshape(N, 10)

    # ---- Tecplot output (POST_process.py:32-69, 101-135) ---------------------

    def _elem1(self):
        """(E, 8) 1-based node indices, Nind[Elem] + 1 as the reference builds them."""
        en = np.asarray(self.Physics.Elem[-8:], dtype=np.int64)
        return (self.Physics.Nind[en] + 1).T

    def Output_field_BIN(self, file_name, Fields, Field_names):
        """Binary Tecplot file Results/<name>/<file_name>.plt: x, y, z and, for
        each step i of Fields (N, F, S), every field name + str(i + 1) (the
        reference's variable list), one FEBRICK zone."""
        from . import tecplot
        self.path = tecplot.results_dir(self.name)
        xyz = np.asarray(self.Physics.Node[-3:], dtype=np.float64).T
        names, values = tecplot.field_table(xyz, Fields, Field_names)
        out = os.path.join(self.path, file_name + ".plt")
        tecplot.write_plt(out, "Solution", names, values, self._elem1() - 1)
        return out

    def Output_field_ASC(self, file_name, Fields, Field_names):
        """ASCII Tecplot file Results/<name>/<file_name>.dat in the reference's
        text layout (FEPOINT, ET=BRICK)."""
        from . import tecplot
        self.path = tecplot.results_dir(self.name)
        xyz = np.asarray(self.Physics.Node[-3:], dtype=np.float64).T
        out = os.path.join(self.path, file_name + ".dat")
        tecplot.write_dat(out, self.name, xyz, Fields, Field_names, self._elem1())
        return out
