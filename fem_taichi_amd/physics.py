"""Mesh + pattern container with the reference's ``Read`` attributes.

``Physics`` carries the duck type that the reference's ``Model`` and ``Post``
consume (READ_physics.py:44-61, 96-97, 114, 134): ``Node`` (>=4, N) with
1-based IDs in row 0 and x, y, z in the last three rows, ``Elem`` (>=8, E)
whose last 8 rows are node ID - 1, ``Bounds``/``Forces`` lists,
``Nind``, ``SmNode``, ``ReNode``, ``CoNode``, ``CoElem``, ``Num_Node``,
``Num_Elem``, ``Num_MaxN``.  ``pre_process()`` builds the pattern on the GPU
(fem_pattern_count / fem_pattern_fill) and keeps the device copies, plus the
node -> element incidence list and the row groups, in ``self.dev``.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .groups import DEFAULT_BRICK, build_row_groups, build_tasks_torch, node_group_ids
from .mesh import is_grid_mesh
from .runtime import device, stream_ptr

# Upper bound on the elements one row group may touch (LDS staging budget).
MAX_GROUP_ELEMS = 48
# Large-displacement assembly: ~3 KB of LDS per element (gradients, F, F F^T,
# stress), so smaller groups: 2 x 2 x 1 node bricks (18 elements), or 4 rows
# of the Morton order, at most 24 elements.
KN_BRICK = (2, 2, 1)
KN_MORTON_ROWS = 4
KN_MAX_GROUP_ELEMS = 24
# Load vector and stress recovery (one thread per element, 392 B of LDS per
# element): 4 x 4 x 4 node bricks touch 125 elements, so each element is
# computed by about 2 groups instead of about 3 with the assembly's bricks;
# 64 rows of the Morton order otherwise, at most 128 elements.
FIELD_BRICK = (4, 4, 4)
FIELD_MORTON_ROWS = 64
FIELD_MAX_GROUP_ELEMS = 128


class DeviceMesh:
    """Device-resident mesh, pattern, incidences and row groups + the C descriptor."""

    def __init__(self, coords, elem, pattern, groups):
        self.t = {}
        dev = device()
        self.t["coords"] = torch.as_tensor(coords, dtype=torch.float64).contiguous().to(dev)
        self.t["elem"] = torch.as_tensor(elem, dtype=torch.int32).contiguous().to(dev)
        for k, v in pattern.items():
            self.t[k] = v
        for k in ("grp_row_ptr", "grp_rows", "grp_elem_ptr", "grp_elems", "grp_enodes", "grp_task_ptr", "tasks"):
            self.t[k] = torch.as_tensor(groups[k], dtype=torch.int32).contiguous().to(dev)
        for k in ("grp_row_rec", "grp_inc"):      # the field kernels' row records (field groups only)
            self.t[k] = torch.as_tensor(groups.get(k, np.zeros(0, np.int32)), dtype=torch.int32).contiguous().to(dev)
        self.num_node = int(coords.shape[0])
        self.num_elem = int(elem.shape[0])
        self.nnzb = int(pattern["re_node"].numel())
        m = _lib.FemMesh()
        m.num_node, m.num_elem, m.nnzb = self.num_node, self.num_elem, self.nnzb
        for name in ("coords", "elem", "sm_node", "re_node", "co_node", "co_elem", "ne_ptr", "ne_list",
                     "grp_row_ptr", "grp_rows", "grp_elem_ptr", "grp_elems", "grp_enodes", "grp_task_ptr",
                     "tasks", "grp_row_rec", "grp_inc"):
            setattr(m, name, self.t[name].data_ptr() if self.t[name].numel() else None)
        m.num_groups = int(groups["num_groups"])
        m.max_group_elems = int(groups["max_group_elems"])
        m.max_group_rows = int(groups["max_group_rows"])
        m.max_group_slots = int(groups["max_group_slots"])
        m.max_group_tasks = int(groups["max_group_tasks"])
        self.c = m
        self.groups_info = {k: groups[k] for k in ("num_groups", "max_group_elems", "max_group_rows",
                                                   "max_group_slots", "max_group_tasks")}

    @property
    def ref(self):
        return ctypes.byref(self.c)

    def __getitem__(self, k):
        return self.t[k]


def gpu_pattern(elem_idx: torch.Tensor, N: int):
    """Run the GPU pattern builder on (E, 8) int32 node indices (device)."""
    dev = device()
    E = int(elem_idx.shape[0])
    st = stream_ptr()
    co_elem = torch.empty(N, dtype=torch.int32, device=dev)
    ne_ptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
    ne_list = torch.empty(max(8 * E, 1), dtype=torch.int32, device=dev)
    co_node = torch.empty(N, dtype=torch.int32, device=dev)
    sm_node = torch.empty(N + 1, dtype=torch.int32, device=dev)
    wb = _lib.lib().fem_pattern_work_bytes(E, N)
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    _lib.call("fem_pattern_count", elem_idx.data_ptr() if E else None, E, N, co_elem.data_ptr(),
              ne_ptr.data_ptr(), ne_list.data_ptr(), co_node.data_ptr(), sm_node.data_ptr(),
              work.data_ptr(), wb, st)
    nnzb = int(sm_node[N].item())
    re_node = torch.empty(max(nnzb, 1), dtype=torch.int32, device=dev)
    _lib.call("fem_pattern_fill", elem_idx.data_ptr() if E else None, N, ne_ptr.data_ptr(),
              ne_list.data_ptr(), sm_node.data_ptr(), re_node.data_ptr(), st)
    return dict(co_elem=co_elem, ne_ptr=ne_ptr, ne_list=ne_list[:8 * E], co_node=co_node,
                sm_node=sm_node, re_node=re_node[:nnzb])


def row_records(groups, ne_ptr, ne_list, co_elem):
    """The row records of the row-owner load-vector and stress kernels
    (struct fem_mesh grp_row_rec / grp_inc): per position p of grp_rows, 12
    int32 = [node i, CoElem[i], incidence count c, overflow offset o, first 8
    incidences (local element << 3 | local node), -1 padded]; incidences 8..c-1
    of a row go to the overflow list at o.  Incidences keep the ascending
    element order of ne_list."""
    rows = np.asarray(groups["grp_rows"], dtype=np.int64)
    N = len(rows)
    ne_ptr = np.asarray(ne_ptr, dtype=np.int64)
    cnt = np.diff(ne_ptr)
    inc_val = (np.asarray(groups["ne_local"], dtype=np.int64) << 3) | (np.asarray(ne_list, dtype=np.int64) & 7)
    rec = np.full((N, 12), -1, dtype=np.int32)
    c = cnt[rows]
    rec[:, 0] = rows
    rec[:, 1] = np.asarray(co_elem)[rows]
    rec[:, 2] = c
    ovf_n = np.maximum(c - 8, 0)
    ovf_off = np.concatenate([[0], np.cumsum(ovf_n)[:-1]]) if N else np.zeros(0, np.int64)
    rec[:, 3] = ovf_off
    start = ne_ptr[rows]
    for q in range(8):
        has = c > q
        rec[has, 4 + q] = inc_val[start[has] + q]
    total = int(ovf_n.sum())
    inc = np.zeros(total, dtype=np.int32)
    if total:
        p_idx = np.repeat(np.arange(N), ovf_n)
        k = np.arange(total) - np.repeat(ovf_off, ovf_n) + 8
        inc[:] = inc_val[start[p_idx] + k]
    return rec.reshape(-1), inc


class Physics:
    """A mesh with boundary sets, in the reference's ``Read`` layout."""

    def __init__(self, Node, Elem, Bounds=None, Forces=None, name="mesh", node_dims=None,
                 brick=DEFAULT_BRICK, run_pre_process=True):
        self.name = name
        self.Node = np.asarray(Node)
        self.Elem = np.asarray(Elem)
        self.Bounds = list(Bounds) if Bounds is not None else []
        self.Forces = list(Forces) if Forces is not None else []
        self.node_dims = node_dims
        self.brick = brick
        self.dev = None
        if run_pre_process:
            self.pre_process()

    @classmethod
    def adopt(cls, obj):
        """A Physics for ``obj``: ``obj`` itself when it is one, else a Physics
        built from the attributes the reference's ``Read`` carries
        (READ_physics.py:44-61, 96-97, 114, 134: Node, Elem, Bounds, Forces,
        name), with the pattern on the GPU.  The reference's SciPy
        ``pre_process`` sets no device state, so a ``Read`` handed to ``Model``
        or ``Post`` is wrapped here; the wrapper is kept on the object (as
        ``_fem355_physics``) so Model and Post of one Read share one pattern."""
        if isinstance(obj, cls):
            if obj.dev is None:
                obj.pre_process()
            return obj
        cached = getattr(obj, "_fem355_physics", None)
        if isinstance(cached, cls) and cached._src[0] is obj.Node and cached._src[1] is obj.Elem:
            return cached
        for attr in ("Node", "Elem"):
            if not hasattr(obj, attr):
                raise TypeError(f"{type(obj).__name__} has no {attr}: not a Read-shaped mesh")
        phys = cls(obj.Node, obj.Elem, getattr(obj, "Bounds", None), getattr(obj, "Forces", None),
                   name=getattr(obj, "name", "mesh"), node_dims=getattr(obj, "node_dims", None))
        phys._src = (obj.Node, obj.Elem)      # a wrapper is reused only while these are unchanged
        try:
            obj._fem355_physics = phys
        except AttributeError:
            pass
        return phys

    @classmethod
    def from_grid(cls, grid, **kw):
        """From a ``mesh.StructuredGrid`` (keeps its brick-able node layout)."""
        return cls(grid.Node, grid.Elem, grid.Bounds, grid.Forces, name=grid.name,
                   node_dims=grid.node_dims, **kw)

    def pre_process(self):
        """READ_physics.py:42-61 on the GPU; the outputs equal SciPy's bit for bit."""
        self.Num_Node = self.Node.shape[-1]
        self.Num_Elem = self.Elem.shape[-1]
        ids = self.Node[0, :].astype(np.int64)
        self.Num_MaxN = int(ids.max()) if self.Num_Node else 0
        if self.Num_Node and ids.min() < 1:
            raise ValueError("node IDs must be >= 1")
        self.Nind = np.zeros((self.Num_MaxN,), dtype=np.int64)
        self.Nind[ids - 1] = np.arange(self.Num_Node, dtype=np.int64)
        elem_ids = np.asarray(self.Elem[-8:], dtype=np.int64)
        if elem_ids.size and (elem_ids.min() < 0 or elem_ids.max() >= self.Num_MaxN):
            raise ValueError("element references a node ID outside the Node table")
        en = self.Nind[elem_ids].T.astype(np.int32)          # (E, 8) node indices
        self.elem_nodes = np.ascontiguousarray(en)
        dev = device()
        elem_t = torch.as_tensor(self.elem_nodes).to(dev)
        pat = gpu_pattern(elem_t, self.Num_Node)
        self.SmNode = pat["sm_node"].cpu().numpy()
        self.ReNode = pat["re_node"].cpu().numpy()
        self.CoNode = pat["co_node"].cpu().numpy()
        self.CoElem = pat["co_elem"].cpu().numpy()
        ne_ptr = pat["ne_ptr"].cpu().numpy()
        ne_list = pat["ne_list"].cpu().numpy()
        coords = np.ascontiguousarray(np.asarray(self.Node[-3:], dtype=np.float64).T)
        gid = node_group_ids(self.Num_Node, node_dims=self.node_dims, xyz=coords, brick=self.brick)
        groups = build_row_groups(self.Num_Node, self.Num_Elem, ne_ptr, ne_list, gid, MAX_GROUP_ELEMS,
                                  co_node=self.CoNode)
        tp, tasks = build_tasks_torch(self.Num_Node, self.elem_nodes, self.SmNode, self.ReNode, ne_ptr, ne_list,
                                      groups, device=device())
        groups["grp_task_ptr"], groups["tasks"] = tp, tasks
        groups["grp_enodes"] = np.ascontiguousarray(self.elem_nodes[groups["grp_elems"]].reshape(-1))
        groups["max_group_tasks"] = int(np.diff(tp).max()) if len(tp) > 1 else 0
        lds = _lib.lib().fem_hex8_assemble_lds_bytes(groups["max_group_elems"], groups["max_group_rows"],
                                                      _lib.DTYPE_F64, 1, 1)
        if lds > 160 * 1024:
            raise ValueError(f"row groups need {lds} B of LDS; use a smaller brick")
        self.dev = DeviceMesh(coords, self.elem_nodes, pat, groups)
        # node grid of the structured-grid kernels: x-fastest node ids and the
        # element numbering of mesh.grid_connectivity (checked, not assumed)
        self.dev.grid_nodes = None
        if self.node_dims is not None:
            dims = tuple(int(d) for d in self.node_dims)
            if is_grid_mesh(self.elem_nodes, self.Num_Node, tuple(d - 1 for d in dims)):
                self.dev.grid_nodes = dims
        self._symbolic = (coords, pat, ne_ptr, ne_list)
        self._kn_dev = None
        self._fields_dev = None
        return self

    @property
    def fields_dev(self):
        """The same mesh and pattern with the larger row groups of the
        load-vector and stress kernels (no task lists; built on first use)."""
        if self._fields_dev is None:
            coords, pat, ne_ptr, ne_list = self._symbolic
            gid = node_group_ids(self.Num_Node, node_dims=self.node_dims, xyz=coords, brick=FIELD_BRICK,
                                 rows=FIELD_MORTON_ROWS)
            groups = build_row_groups(self.Num_Node, self.Num_Elem, ne_ptr, ne_list, gid, FIELD_MAX_GROUP_ELEMS,
                                      co_node=self.CoNode)
            groups["grp_task_ptr"] = np.zeros(int(groups["num_groups"]) + 1, dtype=np.int32)
            groups["tasks"] = np.zeros(0, dtype=np.int32)
            groups["max_group_tasks"] = 0
            groups["grp_enodes"] = np.ascontiguousarray(self.elem_nodes[groups["grp_elems"]].reshape(-1))
            groups["grp_row_rec"], groups["grp_inc"] = row_records(groups, ne_ptr, ne_list, self.CoElem)
            self._fields_dev = DeviceMesh(coords, self.elem_nodes, pat, groups)
        return self._fields_dev

    @property
    def kn_dev(self):
        """The same mesh and pattern with the smaller row groups of the
        large-displacement assembly (built on first use)."""
        if self._kn_dev is None:
            coords, pat, ne_ptr, ne_list = self._symbolic
            gid = node_group_ids(self.Num_Node, node_dims=self.node_dims, xyz=coords, brick=KN_BRICK,
                                 rows=KN_MORTON_ROWS)
            groups = build_row_groups(self.Num_Node, self.Num_Elem, ne_ptr, ne_list, gid, KN_MAX_GROUP_ELEMS,
                                      co_node=self.CoNode)
            tp, tasks = build_tasks_torch(self.Num_Node, self.elem_nodes, self.SmNode, self.ReNode, ne_ptr,
                                          ne_list, groups, device=device())
            groups["grp_task_ptr"], groups["tasks"] = tp, tasks
            groups["grp_enodes"] = np.ascontiguousarray(self.elem_nodes[groups["grp_elems"]].reshape(-1))
            groups["max_group_tasks"] = int(np.diff(tp).max()) if len(tp) > 1 else 0
            lds = _lib.lib().fem_hex8_kn_lds_bytes(groups["max_group_elems"], groups["max_group_rows"])
            if lds > 160 * 1024:
                raise ValueError(f"large-displacement row groups need {lds} B of LDS")
            self._kn_dev = DeviceMesh(coords, self.elem_nodes, pat, groups)
        return self._kn_dev
