"""Row groups for the row-owner kernels (one-time symbolic work, host side).

A row group is a spatially compact set of node rows; the assembly kernel
stages the elements touching a group's rows in LDS and writes every block of
those rows once.  Structured grids are cut into node bricks; other meshes are
cut along a Morton (Z-order) sort of the node coordinates.  Groups whose
element count would overflow the LDS budget are split in halves until they
fit.
"""
from __future__ import annotations

import numpy as np

# Node brick of a structured grid: 4 x 2 x 2 rows touch at most 5*3*3 = 45
# elements, which keeps the FP64 staging of a group near 76 KB (two
# work-groups per CU's 160 KB of LDS).
DEFAULT_BRICK = (4, 2, 2)
DEFAULT_MORTON_ROWS = 8


def _morton_order(xyz: np.ndarray, bits: int = 10) -> np.ndarray:
    lo = xyz.min(axis=0)
    span = np.maximum(xyz.max(axis=0) - lo, 1e-300)
    q = np.floor((xyz - lo) / span * ((1 << bits) - 1)).astype(np.uint64)
    code = np.zeros(len(xyz), dtype=np.uint64)
    for b in range(bits):
        for d in range(3):
            code |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    return np.argsort(code, kind="stable")


def node_group_ids(N, node_dims=None, xyz=None, brick=DEFAULT_BRICK, rows=DEFAULT_MORTON_ROWS):
    """Group id of every node (groups numbered in processing order)."""
    if node_dims is not None:
        nxn, nyn, nzn = node_dims
        assert nxn * nyn * nzn == N
        bx, by, bz = brick
        idx = np.arange(N)
        i, j, k = idx % nxn, (idx // nxn) % nyn, idx // (nxn * nyn)
        gx, gy = -(-nxn // bx), -(-nyn // by)
        return (i // bx) + gx * ((j // by) + gy * (k // bz))
    order = _morton_order(np.asarray(xyz, dtype=np.float64))
    gid = np.empty(N, dtype=np.int64)
    gid[order] = np.arange(N) // rows
    return gid


def build_row_groups(N, E, ne_ptr, ne_list, gid, max_elems, co_node=None):
    """Row groups from a node -> group map.

    Returns dict with grp_row_ptr, grp_rows, grp_elem_ptr, grp_elems,
    ne_local (int32 arrays), num_groups, max_group_elems, max_group_rows.
    Groups whose element union exceeds ``max_elems`` are split in halves.
    """
    ne_ptr = np.asarray(ne_ptr, dtype=np.int64)
    inc_node = np.repeat(np.arange(N, dtype=np.int64), np.diff(ne_ptr))
    inc_e = (np.asarray(ne_list, dtype=np.int64) >> 3)
    gid = np.asarray(gid, dtype=np.int64).copy()
    while True:
        # renumber groups densely, keeping their order
        _, gid = np.unique(gid, return_inverse=True)
        gid = gid.astype(np.int64)
        G = int(gid.max()) + 1 if N else 0
        keys = gid[inc_node] * max(E, 1) + inc_e
        uniq = np.unique(keys)
        ug = uniq // max(E, 1)
        counts = np.bincount(ug, minlength=G)
        bad = np.nonzero(counts > max_elems)[0]
        if len(bad) == 0:
            break
        # split each oversized group: second half of its rows gets a new id
        order = np.argsort(gid, kind="stable")
        starts = np.searchsorted(gid[order], np.arange(G))
        ends = np.searchsorted(gid[order], np.arange(G), side="right")
        newgid = gid * 2
        for g in bad:
            rows = order[starts[g]:ends[g]]
            if len(rows) <= 1:
                raise ValueError(f"a single node touches {counts[g]} elements (> {max_elems})")
            newgid[rows[len(rows) // 2:]] += 1
        gid = newgid
    order = np.argsort(gid, kind="stable")
    grp_row_ptr = np.concatenate([[0], np.cumsum(np.bincount(gid, minlength=G))]).astype(np.int32)
    grp_elem_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    grp_elems = (uniq % max(E, 1)).astype(np.int32)
    ne_local = (np.searchsorted(uniq, keys) - grp_elem_ptr[gid[inc_node]]).astype(np.int32)
    slots = np.bincount(gid, weights=co_node, minlength=G) if (co_node is not None and G) else np.zeros(1)
    incs = np.bincount(gid, weights=np.diff(ne_ptr), minlength=G) if G else np.zeros(1)
    return dict(grp_row_ptr=grp_row_ptr, grp_rows=order.astype(np.int32), grp_elem_ptr=grp_elem_ptr,
                grp_elems=grp_elems, ne_local=ne_local, num_groups=G,
                max_group_elems=int(counts.max()) if G else 0,
                max_group_rows=int(np.diff(grp_row_ptr).max()) if G else 0,
                max_group_slots=int(slots.max()) if G else 0,
                max_group_incs=int(incs.max()) if G else 0)
