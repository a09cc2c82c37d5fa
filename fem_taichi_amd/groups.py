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

# Node brick of a structured grid: 2 x 4 x 2 rows touch at most 3*5*3 = 45
# elements, which keeps the FP64 staging of a group near 76 KB (two
# work-groups per CU's 160 KB of LDS).  Of the three orientations of the
# 16-row brick, 2 x 4 x 2 assembles about 2% faster than 4 x 2 x 2 on both the
# 60x60x55 slab and the 400x50x50 blade (measured; the values are identical).
DEFAULT_BRICK = (2, 4, 2)
DEFAULT_MORTON_ROWS = 8
# max_group_slots when the row block counts are not given (int32 maximum)
MAX_SLOTS_UNKNOWN = 2 ** 31 - 1


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
        if nxn * nyn * nzn != N:
            raise ValueError(f"node_dims {tuple(node_dims)} do not multiply to the {N} nodes")
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
    # without CoNode the slot count is unknown: report the int32 maximum, which
    # keeps the assembly's LDS-staged copy-out off (it needs every group's
    # blocks to fit its gradient area)
    slots = np.bincount(gid, weights=co_node, minlength=G) if (co_node is not None and G) else None
    incs = np.bincount(gid, weights=np.diff(ne_ptr), minlength=G) if G else np.zeros(1)
    return dict(grp_row_ptr=grp_row_ptr, grp_rows=order.astype(np.int32), grp_elem_ptr=grp_elem_ptr,
                grp_elems=grp_elems, ne_local=ne_local, num_groups=G,
                max_group_elems=int(counts.max()) if G else 0,
                max_group_rows=int(np.diff(grp_row_ptr).max()) if G else 0,
                max_group_slots=(int(slots.max()) if slots is not None else MAX_SLOTS_UNKNOWN) if G else 0,
                max_group_incs=int(incs.max()) if G else 0)


# Order of the tasks inside a group: class (20 - log2 T for blocks of T
# lanes, decreasing T; 21 for the dual tasks after them), then the block (or
# dual pair) index, then the lane within the block.  Packed into one int64 as
# class * 2^52 + index * 64 + lane: the index may be any block number of the
# mesh (< 2^46), so the class always dominates.
_KEY_CLASS = 1 << 52


def task_sort_key(cls, idx, lane):
    """Sort key of a task (numpy or torch int64 arrays, or ints)."""
    return cls * _KEY_CLASS + idx * 64 + lane


def _next_pow2(a):
    a = np.maximum(a, 1)
    return (1 << np.ceil(np.log2(a)).astype(np.int64)).astype(np.int64)


def build_tasks(N, elem, sm_node, re_node, ne_ptr, ne_list, groups):
    """Contribution tasks of the row-owner assembly kernel (one-time symbolic work).

    Every output block (row i, slot m) is the sum over the elements e that
    contain both node i (local index j) and the column node (local index k)
    of the element Gram block G_e(j, k).  The contributions of a block are
    split into T = next_pow2(ceil(n/2)) tasks of at most two contributions;
    the T lanes of one block are adjacent and T-aligned (tasks of a group are
    ordered by decreasing T), so the kernel combines them with a fixed
    butterfly of lane shuffles.  Blocks with a single contribution are paired
    into "dual" tasks with two outputs.

    Returns grp_task_ptr (G+1) and tasks (ntask, 2) int32:
      [0] out_a | out_b << 16, out = row_local << 8 | m (m: slot within the
          row); out_b = 0xFFFF unless the task is dual
      [1] c1 | c2 << 14 | log2(T) << 28 | dual << 31, c = le << 6 | j << 3 | k,
          c = 0x3FFF when absent.
    """
    ne_ptr = np.asarray(ne_ptr, dtype=np.int64)
    elem = np.asarray(elem, dtype=np.int64)
    sm_node = np.asarray(sm_node, dtype=np.int64)
    G = int(groups["num_groups"])
    grp_row_ptr = groups["grp_row_ptr"].astype(np.int64)
    grp_rows = groups["grp_rows"].astype(np.int64)
    gid = np.repeat(np.arange(G, dtype=np.int64), np.diff(grp_row_ptr))
    g_of_node = np.empty(N, dtype=np.int64)
    r_of_node = np.empty(N, dtype=np.int64)
    g_of_node[grp_rows] = gid
    r_of_node[grp_rows] = np.arange(N, dtype=np.int64) - grp_row_ptr[gid]
    if int(np.diff(grp_row_ptr).max(initial=0)) > 255:
        raise ValueError("row groups of more than 255 rows")
    inc_node = np.repeat(np.arange(N, dtype=np.int64), np.diff(ne_ptr))
    ne_list = np.asarray(ne_list, dtype=np.int64)
    inc_e = ne_list >> 3
    inc_j = ne_list & 7
    inc_le = np.asarray(groups["ne_local"], dtype=np.int64)
    if inc_le.size and inc_le.max() > 254:
        raise ValueError("more than 255 elements in a row group")
    # expand to (incidence, k)
    I = len(inc_node)
    node = np.repeat(inc_node, 8)
    k = np.tile(np.arange(8, dtype=np.int64), I)
    e = np.repeat(inc_e, 8)
    col = elem[e, k]
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(sm_node))
    keys = rows * N + np.asarray(re_node, dtype=np.int64)
    slot = np.searchsorted(keys, node * N + col)
    m = slot - sm_node[node]
    if m.size and m.max() > 255:
        raise ValueError("a row with more than 256 neighbour blocks")
    g = g_of_node[node]
    r = r_of_node[node]
    cc = (np.repeat(inc_le, 8) << 6) | (np.repeat(inc_j, 8) << 3) | k
    order = np.lexsort((e, m, r, g))
    g, r, m, cc = g[order], r[order], m[order], cc[order]
    out_key = (r << 8) | m
    # runs of equal (g, out_key) = one output block
    skey = g * 65536 + out_key
    start = np.nonzero(np.concatenate([[True], skey[1:] != skey[:-1]]))[0]
    n = np.diff(np.concatenate([start, [len(skey)]]))
    if n.size and n.max() > 64:
        raise ValueError("a node pair shared by more than 64 elements")
    T = _next_pow2((n + 1) // 2)
    sg, sout = g[start], out_key[start]
    NONE = 0x3FFF
    recs = []   # arrays of (g, sortkey, outA, outB, word2)
    multi = n >= 2
    # non-single blocks: T tasks each, task q holds contributions 2q, 2q+1
    ms = np.nonzero(multi)[0]
    if len(ms):
        Tm = T[ms]
        tb = np.repeat(ms, Tm)
        q = np.arange(Tm.sum()) - np.repeat(np.cumsum(Tm) - Tm, Tm)
        i1 = start[tb] + 2 * q
        i2 = i1 + 1
        has1 = 2 * q < n[tb]
        has2 = 2 * q + 1 < n[tb]
        c1 = np.where(has1, cc[np.minimum(i1, len(cc) - 1)], NONE)
        c2 = np.where(has2, cc[np.minimum(i2, len(cc) - 1)], NONE)
        lt = np.log2(T[tb]).astype(np.int64)
        w2 = c1 | (c2 << 14) | (lt << 28)
        # decreasing T inside a group keeps every block's lanes T-aligned
        recs.append((sg[tb], task_sort_key(20 - lt, tb, q), sout[tb], np.full(len(tb), -1), w2))
    ss = np.nonzero(~multi)[0]
    if len(ss):
        # pair singles of the same group into dual tasks
        gs = sg[ss]
        o = np.argsort(gs, kind="stable")
        ss, gs = ss[o], gs[o]
        first = np.concatenate([[True], gs[1:] != gs[:-1]])
        pos = np.arange(len(ss)) - np.maximum.accumulate(np.where(first, np.arange(len(ss)), 0))
        a_idx = np.nonzero(pos % 2 == 0)[0]
        b_idx = a_idx + 1
        b_ok = (b_idx < len(ss))
        b_ok[b_ok] &= (gs[b_idx[b_ok]] == gs[a_idx[b_ok]])
        cA = cc[start[ss[a_idx]]]
        cB = np.where(b_ok, cc[start[ss[np.minimum(b_idx, len(ss) - 1)]]], NONE)
        oB = np.where(b_ok, sout[ss[np.minimum(b_idx, len(ss) - 1)]], -1)
        w2 = cA | (cB << 14) | (1 << 31)
        # after every multi-task block (T = 1 needs no alignment)
        recs.append((gs[a_idx], task_sort_key(21, a_idx.astype(np.int64), 0), sout[ss[a_idx]], oB, w2))
    gg = np.concatenate([x[0] for x in recs])
    sk = np.concatenate([x[1] for x in recs])
    oa = np.concatenate([x[2] for x in recs])
    ob = np.concatenate([x[3] for x in recs])
    w2 = np.concatenate([x[4] for x in recs])
    o = np.lexsort((sk, gg))
    w0 = oa[o] | (np.where(ob[o] < 0, 0xFFFF, ob[o]) << 16)
    tasks = np.stack([w0, w2[o]], axis=1).astype(np.int64)
    tasks = tasks.astype(np.uint32).view(np.int32)
    grp_task_ptr = np.concatenate([[0], np.cumsum(np.bincount(gg, minlength=G))]).astype(np.int32)
    return grp_task_ptr, np.ascontiguousarray(tasks)


def build_tasks_torch(N, elem, sm_node, re_node, ne_ptr, ne_list, groups, device="cpu"):
    """build_tasks with torch on ``device`` (the GPU when there is one): the
    same task lists, element for element, from sorts and searches that run in
    milliseconds on the device instead of a minute of host lexsort on a
    million-node mesh.  Inputs and outputs are numpy arrays as for
    build_tasks."""
    import torch
    d = torch.device(device)
    i64 = torch.int64

    def t(a):
        return torch.as_tensor(np.ascontiguousarray(a)).to(d, i64)

    def first_flags(x):
        f = torch.ones_like(x, dtype=torch.bool)
        if x.numel() > 1:
            f[1:] = x[1:] != x[:-1]
        return f

    ne_ptr_t, elem_t, sm_t = t(ne_ptr), t(elem), t(sm_node)
    G = int(groups["num_groups"])
    grp_row_ptr, grp_rows = t(groups["grp_row_ptr"]), t(groups["grp_rows"])
    nrow = grp_row_ptr[1:] - grp_row_ptr[:-1]
    if int(nrow.max()) > 255 if nrow.numel() else False:
        raise ValueError("row groups of more than 255 rows")
    gid = torch.repeat_interleave(torch.arange(G, device=d), nrow)
    g_of_node = torch.empty(N, dtype=i64, device=d)
    r_of_node = torch.empty(N, dtype=i64, device=d)
    g_of_node[grp_rows] = gid
    r_of_node[grp_rows] = torch.arange(N, device=d) - grp_row_ptr[gid]
    inc_node = torch.repeat_interleave(torch.arange(N, device=d), ne_ptr_t[1:] - ne_ptr_t[:-1])
    ne_list_t = t(ne_list)
    inc_e, inc_j = ne_list_t >> 3, ne_list_t & 7
    inc_le = t(groups["ne_local"])
    if inc_le.numel() and int(inc_le.max()) > 254:
        raise ValueError("more than 255 elements in a row group")
    I = inc_node.numel()
    node = torch.repeat_interleave(inc_node, 8)
    k = torch.arange(8, device=d).repeat(I)
    e = torch.repeat_interleave(inc_e, 8)
    col = elem_t[e, k]
    rows = torch.repeat_interleave(torch.arange(N, device=d), sm_t[1:] - sm_t[:-1])
    keys = rows * N + t(re_node)
    slot = torch.searchsorted(keys, node * N + col)
    m = slot - sm_t[node]
    if m.numel() and int(m.max()) > 255:
        raise ValueError("a row with more than 256 neighbour blocks")
    g = g_of_node[node]
    r = r_of_node[node]
    cc = (torch.repeat_interleave(inc_le, 8) << 6) | (torch.repeat_interleave(inc_j, 8) << 3) | k
    del node, col, slot, rows, keys
    # lexsort by (g, r, m), then e: a stable sort on e, then a stable sort on the (g, r, m) key
    o1 = torch.sort(e, stable=True).indices
    o2 = torch.sort(((g << 16) | (r << 8) | m)[o1], stable=True).indices
    order = o1[o2]
    g, r, m, cc = g[order], r[order], m[order], cc[order]
    del e, order, o1, o2
    out_key = (r << 8) | m
    skey = g * 65536 + out_key
    start = torch.nonzero(first_flags(skey)).squeeze(1)
    n = torch.diff(torch.cat([start, torch.tensor([skey.numel()], device=d)]))
    if n.numel() and int(n.max()) > 64:
        raise ValueError("a node pair shared by more than 64 elements")
    half = torch.clamp((n + 1) // 2, min=1)
    T = torch.ones_like(half)
    while bool((T < half).any()):
        T = torch.where(T < half, T * 2, T)
    sg, sout = g[start], out_key[start]
    NONE = 0x3FFF
    recs = []
    multi = n >= 2
    ms = torch.nonzero(multi).squeeze(1)
    if ms.numel():
        Tm = T[ms]
        tb = torch.repeat_interleave(ms, Tm)
        q = torch.arange(int(Tm.sum()), device=d) - torch.repeat_interleave(torch.cumsum(Tm, 0) - Tm, Tm)
        i1 = start[tb] + 2 * q
        i2 = i1 + 1
        has1 = 2 * q < n[tb]
        has2 = 2 * q + 1 < n[tb]
        last = cc.numel() - 1
        c1 = torch.where(has1, cc[torch.clamp(i1, max=last)], torch.full_like(i1, NONE))
        c2 = torch.where(has2, cc[torch.clamp(i2, max=last)], torch.full_like(i2, NONE))
        lt = torch.round(torch.log2(T[tb].double())).to(i64)
        w2 = c1 | (c2 << 14) | (lt << 28)
        recs.append((sg[tb], task_sort_key(20 - lt, tb, q), sout[tb], torch.full_like(tb, -1), w2))
    ss = torch.nonzero(~multi).squeeze(1)
    if ss.numel():
        gs = sg[ss]
        o = torch.sort(gs, stable=True).indices
        ss, gs = ss[o], gs[o]
        ar = torch.arange(ss.numel(), device=d)
        pos = ar - torch.cummax(torch.where(first_flags(gs), ar, torch.zeros_like(ar)), 0).values
        a_idx = torch.nonzero(pos % 2 == 0).squeeze(1)
        b_idx = a_idx + 1
        lastb = ss.numel() - 1
        b_ok = (b_idx <= lastb) & (gs[torch.clamp(b_idx, max=lastb)] == gs[a_idx])
        cA = cc[start[ss[a_idx]]]
        cB = torch.where(b_ok, cc[start[ss[torch.clamp(b_idx, max=lastb)]]], torch.full_like(a_idx, NONE))
        oB = torch.where(b_ok, sout[ss[torch.clamp(b_idx, max=lastb)]], torch.full_like(a_idx, -1))
        w2 = cA | (cB << 14) | (1 << 31)
        recs.append((gs[a_idx], task_sort_key(21, a_idx, 0), sout[ss[a_idx]], oB, w2))
    gg = torch.cat([x[0] for x in recs])
    sk = torch.cat([x[1] for x in recs])
    oa = torch.cat([x[2] for x in recs])
    ob = torch.cat([x[3] for x in recs])
    w2 = torch.cat([x[4] for x in recs])
    o1 = torch.sort(sk, stable=True).indices
    o = o1[torch.sort(gg[o1], stable=True).indices]
    w0 = oa[o] | (torch.where(ob[o] < 0, torch.full_like(ob, 0xFFFF)[o], ob[o]) << 16)
    tasks = torch.stack([w0, w2[o]], dim=1) & 0xFFFFFFFF
    tasks = torch.where(tasks >= (1 << 31), tasks - (1 << 32), tasks).to(torch.int32)
    grp_task_ptr = torch.cat([torch.zeros(1, dtype=i64, device=d),
                              torch.cumsum(torch.bincount(gg, minlength=G), 0)]).to(torch.int32)
    return grp_task_ptr.cpu().numpy(), np.ascontiguousarray(tasks.cpu().numpy())
