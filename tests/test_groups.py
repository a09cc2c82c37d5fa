"""Host logic of the row-owner assembly: row groups and contribution tasks.

The assembly kernel's phase 2 is emulated here lane by lane (64-lane waves,
the T-lane shuffle butterfly with its wave-uniform early exit, head-lane
stores), with the oracle's element blocks standing in for the Gram blocks the
kernel builds in LDS.  Every output block must be written exactly once and
equal the oracle's assembled K (and M).  CPU only.
"""
import numpy as np
import pytest

from oracle import fem_oracle as O
from fem_taichi_amd import groups as Gr
from fem_taichi_amd import mesh

NONE = 0x3FFF


def incidences(en, N):
    """ne_ptr / ne_list (e << 3 | j, sorted per node) of an (E, 8) element table."""
    e = np.repeat(np.arange(len(en)), 8)
    j = np.tile(np.arange(8), len(en))
    node = en.reshape(-1)
    o = np.lexsort((e * 8 + j, node))
    ne_ptr = np.concatenate([[0], np.cumsum(np.bincount(node, minlength=N))])
    return ne_ptr, (e * 8 + j)[o]


def emulate(tasks, tp, groups, blocks_e, SmNode, nslots, wave=64, threads=512, co_node=None):
    """Phase 2 of the kernel on element blocks ``blocks_e`` (E, 8, 8, ...).

    With ``co_node`` (blocks per row) a single-pass group with at most 64 rows
    takes the staged path: head lanes write their block to the staging copy
    at (prefix of the group's row block counts)[r] + m, then each row's run is
    copied out to its BSR slots; every staging entry must be written once."""
    out = np.zeros((nslots,) + blocks_e.shape[3:])
    hits = np.zeros(nslots, dtype=np.int64)
    gep, ge = groups["grp_elem_ptr"], groups["grp_elems"]
    grp, grows = groups["grp_row_ptr"], groups["grp_rows"]
    tw = tasks.view(np.uint32)
    zero = np.zeros(blocks_e.shape[3:])

    def contrib(g, c):
        if c == NONE:
            return zero
        le, j, k = c >> 6, (c >> 3) & 7, c & 7
        assert le < gep[g + 1] - gep[g]
        return blocks_e[ge[gep[g] + le], j, k]

    for g in range(len(tp) - 1):
        t0, nt = tp[g], tp[g + 1] - tp[g]
        rows = grows[grp[g]:grp[g + 1]]
        staged = co_node is not None and nt <= threads and len(rows) <= 64
        if staged:
            co = co_node[rows]
            pre = np.concatenate([[0], np.cumsum(co)[:-1]])
            stage = np.zeros((int(co.sum()),) + blocks_e.shape[3:])
            stage_hits = np.zeros(int(co.sum()), dtype=np.int64)
        for base in range(0, nt, threads):
            for w0 in range(0, threads, wave):
                lanes = range(w0, w0 + wave)
                val, T, outA, outB, dual, c2v, valid = [], [], [], [], [], [], []
                for tid in lanes:
                    t = base + tid
                    ok = t < nt
                    rec = tw[t0 + t] if ok else np.array([0xFFFF0000, NONE | NONE << 14], dtype=np.uint32)
                    w2 = int(rec[1])
                    c1, c2 = w2 & 0x3FFF, (w2 >> 14) & 0x3FFF
                    d = bool(w2 >> 31)
                    Tn = 1 if d else 1 << ((w2 >> 28) & 7)
                    v1, v2 = contrib(g, c1), contrib(g, c2)
                    val.append(v1 if d else v1 + v2)
                    c2v.append(v2)
                    T.append(Tn)
                    outA.append(int(rec[0]) & 0xFFFF)
                    ob = int(rec[0]) >> 16
                    outB.append(-1 if ob == 0xFFFF else ob)
                    dual.append(d)
                    valid.append(ok)
                # the kernel's steps: lane ^1, ^2 (quad_perm), ^7 (row_half_mirror),
                # ^15 (row_mirror), ^16; step k runs for lanes with T > 2^k and
                # is skipped when no lane of the wave needs it
                for s, xor in ((1, 1), (2, 2), (4, 7), (8, 15), (16, 16)):
                    if not any(s < x for x in T):
                        continue
                    new = list(val)
                    for li in range(wave):
                        if s < T[li]:
                            new[li] = val[li] + val[li ^ xor]
                    val = new
                for li in range(wave):
                    if not valid[li] or (li & (T[li] - 1)) != 0:
                        continue
                    for o, v in [(outA[li], val[li])] + ([(outB[li], c2v[li])] if dual[li] and outB[li] >= 0 else []):
                        r, m = o >> 8, o & 255
                        if staged:
                            stage[pre[r] + m] = v
                            stage_hits[pre[r] + m] += 1
                            continue
                        node = grows[grp[g] + r]
                        slot = SmNode[node] + m
                        out[slot] = v
                        hits[slot] += 1
        if staged:
            assert np.all(stage_hits == 1)
            for r, node in enumerate(rows):
                out[SmNode[node]:SmNode[node] + co[r]] = stage[pre[r]:pre[r] + co[r]]
                hits[SmNode[node]:SmNode[node] + co[r]] += 1
    return out, hits


def check_block_lanes(tasks, tp):
    """The T tasks of a block are adjacent, T-aligned within their group, share outA."""
    w2 = tasks[:, 1].view(np.uint32)
    outa = tasks[:, 0].view(np.uint32) & 0xFFFF
    for g in range(len(tp) - 1):
        t, end = tp[g], tp[g + 1]
        while t < end:
            d = bool(w2[t] >> 31)
            Tn = 1 if d else 1 << ((int(w2[t]) >> 28) & 7)
            assert (t - tp[g]) % Tn == 0
            assert np.all(outa[t:t + Tn] == outa[t])
            t += Tn


def crowded_mesh():
    """Two neighbouring hexes, duplicated 20 and 26 times with a lognormal
    modulus per copy: node pairs with 20, 26 and 46 contributions (T = 16, 16,
    32), the wide butterfly steps structured meshes never reach."""
    g = mesh.structured_block(2, 1, 1, 1.0, 0.1, 5)
    elem = np.concatenate([np.repeat(g.Elem[-8:, :1], 20, axis=1), np.repeat(g.Elem[-8:, 1:2], 26, axis=1)],
                          axis=1)
    E = 2.0e5 * np.exp(0.3 * np.random.default_rng(4).standard_normal(elem.shape[1]))
    return g, elem, E


CASES = [
    ("brick", dict(dims=(9, 5, 6), jitter=0.1, seed=3), True),
    ("morton", dict(dims=(7, 6, 5), jitter=0.15, seed=4), False),
    ("permuted", dict(dims=(5, 4, 7), jitter=0.1, seed=5), False),
]


@pytest.mark.parametrize("name,kw,brick", CASES, ids=[c[0] for c in CASES])
def test_tasks_assemble_like_oracle(name, kw, brick):
    g = mesh.structured_block(*kw["dims"], 1.0, kw["jitter"], kw["seed"], E_sigma=0.3)
    Node, Elem = (mesh.permute_ids(g, seed=kw["seed"]) if name == "permuted" else (g.Node, g.Elem))
    pp = O.pre_process(Node, Elem)
    en = O.element_nodes(Elem, pp["Nind"])
    N = Node.shape[1]
    coords = np.asarray(Node[-3:], dtype=np.float64).T
    ne_ptr, ne_list = incidences(en, N)
    if brick:
        nx, ny, nz = kw["dims"]
        gid = Gr.node_group_ids(N, node_dims=(nx + 1, ny + 1, nz + 1))
    else:
        gid = Gr.node_group_ids(N, xyz=coords)
    groups = Gr.build_row_groups(N, len(en), ne_ptr, ne_list, gid, 48, co_node=pp["CoNode"])
    assert groups["max_group_elems"] <= 48
    tp, tasks = Gr.build_tasks(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
    check_block_lanes(tasks, tp)

    Coord = O.element_coords(Node, Elem, pp["Nind"])
    EX = g.Materials["EX"]
    Ke = O.element_K(Coord, EX, g.Materials["PO"])
    nnzb = int(pp["SmNode"][-1])
    Kt, hits = emulate(tasks, tp, groups, Ke, pp["SmNode"], nnzb)
    assert np.all(hits == 1)
    Kref = O.assemble_K_bsr(Ke, en, pp["SmNode"], pp["ReNode"])
    err = np.linalg.norm(Kt.reshape(-1) - Kref.reshape(-1)) / np.linalg.norm(Kref)
    assert err < 1e-14, err
    Ks, hits = emulate(tasks, tp, groups, Ke, pp["SmNode"], nnzb, co_node=pp["CoNode"])
    assert np.all(hits == 1)
    assert np.array_equal(Ks, Kt)

    Me = O.element_M(Coord, g.Materials["DE"])
    Mt, hits = emulate(tasks, tp, groups, Me, pp["SmNode"], nnzb)
    assert np.all(hits == 1)
    Mref = O.assemble_M_block(Me, en, pp["SmNode"], pp["ReNode"])
    assert np.linalg.norm(Mt - Mref) / np.linalg.norm(Mref) < 1e-14


def test_tasks_reject_crowded_node_pair():
    # 65 elements on the same 8 nodes: a node pair with 65 contributions cannot
    # be summed by a 32-lane butterfly of pairs
    N = 8
    en = np.tile(np.arange(8, dtype=np.int64), (65, 1))
    rows = [set() for _ in range(N)]
    for nodes in en:
        for a in nodes:
            rows[a].update(int(b) for b in nodes)
    sm = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    re = np.concatenate([sorted(r) for r in rows])
    ne_ptr, ne_list = incidences(en, N)
    gid = np.zeros(N, dtype=np.int64)
    groups = Gr.build_row_groups(N, 65, ne_ptr, ne_list, gid, 200)
    with pytest.raises(ValueError, match="more than 64"):
        Gr.build_tasks(N, en, sm, re, ne_ptr, ne_list, groups)


def test_tasks_ansys_mesh2_morton():
    """The reference's Mesh/2.cdb through Morton row groups and task lists."""
    from fixtures import load_mesh2
    Node, Elem, _, _ = load_mesh2()
    pp = O.pre_process(Node, Elem)
    en = O.element_nodes(Elem, pp["Nind"])
    N = Node.shape[1]
    ne_ptr, ne_list = incidences(en, N)
    gid = Gr.node_group_ids(N, xyz=np.asarray(Node[-3:], dtype=np.float64).T)
    groups = Gr.build_row_groups(N, len(en), ne_ptr, ne_list, gid, 48, co_node=pp["CoNode"])
    tp, tasks = Gr.build_tasks(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
    assert np.diff(tp).max() <= 512
    check_block_lanes(tasks, tp)
    Coord = O.element_coords(Node, Elem, pp["Nind"])
    Ke = O.element_K(Coord, 2.0e5, 0.3)
    Kt, hits = emulate(tasks, tp, groups, Ke, pp["SmNode"], int(pp["SmNode"][-1]))
    assert np.all(hits == 1)
    Kref = O.assemble_K_bsr(Ke, en, pp["SmNode"], pp["ReNode"])
    assert np.linalg.norm(Kt.reshape(-1) - Kref.reshape(-1)) / np.linalg.norm(Kref) < 1e-14


def test_tasks_crowded_wide_butterfly():
    g, elem, E = crowded_mesh()
    Node = g.Node
    pp = O.pre_process(Node, elem)
    en = O.element_nodes(elem, pp["Nind"])
    N = Node.shape[1]
    ne_ptr, ne_list = incidences(en, N)
    gid = Gr.node_group_ids(N, xyz=np.asarray(Node[-3:], dtype=np.float64).T)
    groups = Gr.build_row_groups(N, len(en), ne_ptr, ne_list, gid, 48, co_node=pp["CoNode"])
    tp, tasks = Gr.build_tasks(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
    lt = (tasks[:, 1].view(np.uint32) >> 28) & 7
    assert lt.max() == 5 and 4 in set(lt.tolist())      # T = 32 and T = 16 blocks present
    check_block_lanes(tasks, tp)
    Ke = O.element_K(O.element_coords(Node, elem, pp["Nind"]), E, 0.3)
    Kt, hits = emulate(tasks, tp, groups, Ke, pp["SmNode"], int(pp["SmNode"][-1]))
    assert np.all(hits == 1)
    Kref = O.assemble_K_bsr(Ke, en, pp["SmNode"], pp["ReNode"])
    assert np.linalg.norm(Kt.reshape(-1) - Kref.reshape(-1)) / np.linalg.norm(Kref) < 1e-14


@pytest.mark.parametrize("name,kw,brick", CASES, ids=[c[0] for c in CASES])
def test_tasks_torch_equal_numpy(name, kw, brick):
    """build_tasks_torch (the device form used on GPU machines, run here on
    the CPU) gives the numpy build_tasks' lists exactly."""
    g = mesh.structured_block(*kw["dims"], 1.0, kw["jitter"], kw["seed"], E_sigma=0.3)
    Node, Elem = (mesh.permute_ids(g, seed=kw["seed"]) if name == "permuted" else (g.Node, g.Elem))
    pp = O.pre_process(Node, Elem)
    en = O.element_nodes(Elem, pp["Nind"])
    N = Node.shape[1]
    ne_ptr, ne_list = incidences(en, N)
    if brick:
        nx, ny, nz = kw["dims"]
        gid = Gr.node_group_ids(N, node_dims=(nx + 1, ny + 1, nz + 1))
    else:
        gid = Gr.node_group_ids(N, xyz=np.asarray(Node[-3:], dtype=np.float64).T)
    groups = Gr.build_row_groups(N, len(en), ne_ptr, ne_list, gid, 48, co_node=pp["CoNode"])
    a = Gr.build_tasks(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
    b = Gr.build_tasks_torch(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
    assert np.array_equal(a[0], b[0]) and a[1].dtype == b[1].dtype and np.array_equal(a[1], b[1])


def test_tasks_torch_equal_numpy_crowded_and_mesh2():
    from fixtures import load_mesh2
    g, elem, E = crowded_mesh()
    meshes = [(g.Node, elem)]
    Node, Elem, _, _ = load_mesh2()
    meshes.append((Node, Elem))
    for Node, Elem in meshes:
        pp = O.pre_process(Node, Elem)
        en = O.element_nodes(Elem, pp["Nind"])
        N = Node.shape[1]
        ne_ptr, ne_list = incidences(en, N)
        gid = Gr.node_group_ids(N, xyz=np.asarray(Node[-3:], dtype=np.float64).T)
        groups = Gr.build_row_groups(N, len(en), ne_ptr, ne_list, gid, 64, co_node=pp["CoNode"])
        a = Gr.build_tasks(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
        b = Gr.build_tasks_torch(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_task_sort_key_orders_by_class_first_at_any_block_index():
    """Task order inside a group is (class, block index, lane); block indices
    run over the whole mesh, beyond 2^24 blocks on million-node meshes, and
    must never spill into the class (a packed key that let them did
    misalign the T-lane blocks of a group straddling 2^24 blocks)."""
    rng = np.random.default_rng(0)
    n = 200000
    cls = rng.integers(15, 22, n)
    idx = rng.integers(0, 1 << 40, n)
    idx[:1000] = (1 << 24) + rng.integers(-8, 8, 1000)     # around the old spill point
    lane = rng.integers(0, 32, n)
    key = Gr.task_sort_key(cls, idx, lane)
    assert np.array_equal(np.argsort(key, kind="stable"), np.lexsort((lane, idx, cls)))
    import torch
    kt = Gr.task_sort_key(torch.as_tensor(cls), torch.as_tensor(idx), torch.as_tensor(lane))
    assert np.array_equal(kt.numpy(), key)
