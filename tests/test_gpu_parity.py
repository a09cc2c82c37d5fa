"""GPU parity of the HIP path against the CPU oracle (oracle/fem_oracle.py).

Tolerances (the FP64 bar): K, M, F within 1e-12 relative Frobenius with the
1e20 penalty entries compared exactly; stresses within 1e-10; integer
pattern arrays bit-exact.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu

K_TOL = 1e-12
S_TOL = 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def ansys_mesh2():
    """The reference's own Mesh/2.cdb (unstructured hex, gapped node IDs),
    root sets fixed, steel, omega = (0, 0, 100): Morton row groups."""
    from types import SimpleNamespace
    from fixtures import STEEL, load_mesh2
    Node, Elem, Bounds, Forces = load_mesh2()
    return SimpleNamespace(Node=Node, Elem=Elem, Bounds=Bounds, Forces=Forces, Materials=dict(STEEL),
                           EX=STEEL["EX"], PO=STEEL["PO"], DE=STEEL["DE"], Omegas=(0, 0, 100),
                           name="ansys_mesh2", node_dims=None)


def crowded():
    """Two hexes duplicated 20 and 26 times (per-copy modulus): node pairs
    with 20-46 contributions drive the 16- and 32-lane block sums."""
    from types import SimpleNamespace
    g = mesh.structured_block(2, 1, 1, 1.0, 0.1, 5)
    elem = np.concatenate([np.repeat(g.Elem[-8:, :1], 20, axis=1), np.repeat(g.Elem[-8:, 1:2], 26, axis=1)],
                          axis=1)
    E = 2.0e5 * np.exp(0.3 * np.random.default_rng(4).standard_normal(elem.shape[1]))
    mats = {"EX": E, "PO": 0.3, "DE": 7.8e-9}
    return SimpleNamespace(Node=g.Node, Elem=elem, Bounds=g.Bounds, Forces=g.Forces, Materials=mats, EX=E,
                           PO=0.3, DE=7.8e-9, Omegas=(0, 0, 100), name="crowded", node_dims=None)


def cases():
    return {
        "single_hex": mesh.structured_block(1, 1, 1, 1.0, 0.0, 0),   # one group, one element, z = 0 face fixed
        "box_3x4x5": mesh.structured_block(3, 4, 5, 1.0, 0.0, 0),
        "jit_6x5x7_lognormal": mesh.structured_block(6, 5, 7, 1.0, 0.10, 1, E_sigma=0.3),
        "twisted_10x4x4": mesh.twisted_blade(10, 4, 4, seed=3),
        "ansys_mesh2": ansys_mesh2(),
        "crowded_46": crowded(),
    }


@pytest.fixture(scope="module", params=list(cases().keys()))
def setup(request, gpu):
    from fem_taichi_amd import Model, Physics
    g = cases()[request.param]
    print(f"\ncase {request.param}", flush=True)
    phys = Physics.from_grid(g)
    pp = O.pre_process(g.Node, g.Elem)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    return g, phys, pp, model


def test_pattern_bit_exact(setup):
    g, phys, pp, _ = setup
    for k in ("SmNode", "ReNode", "CoNode", "CoElem"):
        assert np.array_equal(getattr(phys, k).astype(np.int64), pp[k].astype(np.int64)), k


def test_K_csr(setup):
    g, phys, pp, model = setup
    K = model.Get_MatrixKL()
    Kref = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    assert K.indices.dtype == np.int32 and K.indptr.dtype == np.int32
    assert np.array_equal(K.indptr, Kref.indptr)
    assert np.array_equal(K.indices, Kref.indices)
    pen = Kref.data >= 1e19
    assert np.array_equal(K.data[pen], Kref.data[pen])
    assert rel(K.data[~pen], Kref.data[~pen]) < K_TOL


def test_K_bsr_matches_csr(setup):
    g, phys, pp, model = setup
    vals = model.assemble_K("f64", "bsr").cpu().numpy().reshape(-1, 3, 3)
    Vref = O.model_K_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    pen = Vref >= 1e19
    assert np.array_equal(vals[pen], Vref[pen])
    assert rel(vals[~pen], Vref[~pen]) < K_TOL
    n = 3 * phys.Num_Node
    B = sp.bsr_matrix((vals, phys.ReNode, phys.SmNode), shape=(n, n)).tocsr()
    assert (B != model.Get_MatrixKL()).nnz == 0


def test_K_deterministic(setup):
    _, _, _, model = setup
    a = model.assemble_K("f64", "bsr").cpu().numpy()
    b = model.assemble_K("f64", "bsr").cpu().numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_K_f32(setup):
    g, phys, pp, model = setup
    v32 = model.assemble_K("f32", "bsr").cpu().numpy().astype(np.float64).reshape(-1, 3, 3)
    Vref = O.model_K_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    free = Vref < 1e19
    assert rel(v32[free], Vref[free]) < 1e-6


def test_M_csr(setup):
    g, phys, pp, model = setup
    M = model.Get_MatrixM()
    Mref = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    assert np.array_equal(M.indptr, Mref.indptr) and np.array_equal(M.indices, Mref.indices)
    assert rel(M.data, Mref.data) < K_TOL


def test_F(setup):
    g, phys, pp, model = setup
    F = model.Get_VectorF()
    Fref = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    assert rel(F, Fref) < K_TOL


def test_spmv(setup):
    import torch
    from fem_taichi_amd import solvers
    g, phys, pp, model = setup
    K = model.Get_MatrixKL()
    rng = np.random.default_rng(0)
    x = rng.standard_normal(K.shape[0])
    op = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    y = op.matvec(torch.as_tensor(x).cuda()).cpu().numpy()
    assert rel(y, K @ x) < 1e-14
    op32 = solvers.BsrOperator(model.dev, model.K_bsr("f32"))
    y32 = op32.matvec(torch.as_tensor(x).float().cuda()).double().cpu().numpy()
    assert rel(y32, K @ x) < 1e-5


def test_static_pcg_and_stress(setup):
    from fem_taichi_amd import Post, Static_solver
    g, phys, pp, model = setup
    solver = Static_solver(model, 0)
    U = solver.Solve_cpu(tol=1e-10)
    Kref = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    Fref = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    assert O.static_residual(Kref, U, Fref) < 1e-9
    S = Post(g.name, phys, g.Materials, U.reshape(-1, 3, 1)).Get_Stress()
    Sref = O.get_stress(g.Node, g.Elem, pp, g.EX, g.PO, U)
    assert rel(S[:, :6], Sref[:, :6]) < S_TOL
    assert rel(S[:, 6:9], Sref[:, 6:9]) < S_TOL
    assert rel(S[:, 9], Sref[:, 9]) < 1e-8


def test_static_mixed(setup):
    from fem_taichi_amd import Static_solver
    g, phys, pp, model = setup
    solver = Static_solver(model, 0)
    U = solver.Solve_cpu(method="mixed", tol=1e-10)
    Kref = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    Fref = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    assert O.static_residual(Kref, U, Fref) < 1e-9


@pytest.mark.gpu
def test_static_mixed_fallback(setup):
    """patience=0 declares the FP32-matrix phase stalled at its first FP64
    residual replacement; the FP64 PCG that finishes the solve must still
    reach the tolerance on the oracle's K and F."""
    from fem_taichi_amd import solvers
    g, phys, pp, model = setup
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    K32 = solvers.BsrOperator(model.dev, model.K_bsr("f32"))
    U, info = solvers.solve_mixed(K64, K32, model.F_device(), tol=1e-10, patience=0)
    assert info.outer == 2 and info.history[0]["stalled"]
    Kref = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    Fref = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    assert O.static_residual(Kref, U.cpu().numpy(), Fref) < 1e-9


@pytest.mark.parametrize("nrhs", [1, 3, 8, 10, 16, 64])
def test_spmm_multi(setup, nrhs):
    """fem_bsr_spmm (K staged in LDS for 16 columns, two columns per thread for
    other even counts from 8 on, thread per (row, column) otherwise) and
    fem_mass_spmm against scipy."""
    import torch
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    g, phys, pp, model = setup
    K = model.Get_MatrixKL()
    M = model.Get_MatrixM()
    n = K.shape[0]
    X = np.random.default_rng(nrhs).standard_normal((n, nrhs))
    Xd = torch.as_tensor(X).cuda().contiguous()
    Y = torch.empty_like(Xd)
    _lib.call("fem_bsr_spmm", model.dev.ref, model.K_bsr("f64").data_ptr(), Xd.data_ptr(), Y.data_ptr(), nrhs,
              stream_ptr())
    assert rel(Y.cpu().numpy(), K @ X) < 1e-14
    if nrhs == 16:
        # the 16-column kernel (K staged in LDS) sums every column like the
        # two-column kernel does for 8: bit-identical to the two halves
        for h in range(2):
            Xh = Xd[:, 8 * h:8 * h + 8].contiguous()
            Yh = torch.empty_like(Xh)
            _lib.call("fem_bsr_spmm", model.dev.ref, model.K_bsr("f64").data_ptr(), Xh.data_ptr(), Yh.data_ptr(), 8,
                      stream_ptr())
            assert torch.equal(Yh, Y[:, 8 * h:8 * h + 8])
    _lib.call("fem_mass_spmm", model.dev.ref, model.M_blocks().data_ptr(), Xd.data_ptr(), Y.data_ptr(), nrhs,
              stream_ptr())
    assert rel(Y.cpu().numpy(), M @ X) < 1e-14


def test_tasks_on_device_equal_host(gpu):
    """The assembly task lists built on the GPU (groups.build_tasks_torch, what
    Physics uses) equal the host numpy build on a 24x20x18 jittered slab."""
    from fem_taichi_amd import Physics, groups as Gr
    g = mesh.structured_block(24, 20, 18, 1.0, 0.1, 3, E_sigma=0.3)
    phys = Physics.from_grid(g)
    pp = O.pre_process(g.Node, g.Elem)
    en = O.element_nodes(g.Elem, pp["Nind"])
    N = g.Node.shape[1]
    e = np.repeat(np.arange(len(en)), 8)
    j = np.tile(np.arange(8), len(en))
    node = en.reshape(-1)
    o = np.lexsort((e * 8 + j, node))
    ne_ptr = np.concatenate([[0], np.cumsum(np.bincount(node, minlength=N))])
    ne_list = (e * 8 + j)[o]
    gid = Gr.node_group_ids(N, node_dims=(25, 21, 19))
    groups = Gr.build_row_groups(N, len(en), ne_ptr, ne_list, gid, 48, co_node=pp["CoNode"])
    a = Gr.build_tasks(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups)
    b = Gr.build_tasks_torch(N, en, pp["SmNode"], pp["ReNode"], ne_ptr, ne_list, groups, device="cuda")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert phys.dev.groups_info["num_groups"] == groups["num_groups"]


def test_pattern_input_checks(gpu):
    """The pattern builder's own checks turn bad connectivity into errors
    through the C ABI (a node index outside [0, N): FEM_E_INVALID; a node
    with more than 128 distinct neighbours: FEM_E_LIMIT), and an empty
    element table gives an empty pattern."""
    import torch
    from fem_taichi_amd import _lib
    from fem_taichi_amd.physics import gpu_pattern
    fan = np.zeros((20, 8), np.int32)            # 20 hexes around node 0: 141 neighbours
    fan[:, 1:] = 1 + np.arange(140).reshape(20, 7)
    with pytest.raises(_lib.FemError, match="neighbour") as ei:
        gpu_pattern(torch.as_tensor(fan).cuda(), 141)
    assert ei.value.code == _lib.FEM_E_LIMIT
    ok = gpu_pattern(torch.as_tensor(fan[:18]).cuda(), 127)   # 127 neighbours: within the limit
    nbrs = [set() for _ in range(127)]
    for row in fan[:18]:
        for a in row:
            nbrs[a].update(row.tolist())
    assert ok["co_node"].cpu().tolist() == [len(s) for s in nbrs]
    assert ok["re_node"][:127].cpu().tolist() == list(range(127))
    bad = fan[:2].copy()
    bad[1, 3] = 500
    with pytest.raises(_lib.FemError, match="outside"):
        gpu_pattern(torch.as_tensor(bad).cuda(), 141)
    bad[1, 3] = -1
    with pytest.raises(_lib.FemError, match="outside"):
        gpu_pattern(torch.as_tensor(bad).cuda(), 141)
    empty = gpu_pattern(torch.zeros((0, 8), dtype=torch.int32).cuda(), 5)
    assert empty["co_node"].cpu().tolist() == [0] * 5 and int(empty["sm_node"][5].item()) == 0
