"""GPU: the structured-grid 16-column kernels (fem_bsr_spmm16_grid,
fem_mg_smooth16_grid) against scipy (FP64 K X) and against the
pattern-agnostic kernels they stand in for (fem_bsr_spmm,
fem_mg_smooth_multi), on grids with partial 8 x 4 tiles, single-node-thick
axes and several z segments.  Bar: FP64 within 1e-14 relative of scipy's
product; FP32 passes within 1e-6 of the FP32 pattern-agnostic pass (the sums
run in another order); repeat calls bit-identical."""
import numpy as np
import pytest
import scipy.sparse as sp

from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu

GRIDS = [(1, 1, 1), (7, 3, 2), (8, 4, 5), (13, 9, 6), (17, 5, 40), (3, 11, 2), (30, 20, 25)]


def build(dims, seed=3):
    from fem_taichi_amd import Model, Physics
    g = mesh.structured_block(*dims, 1.0, 0.1, seed, E_sigma=0.3)
    phys = Physics.from_grid(g)
    assert phys.dev.grid_nodes == tuple(d + 1 for d in dims)
    return g, phys, Model(g.name, phys, g.Materials)


@pytest.mark.parametrize("dims", GRIDS)
def test_spmm16_grid_matches_scipy(gpu, dims):
    import torch
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    print(f"\nspmm16 grid {dims}", flush=True)
    g, phys, model = build(dims)
    N = phys.Num_Node
    k = model.assemble_K("f64", "bsr", penalty=False)
    rng = np.random.default_rng(sum(dims))
    X = rng.standard_normal((3 * N, 16))
    Xd = torch.as_tensor(X).cuda()
    Y = torch.full_like(Xd, np.nan)
    nx, ny, nz = phys.dev.grid_nodes
    _lib.call("fem_bsr_spmm16_grid", phys.dev.ref, nx, ny, nz, k.data_ptr(), Xd.data_ptr(), Y.data_ptr(),
              stream_ptr())
    Kh = sp.bsr_matrix((k.cpu().numpy().reshape(-1, 3, 3), phys.ReNode, phys.SmNode), shape=(3 * N, 3 * N))
    ref = Kh @ X
    got = Y.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref) * 10
    Y2 = torch.empty_like(Xd)
    _lib.call("fem_bsr_spmm", phys.dev.ref, k.data_ptr(), Xd.data_ptr(), Y2.data_ptr(), 16, stream_ptr())
    assert np.linalg.norm(got - Y2.cpu().numpy()) <= 1e-14 * np.linalg.norm(ref) * 10
    Y3 = torch.full_like(Xd, np.nan)
    _lib.call("fem_bsr_spmm16_grid", phys.dev.ref, nx, ny, nz, k.data_ptr(), Xd.data_ptr(), Y3.data_ptr(),
              stream_ptr())
    assert torch.equal(Y, Y3)


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("mode", [0, 1])
def test_smooth16_grid_matches_multi(gpu, dims, mode):
    import torch
    from fem_taichi_amd import _lib, solvers
    from fem_taichi_amd.runtime import stream_ptr
    print(f"\nsmooth16 grid {dims} mode {mode}", flush=True)
    g, phys, model = build(dims, seed=5)
    N = phys.Num_Node
    k32 = model.assemble_K("f32", "bsr")
    dinv, _ = solvers.BsrOperator(phys.dev, k32).block_jacobi()
    rng = np.random.default_rng(7 + mode)
    X = torch.as_tensor(rng.standard_normal((3 * N, 16)).astype(np.float32)).cuda()
    B = torch.as_tensor(rng.standard_normal((3 * N, 16)).astype(np.float32)).cuda()
    out = torch.full_like(X, np.nan)
    ref = torch.full_like(X, np.nan)
    nx, ny, nz = phys.dev.grid_nodes
    w = 0.6
    _lib.call("fem_mg_smooth16_grid", phys.dev.ref, nx, ny, nz, k32.data_ptr(), dinv.data_ptr(), w, X.data_ptr(),
              B.data_ptr(), out.data_ptr(), mode, stream_ptr())
    _lib.call("fem_mg_smooth_multi", phys.dev.ref, k32.data_ptr(), dinv.data_ptr(), w, X.data_ptr(), B.data_ptr(),
              ref.data_ptr(), 16, mode, stream_ptr())
    o, r = out.double().cpu().numpy(), ref.double().cpu().numpy()
    assert np.all(np.isfinite(o))
    assert np.linalg.norm(o - r) <= 1e-6 * np.linalg.norm(r)
    out2 = torch.full_like(X, np.nan)
    _lib.call("fem_mg_smooth16_grid", phys.dev.ref, nx, ny, nz, k32.data_ptr(), dinv.data_ptr(), w, X.data_ptr(),
              B.data_ptr(), out2.data_ptr(), mode, stream_ptr())
    assert torch.equal(out, out2)


def test_grid_kernels_reject_mismatched_grid(gpu):
    import torch
    from fem_taichi_amd import _lib
    g, phys, model = build((4, 3, 2))
    k = model.assemble_K("f64", "bsr")
    X = torch.zeros(3 * phys.Num_Node, 16, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.FemError):
        _lib.call("fem_bsr_spmm16_grid", phys.dev.ref, 5, 4, 2, k.data_ptr(), X.data_ptr(), X.clone().data_ptr(),
                  None)


def test_permuted_mesh_keeps_the_general_kernels(gpu):
    """A mesh whose node ids are not the grid numbering gets no grid dims, so
    the solvers keep the pattern-agnostic kernels."""
    from fem_taichi_amd import Physics
    g = mesh.structured_block(3, 3, 3)
    node, elem = mesh.permute_ids(g, seed=1)
    phys = Physics(node, elem, node_dims=g.node_dims)
    assert phys.dev.grid_nodes is None


@pytest.mark.parametrize("dims", GRIDS)
def test_mass16_grid_matches_scipy(gpu, dims):
    import torch
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    g, phys, model = build(dims, seed=9)
    N = phys.Num_Node
    mb = model.M_blocks()
    rng = np.random.default_rng(11 + sum(dims))
    X = rng.standard_normal((3 * N, 16))
    Xd = torch.as_tensor(X).cuda()
    Y = torch.full_like(Xd, np.nan)
    nx, ny, nz = phys.dev.grid_nodes
    _lib.call("fem_mass_spmm16_grid", phys.dev.ref, nx, ny, nz, mb.data_ptr(), Xd.data_ptr(), Y.data_ptr(),
              stream_ptr())
    Ms = sp.csr_matrix((mb.cpu().numpy(), phys.ReNode, phys.SmNode), shape=(N, N))
    ref = (Ms @ X.reshape(N, 48)).reshape(3 * N, 16)          # the scalar block on each DOF diagonal
    assert np.linalg.norm(Y.cpu().numpy() - ref) <= 1e-13 * np.linalg.norm(ref)
    Y2 = torch.empty_like(Xd)
    _lib.call("fem_mass_spmm", phys.dev.ref, mb.data_ptr(), Xd.data_ptr(), Y2.data_ptr(), 16, stream_ptr())
    assert np.linalg.norm(Y.cpu().numpy() - Y2.cpu().numpy()) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("dims", GRIDS)
def test_spmm16_grid_pq_matches_product(gpu, dims):
    """fem_bsr_spmm16_grid_pq: Y bit-identical to fem_bsr_spmm16_grid's, and
    the per-work-group partials summing, per column, to X_c . Y_c within
    1e-13 of numpy's; repeat calls bit-identical; a partial buffer shorter
    than 16 x fem_spmm16_grid_partials is refused."""
    import torch
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    print(f"\nspmm16 grid pq {dims}", flush=True)
    g, phys, model = build(dims, seed=13)
    N = phys.Num_Node
    k = model.assemble_K("f64", "bsr", penalty=False)
    rng = np.random.default_rng(17 + sum(dims))
    Xd = torch.as_tensor(rng.standard_normal((3 * N, 16))).cuda()
    nx, ny, nz = phys.dev.grid_nodes
    nblk = int(_lib.lib().fem_spmm16_grid_partials(nx, ny, nz))
    assert nblk >= 1
    part = torch.full((nblk * 16,), np.nan, dtype=torch.float64, device="cuda")
    Y = torch.full_like(Xd, np.nan)
    _lib.call("fem_bsr_spmm16_grid_pq", phys.dev.ref, nx, ny, nz, k.data_ptr(), Xd.data_ptr(), Y.data_ptr(),
              part.data_ptr(), part.numel(), stream_ptr())
    Y0 = torch.full_like(Xd, np.nan)
    _lib.call("fem_bsr_spmm16_grid", phys.dev.ref, nx, ny, nz, k.data_ptr(), Xd.data_ptr(), Y0.data_ptr(),
              stream_ptr())
    assert torch.equal(Y, Y0)
    ph = part.cpu().numpy().reshape(nblk, 16)
    assert np.all(np.isfinite(ph))
    X, Yh = Xd.cpu().numpy(), Y.cpu().numpy()
    ref = (X * Yh).sum(0)
    scale = (np.abs(X) * np.abs(Yh)).sum(0)
    assert np.all(np.abs(ph.sum(0) - ref) <= 1e-13 * scale)
    part2 = torch.full_like(part, np.nan)
    _lib.call("fem_bsr_spmm16_grid_pq", phys.dev.ref, nx, ny, nz, k.data_ptr(), Xd.data_ptr(), Y.data_ptr(),
              part2.data_ptr(), part2.numel(), stream_ptr())
    assert torch.equal(part, part2)
    with pytest.raises(_lib.FemError):
        _lib.call("fem_bsr_spmm16_grid_pq", phys.dev.ref, nx, ny, nz, k.data_ptr(), Xd.data_ptr(), Y.data_ptr(),
                  part2.data_ptr(), part2.numel() - 1, stream_ptr())
