"""Full-size checks on the assembly workload, the 60x60x55 slab (198,000
elements, 10 % jitter, lognormal E): properties that hold at any size, so they
need no oracle run at this size.

- BSR and the reference's CSR layout hold bit-identical values, and two
  assemblies are bit-identical;
- the unpenalised K annihilates the six rigid-body modes (translations and
  infinitesimal rotations: their discrete strain is exactly zero), is
  symmetric, and the penalised K is positive on random vectors;
- the FP32 K is within 1e-6 of the FP64 K;
- the consistent mass of the unjittered 60x60x55 box sums to rho * V, and
  its centrifugal load vector to rho |omega|^2 V (xc, yc, 0);
- an affine displacement u = G x (reproduced exactly by trilinear elements)
  gives each node the stress E_bar * (lhat tr(eps) I + 2 mhat eps), eps =
  sym(G), with E_bar the mean modulus of the node's elements.

And on the static-solve and modal workload, the twisted blade 400x50x50
(3.13M DOF): the multigrid solve's FP64 residual and its fixed root face;
the four lowest eigenvalues independent of the start block, M-orthogonal
Ritz vectors and small shift-invert residuals.
"""
import numpy as np
import pytest

from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slab(gpu):
    from fem_taichi_amd import Model, Physics
    g = mesh.slab_60x60x55()
    print("\ncase slab_60x60x55", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    return g, phys, model


def test_layouts_and_repeat_bit_identical(slab):
    import torch
    g, phys, model = slab
    bsr = model.assemble_K("f64", "bsr", penalty=False)
    csr = model.assemble_K("f64", "csr", penalty=False)
    again = model.assemble_K("f64", "bsr", penalty=False)
    assert torch.equal(bsr.view(torch.int64), again.view(torch.int64))
    d = bsr.device
    sm = torch.as_tensor(phys.SmNode.astype(np.int64), device=d)
    co = torch.as_tensor(phys.CoNode.astype(np.int64), device=d)
    nnzb = int(sm[-1])
    row = torch.repeat_interleave(torch.arange(phys.Num_Node, device=d), co)
    m = torch.arange(nnzb, device=d) - sm[row]
    p = torch.arange(3, device=d).view(1, 3, 1)
    q = torch.arange(3, device=d).view(1, 1, 3)
    idx = 9 * sm[row].view(-1, 1, 1) + 3 * p * co[row].view(-1, 1, 1) + 3 * m.view(-1, 1, 1) + q
    assert torch.equal(csr[idx.reshape(-1)].view(torch.int64), bsr.view(torch.int64))


def test_rigid_modes_symmetry_positivity(slab):
    import torch
    from fem_taichi_amd import solvers
    g, phys, model = slab
    K = solvers.BsrOperator(model.dev, model.assemble_K("f64", "bsr", penalty=False))
    d = K.vals.device
    x = torch.as_tensor(np.asarray(g.Node[-3:], dtype=np.float64).T, device=d)   # (N, 3)
    gen = torch.Generator(device=d)
    gen.manual_seed(3)
    a = torch.randn(x.numel(), dtype=torch.float64, device=d, generator=gen)
    b = torch.randn(x.numel(), dtype=torch.float64, device=d, generator=gen)
    Ka, Kb = K.matvec(a), K.matvec(b)
    scale = float(Ka.norm() / a.norm())          # a stand-in for ||K||
    modes = []
    for t in range(3):                            # translations
        u = torch.zeros_like(x)
        u[:, t] = 1.0
        modes.append(u)
    for w in torch.eye(3, dtype=torch.float64, device=d):    # rotations w x x about the slab centre
        modes.append(torch.cross(w.expand_as(x), x - x.mean(0), dim=1))
    for u in modes:
        r = K.matvec(u.reshape(-1).contiguous())
        assert float(r.norm() / u.norm()) < 1e-12 * scale
    asym = abs(float(a @ Kb) - float(b @ Ka))
    assert asym < 1e-12 * float(a.norm() * Kb.norm())
    Kp = solvers.BsrOperator(model.dev, model.assemble_K("f64", "bsr"))
    assert float(a @ Kp.matvec(a)) > 0 and float(b @ Kp.matvec(b)) > 0


def test_box_mass_sums_to_rho_volume(gpu):
    from fem_taichi_amd import Model, Physics
    g = mesh.structured_block(60, 60, 55, 1.0, 0.0, 1, E_sigma=0.3)
    model = Model(g.name, Physics.from_grid(g), g.Materials, Omegas=g.Omegas)
    total = float(model.M_blocks().sum())
    rho_v = g.Materials["DE"] * 60 * 60 * 55
    assert abs(total - rho_v) < 1e-12 * rho_v


def test_affine_displacement_stress(slab):
    from fem_taichi_amd import Post
    g, phys, model = slab
    xyz = np.asarray(g.Node[-3:], dtype=np.float64).T
    G = np.array([[2.0, 0.5, -0.3], [0.1, -1.0, 0.7], [0.4, -0.2, 1.5]]) * 1e-4
    U = xyz @ G.T                                  # u_a = G_ab x_b
    S = Post(g.name, phys, g.Materials, U.reshape(-1, 3, 1)).Get_Stress()
    nu = float(g.Materials["PO"])
    lhat, mhat = nu / ((1 + nu) * (1 - 2 * nu)), 1.0 / (2 * (1 + nu))
    eps = 0.5 * (G + G.T)
    s_unit = lhat * np.trace(eps) * np.eye(3) + 2 * mhat * eps
    en = np.asarray(g.Elem[-8:], dtype=np.int64).T       # node index = ID - 1 on these grids
    E = np.asarray(g.Materials["EX"], dtype=np.float64)
    N = xyz.shape[0]
    ebar = np.bincount(en.reshape(-1), weights=np.repeat(E, 8), minlength=N) / np.bincount(en.reshape(-1), minlength=N)
    comps = np.array([s_unit[0, 0], s_unit[1, 1], s_unit[2, 2], s_unit[0, 1], s_unit[1, 2], s_unit[0, 2]])
    expect = ebar[:, None] * comps[None, :]
    assert np.linalg.norm(S[:, :6] - expect) / np.linalg.norm(expect) < 1e-10
    ev = np.sort(np.linalg.eigvalsh(s_unit))[::-1]
    assert np.linalg.norm(S[:, 6:9] - ebar[:, None] * ev[None, :]) / np.linalg.norm(ebar[:, None] * ev) < 1e-10


@pytest.mark.parametrize("case", ["hydrostatic", "uniaxial", "shear", "rotation"])
def test_affine_stress_degenerate_principal(slab, case):
    """Principal stresses and the reference's von Mises term where the
    closed-form 3x3 solve meets repeated roots: a triple root (hydrostatic
    strain), a double root (uniaxial strain), roots (t, 0, -t) (pure shear)
    and a zero tensor (a rigid rotation, whose stresses are rounding
    noise)."""
    from fem_taichi_amd import Post
    g, phys, model = slab
    xyz = np.asarray(g.Node[-3:], dtype=np.float64).T
    G = {"hydrostatic": np.eye(3), "uniaxial": np.diag([1.0, 0.0, 0.0]),
         "shear": np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]]),
         "rotation": np.array([[0.0, -1.0, 0.5], [1.0, 0.0, -0.3], [-0.5, 0.3, 0.0]])}[case] * 1e-4
    S = Post(g.name, phys, g.Materials, (xyz @ G.T).reshape(-1, 3, 1)).Get_Stress()
    nu = float(g.Materials["PO"])
    lhat, mhat = nu / ((1 + nu) * (1 - 2 * nu)), 1.0 / (2 * (1 + nu))
    eps = 0.5 * (G + G.T)
    s_unit = lhat * np.trace(eps) * np.eye(3) + 2 * mhat * eps
    en = np.asarray(g.Elem[-8:], dtype=np.int64).T
    E = np.asarray(g.Materials["EX"], dtype=np.float64)
    N = xyz.shape[0]
    ebar = np.bincount(en.reshape(-1), weights=np.repeat(E, 8), minlength=N) / np.bincount(en.reshape(-1), minlength=N)
    if case == "rotation":
        assert np.max(np.abs(S)) < 1e-12 * E.max() * 1e-4
        return
    ev = np.sort(np.linalg.eigvalsh(s_unit))[::-1]
    expect = ebar[:, None] * ev[None, :]
    assert np.linalg.norm(S[:, 6:9] - expect) / np.linalg.norm(expect) < 1e-10
    vm = np.sqrt((ev[0] - ev[1]) ** 2 + (ev[1] - ev[2]) ** 2 + (ev[2] - ev[1]) ** 2)   # as written in the reference
    if vm > 0:
        assert np.linalg.norm(S[:, 9] - ebar * vm) / np.linalg.norm(ebar * vm) < 1e-8
    else:
        assert np.max(np.abs(S[:, 9])) < 1e-6 * np.max(np.abs(expect))


@pytest.fixture(scope="module")
def blade(gpu):
    """The static-solve and modal workload: twisted blade 400x50x50 (1.04M
    nodes, 3.13M DOF), root face fixed, omega = (0, 0, 100) rad/s."""
    from fem_taichi_amd import Model, Physics
    g = mesh.twisted_blade()
    print("\ncase twisted_blade_400x50x50", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    return g, phys, model


def _free_dofs(phys, kvals):
    """DOFs whose diagonal is not the 1e20 penalty."""
    import torch
    N = phys.Num_Node
    sm, re = phys.SmNode.astype(np.int64), phys.ReNode.astype(np.int64)
    keys = np.repeat(np.arange(N), np.diff(sm)) * N + re
    diag = np.searchsorted(keys, np.arange(N) * (N + 1))
    d = kvals.view(-1, 3, 3)[torch.as_tensor(diag, device=kvals.device)]
    return torch.diagonal(d, dim1=1, dim2=2).reshape(-1) < 1e19


def test_blade_static_multigrid(blade):
    import torch
    from fem_taichi_amd import Static_solver, solvers
    g, phys, model = blade
    s = Static_solver(model, 0)
    U = s.Solve_cpu(method="mg", tol=1e-8)
    K = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    F = model.F_device()
    Ud = torch.as_tensor(U, device=F.device)
    assert float((F - K.matvec(Ud)).norm() / F.norm()) < 1e-8
    free = _free_dofs(phys, model.K_bsr("f64"))
    assert int((~free).sum()) == 3 * 51 * 51                   # the root face x = 0
    assert float(Ud[~free].abs().max()) < 1e-12 * float(Ud.abs().max())


def test_blade_modes(blade):
    """20 modes, block 8: the four lowest eigenvalues agree between two start
    blocks to 1e-8, the Ritz vectors are M-orthogonal, and each pair meets the
    shift-invert residual lambda ||K^-1 M x - x / lambda|| / ||x|| <= 1e-6,
    with K^-1 applied by the block multigrid solve of the eigen-solver (its
    FP64 residual is checked in test_blade_static_multigrid).  The plain
    residual K x - lambda M x is not a useful measure here: the error left in
    x is multiplied by up to lambda_max / lambda_1 (about 1e6 on this mesh)."""
    import torch
    from fem_taichi_amd import Eigen_solver, _lib
    from fem_taichi_amd.runtime import stream_ptr
    g, phys, model = blade
    es = Eigen_solver(model)
    hz0, X = es.Solve_cpu(num_eigen=20, num_block=8, seed=0)
    hz1, _ = Eigen_solver(model).Solve_cpu(num_eigen=20, num_block=8, seed=1)
    assert np.all(np.diff(hz0) >= 0)
    assert np.max(np.abs(hz0[:4] - hz1[:4]) / hz0[:4]) < 1e-8
    lam = (2 * np.pi * hz0[:4]) ** 2
    Xd = torch.as_tensor(X[:, :4]).cuda().contiguous()
    MX = torch.empty_like(Xd)
    _lib.call("fem_mass_spmm", model.dev.ref, model.M_blocks().data_ptr(), Xd.data_ptr(), MX.data_ptr(), 4,
              stream_ptr())
    gram = (Xd.T @ MX).cpu().numpy()
    dg = np.sqrt(np.diag(gram))
    assert np.abs(gram / np.outer(dg, dg) - np.eye(4)).max() < 1e-8      # M-orthogonal, whatever the scaling
    Y = es.ops.solve(MX, 1e-12)
    lam_t = torch.as_tensor(lam, device=Xd.device)
    res = (lam_t * (Y - Xd / lam_t).norm(dim=0) / Xd.norm(dim=0)).cpu().numpy()
    print("shift-invert residuals", res, flush=True)
    assert np.all(res < 1e-6), res


def test_f32_k_within_stated_bound(slab):
    """The FP32 K of the full slab against its FP64 K: 1e-6 relative."""
    import torch
    g, phys, model = slab
    k64 = model.assemble_K("f64", "bsr", penalty=False)
    k32 = model.assemble_K("f32", "bsr", penalty=False).double()
    assert float((k32 - k64).norm() / k64.norm()) < 1e-6


def test_box_centrifugal_load_total(gpu):
    """Sum of the load vector of the unjittered 60x60x55 box spinning at
    omega = (0, 0, 100) rad/s, no nodal loads: rho * |omega|^2 * V * (xc, yc, 0),
    with (xc, yc) the box centre (the partition of unity sums the element
    integrals exactly)."""
    from fem_taichi_amd import Model, Physics
    g = mesh.structured_block(60, 60, 55, 1.0, 0.0, 1, load_face=None, omegas=(0.0, 0.0, 100.0))
    model = Model(g.name, Physics.from_grid(g), g.Materials, Omegas=g.Omegas)
    F = model.F_device().view(-1, 3).sum(0).cpu().numpy()
    rv = g.Materials["DE"] * 100.0 ** 2 * 60 * 60 * 55
    expect = np.array([rv * 30.0, rv * 30.0, 0.0])
    assert np.linalg.norm(F - expect) < 1e-12 * np.linalg.norm(expect)


def test_blade_task_lanes_aligned(blade):
    """Every block of the blade's device-built task lists (27.5M node-pair
    blocks, beyond 2^24) occupies exactly T adjacent, T-aligned lanes of its
    group, T non-increasing inside a group and the dual tasks last; the
    lane butterfly sums a block correctly only then."""
    import torch
    g, phys, model = blade
    t = phys.dev.t
    tp = t["grp_task_ptr"].long()
    w = t["tasks"].view(-1, 2).long() & 0xFFFFFFFF
    nt = w.shape[0]
    assert int(tp[-1]) == nt
    dual = ((w[:, 1] >> 31) & 1).bool()
    lt = (w[:, 1] >> 28) & 7
    T = torch.where(dual, torch.ones_like(lt), 1 << lt)
    outa = w[:, 0] & 0xFFFF
    G = tp.numel() - 1
    grp = torch.repeat_interleave(torch.arange(G, device=w.device), tp.diff())
    pos = torch.arange(nt, device=w.device) - tp[grp]
    start = torch.ones(nt, dtype=torch.bool, device=w.device)
    start[1:] = (grp[1:] != grp[:-1]) | (outa[1:] != outa[:-1]) | dual[1:] | dual[:-1]
    run = torch.cumsum(start.long(), 0) - 1
    length = torch.bincount(run)
    first = torch.nonzero(start).squeeze(1)
    assert torch.equal(length, T[first])                       # T tasks per block
    assert bool((pos[first] % T[first] == 0).all())           # T-aligned
    assert bool((T == T[first][run]).all())
    cls = torch.where(dual, torch.full_like(lt, 21), 20 - lt)  # 20 - log2 T, then the dual tasks
    same = grp[1:] == grp[:-1]
    assert bool((cls[1:][same] >= cls[:-1][same]).all())


def _structured_pattern(node_dims, device):
    """The node-pair pattern of a structured grid in closed form: row i holds
    the in-grid nodes of its 3x3x3 neighbourhood, ascending, which for the
    x-fastest numbering is the (dz, dy, dx) lexicographic order.  Returns
    (SmNode, ReNode) as int64 device tensors."""
    import torch
    nx, ny, nz = node_dims
    N = nx * ny * nz
    idx = torch.arange(N, device=device)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    off = torch.tensor([(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)], device=device)
    cx = x[:, None] + off[None, :, 0]
    cy = y[:, None] + off[None, :, 1]
    cz = z[:, None] + off[None, :, 2]
    ok = (cx >= 0) & (cx < nx) & (cy >= 0) & (cy < ny) & (cz >= 0) & (cz < nz)
    col = cx + nx * (cy + ny * cz)
    sm = torch.zeros(N + 1, dtype=torch.int64, device=device)
    sm[1:] = torch.cumsum(ok.sum(1), 0)
    return sm, col[ok]


def _check_device_pattern(phys, dims):
    """The GPU pattern of a structured grid equals the closed form entry for
    entry; nnzb = prod(3 n_d - 2); rows strictly ascending, the diagonal in
    every row, the pattern symmetric, 27 entries on interior rows."""
    import torch
    d = phys.dev["re_node"].device
    sm = phys.dev["sm_node"].long()
    re = phys.dev["re_node"].long()
    nx, ny, nz = dims
    nnzb = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    assert int(sm[-1]) == nnzb == re.numel()
    sm_ref, re_ref = _structured_pattern(dims, d)
    assert torch.equal(sm, sm_ref) and torch.equal(re, re_ref)
    N = sm.numel() - 1
    row = torch.repeat_interleave(torch.arange(N, device=d), sm.diff())
    same = row[1:] == row[:-1]
    assert bool((re[1:][same] > re[:-1][same]).all())            # strictly ascending per row
    assert int((re == row).sum()) == N                           # one diagonal per row
    key = torch.sort(row * N + re).values
    tkey = torch.sort(re * N + row).values
    assert torch.equal(key, tkey)                                # symmetric
    co = sm.diff()
    assert int(co.max()) == 27 and int((co == 27).sum()) == (nx - 2) * (ny - 2) * (nz - 2)
    assert torch.equal(phys.dev["co_node"].long(), co)


def test_blade_device_pattern(blade):
    """The 400x50x50 blade's GPU pattern (1.04M nodes, 27.4M blocks)."""
    g, phys, model = blade
    _check_device_pattern(phys, g.node_dims)


@pytest.fixture(scope="module")
def blade_c3(gpu):
    """The modal workload at scale: the twisted blade refined to 500x60x75
    elements (2.32M nodes, 6.97M DOF), no stress stiffening."""
    from fem_taichi_amd import Model, Physics
    g = mesh.twisted_blade(500, 60, 75)
    print("\ncase twisted_blade_500x60x75", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0.0, 0.0, 0.0))
    return g, phys, model


def test_c3_device_pattern(blade_c3):
    """The 500x60x75 blade's GPU pattern (2.32M nodes, 61.4M blocks; 9 nnzb
    = 552.6M CSR entries, below the 2^31 index limit)."""
    g, phys, model = blade_c3
    _check_device_pattern(phys, g.node_dims)
    assert 9 * model.nnzb < 2 ** 31


def test_c3_modes_100_block_16(blade_c3):
    """100 modes by block Lanczos with block 16 on the 2.32M-node blade: the
    lowest five eigenvalues agree between two start blocks to 1e-8 (measured
    1e-12 .. 1e-10; modes 6 to 8 agree to 1e-8 .. 5e-8, the convergence of
    higher modes in a 12-step Krylov space), the Ritz vectors of the lowest
    eight are M-orthogonal with shift-invert residuals <= 1e-6, and no inner
    solve handed back a column above the FP64 floor."""
    import torch
    from fem_taichi_amd import Eigen_solver, _lib
    from fem_taichi_amd.runtime import stream_ptr
    g, phys, model = blade_c3
    es = Eigen_solver(model)
    hz0, X = es.Solve_cpu(num_eigen=100, num_block=16, seed=0, inner_tol=1e-10)
    print("c3 frequencies (Hz)", hz0[:8], "worst inner residual", es.ops.inner_rel_max, flush=True)
    assert hz0.shape == (100,) and np.all(np.diff(hz0) >= 0) and np.all(np.isfinite(hz0))
    assert es.ops.inner_rel_max < 1e-7
    hz1, _ = Eigen_solver(model).Solve_cpu(num_eigen=100, num_block=16, seed=1, inner_tol=1e-10)
    rel = np.abs(hz0[:8] - hz1[:8]) / hz0[:8]
    print("start-block difference of the lowest 8", rel, flush=True)
    assert np.max(rel[:5]) < 1e-8
    k = 8
    lam = (2 * np.pi * hz0[:k]) ** 2
    Xd = torch.as_tensor(X[:, :k]).cuda().contiguous()
    MX = torch.empty_like(Xd)
    _lib.call("fem_mass_spmm", model.dev.ref, model.M_blocks().data_ptr(), Xd.data_ptr(), MX.data_ptr(), k,
              stream_ptr())
    gram = (Xd.T @ MX).cpu().numpy()
    dg = np.sqrt(np.diag(gram))
    assert np.abs(gram / np.outer(dg, dg) - np.eye(k)).max() < 1e-8
    Y = es.ops.solve(MX, 1e-12)
    lam_t = torch.as_tensor(lam, device=Xd.device)
    res = (lam_t * (Y - Xd / lam_t).norm(dim=0) / Xd.norm(dim=0)).cpu().numpy()
    print("c3 shift-invert residuals", res, flush=True)
    assert np.all(res < 1e-6), res


def test_blade_learned_shipped_weights(blade):
    """The learned form (FP32 multigrid V-cycle plus the shipped mesh-conv
    correction inside FP64 flexible CG) on the 1M-node blade: FP64-verified
    relative residual 1e-8 without the block-Jacobi PCG fallback."""
    import torch
    from fem_taichi_amd import learned as L, solvers
    g, phys, model = blade
    F = model.F_device()
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    x, info = L.solve_learned(model, K64, F, tol=1e-8, base="mg")
    h = info.history[0]
    print(f"blade learned: {info.iterations} iterations, {h['net_applications']} network applications, "
          f"{info.seconds:.3f} s", flush=True)
    assert info.method == "learned_mg" and not h["fell_back"] and h["net_applications"] > 0
    assert float(torch.linalg.vector_norm(F - K64.matvec(x)) / torch.linalg.vector_norm(F)) < 1e-8
