"""GPU block Lanczos (SOLVE_eigen.py:45-89) against the oracle's restatement
with exact (SuperLU) inner solves and the same start block, and against
scipy's shift-invert eigsh.  Tolerance: converged eigenvalues within 1e-8
relative (the FP64 bar), M-orthonormal vectors within 1e-8."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu

EIG_TOL = 1e-8


@pytest.fixture(scope="module")
def cantilever(gpu):
    from fem_taichi_amd import Model, Physics
    g = mesh.structured_block(4, 4, 10, 1.0, 0.05, 0)
    print("\ncase cantilever 4x4x10", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0, 0, 0))
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    return g, model, K.tocsc(), M.tocsr()


def test_eigen_solver_matches_oracle(cantilever):
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g, model, K, M = cantilever
    num_eigen, b = 20, 4
    n = K.shape[0]
    u0 = np.random.default_rng(0).uniform(size=(n, b))
    lu = sla.splu(K)
    lam_ref, X_ref, _ = O.lanczos(K, M, lambda B: lu.solve(B), num_eigen, b, u0=u0)
    solver = Eigen_solver(model, 0)
    freq, X = solver.Solve_cpu(num_eigen=num_eigen, num_block=b, u0=u0)     # the oracle's start block
    lam = solver.eigenvalues.cpu().numpy()
    # compare the Ritz pairs the oracle shows converged, by the shift-invert
    # residual lam * ||K^-1 M x - x / lam|| / ||x|| <= 1e-4 (the plain K
    # residual is dominated by the 1e20 penalty rows); the unconverged tail
    # of a short Krylov space moves with the last bits of the inner solves
    res = np.linalg.norm(lu.solve(M @ X_ref) - X_ref / lam_ref, axis=0) * lam_ref / np.linalg.norm(X_ref, axis=0)
    conv = np.nonzero(res <= 1e-4)[0]
    assert len(conv) >= 4
    assert np.allclose(lam[conv], lam_ref[conv], rtol=EIG_TOL, atol=0)
    assert np.allclose(freq, np.sqrt(lam) / 2 / np.pi)
    G = X.T @ (M @ X)
    assert np.abs(G - np.eye(num_eigen)).max() < 1e-8


def test_eigen_solver_matches_eigsh(cantilever):
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g, model, K, M = cantilever
    ref = np.sort(sla.eigsh(K, k=6, M=M, sigma=0, which='LM', return_eigenvectors=False))
    solver = Eigen_solver(model, 0)
    solver.Solve_cpu(num_eigen=12, num_block=6, seed=1)
    lam = solver.eigenvalues.cpu().numpy()
    assert np.allclose(lam[:3], ref[:3], rtol=EIG_TOL, atol=0)


def test_block_lancozs_scipy_api(cantilever):
    from fem_taichi_amd.solve_eigen import Block_lancozs, Eigen_solver
    g, model, K, M = cantilever
    lam, X = Block_lancozs(K.tocsr(), M, None, num_eigen=8, num_block=4).Solve_cpu(seed=3)
    s2 = Eigen_solver(model, 0)
    s2.Solve_cpu(num_eigen=8, num_block=4, seed=3)
    assert np.allclose(lam[:4], s2.eigenvalues.cpu().numpy()[:4], rtol=EIG_TOL, atol=0)


def test_zero_padding_quirk(cantilever):
    """num_block not dividing 2*num_eigen leaves T's tail zero: the reference's
    Ritz values there are 1/0 = inf and sort last (SOLVE_eigen.py:47, 61)."""
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g, model, K, M = cantilever
    s = Eigen_solver(model, 0)
    s.Solve_cpu(num_eigen=5, num_block=4, seed=0)
    lam = s.eigenvalues.cpu().numpy()
    assert np.all(np.isfinite(lam)) and np.all(np.diff(lam) >= 0)


def test_eigen_solver_ansys_mesh2(gpu):
    """The reference's Mesh/2.cdb, root sets fixed: the lowest modes against
    scipy's shift-invert eigsh on the oracle's K and M."""
    from fem_taichi_amd import Model, Physics
    from fem_taichi_amd.solve_eigen import Eigen_solver
    from fixtures import STEEL, load_mesh2
    Node, Elem, Bounds, Forces = load_mesh2()
    print("\ncase ansys_mesh2 modes", flush=True)
    phys = Physics(Node, Elem, Bounds, Forces, name="mesh2")
    model = Model("mesh2", phys, dict(STEEL), Omegas=(0, 0, 0))
    pp = O.pre_process(Node, Elem)
    K = O.model_K_csr(Node, Elem, pp, STEEL["EX"], STEEL["PO"], Bounds)
    M = O.model_M_csr(Node, Elem, pp, STEEL["DE"])
    ref = np.sort(sla.eigsh(K.tocsc(), k=6, M=M.tocsr(), sigma=0, which='LM', return_eigenvectors=False))
    # the lowest modes are clustered (1.839644e5, 1.839649e5 twice, 1.839661e5
    # twice): 10 block steps separate them (the oracle's Lanczos with exact
    # solves agrees with eigsh to 1e-13 at these settings, 1e-6 at 16 / 8)
    solver = Eigen_solver(model, 0)
    solver.Solve_cpu(num_eigen=40, num_block=8, seed=2)
    lam = solver.eigenvalues.cpu().numpy()
    print("lambda", lam[:6], "eigsh", ref, flush=True)
    assert np.allclose(lam[:5], ref[:5], rtol=EIG_TOL, atol=0)


def test_eigen_solver_multigrid_inner_matches_pcg_inner(gpu):
    """Inner solves by multigrid flexible CG (a 2-level hierarchy on a twisted
    blade) against the same Lanczos with block-Jacobi PCG inner solves and
    against eigsh on the oracle's K and M: the lowest eigenvalues to 1e-8."""
    from fem_taichi_amd import Model, Physics
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g = mesh.twisted_blade(40, 6, 6)
    print("\ncase blade 40x6x6 modes, inner mg vs pcg", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0, 0, 0))
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    ref = np.sort(sla.eigsh(K.tocsc(), k=4, M=M.tocsr(), sigma=0, which='LM', return_eigenvectors=False))
    s_mg = Eigen_solver(model, 0)
    s_mg.Solve_cpu(num_eigen=16, num_block=4, seed=5, inner="mg")
    assert len(s_mg.ops.pre.hier.dims) >= 2
    s_pcg = Eigen_solver(model, 0)
    s_pcg.Solve_cpu(num_eigen=16, num_block=4, seed=5, inner="pcg")
    lam_mg = s_mg.eigenvalues.cpu().numpy()
    lam_pcg = s_pcg.eigenvalues.cpu().numpy()
    print("mg", lam_mg[:4], "pcg", lam_pcg[:4], "eigsh", ref, "inner its mg", s_mg.ops.iterations[:3],
          "pcg", s_pcg.ops.iterations[:3], flush=True)
    assert np.allclose(lam_mg[:3], ref[:3], rtol=EIG_TOL, atol=0)
    assert np.allclose(lam_mg[:3], lam_pcg[:3], rtol=EIG_TOL, atol=0)


@pytest.mark.parametrize("n,p,q,ld", [(1000, 8, 8, 8), (300000, 40, 8, 56), (7, 3, 5, 3), (50000, 200, 16, 200),
                                      (20000, 600, 20, 640), (5000, 16, 70, 16), (0, 4, 4, 4), (1, 1, 1, 1)])
def test_gram_matches_numpy(gpu, n, p, q, ld):
    """fem_gemm_tn_f64 (through solve_eigen.gram) against numpy's X^T Y, on a
    row-strided view of X as the Lanczos basis is, including the column
    splits beyond the kernel's p <= 512 / q <= 64 and empty inputs; two calls
    give the same bits."""
    import torch
    from fem_taichi_amd.solve_eigen import gram
    print(f"\ngram n {n} p {p} q {q} lda {ld}", flush=True)
    rng = np.random.default_rng(n + p)
    Xw = rng.standard_normal((n, ld))
    Y = rng.standard_normal((n, q)) * 1e3
    Xd = torch.as_tensor(Xw).cuda()[:, :p]
    Yd = torch.as_tensor(Y).cuda()
    C = gram(Xd, Yd).cpu().numpy()
    ref = Xw[:, :p].T @ Y
    assert C.shape == (p, q)
    if n == 0:
        assert np.all(C == 0)
        return
    err = np.abs(C - ref).max() / (np.abs(Xw[:, :p]).max() * np.abs(Y).max() * np.sqrt(n))
    assert err < 1e-14
    assert np.array_equal(gram(Xd, Yd).cpu().numpy(), C)


@pytest.mark.parametrize("inner", ["mg", "pcg"])
def test_block_solve_wider_than_kernels(gpu, inner):
    """A block of 70 right-hand sides (wider than the 64-column kernels) is
    solved in pieces: every column's FP64 residual on the oracle's K below
    1e-8, one iteration record for the call."""
    import torch
    from fem_taichi_amd import Model, Physics
    from fem_taichi_amd.solve_eigen import BlockOps, MultigridBlockOps
    g = mesh.structured_block(6, 5, 8, 1.0, 0.05, 1)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0, 0, 0))
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    n = K.shape[0]
    B = np.random.default_rng(7).standard_normal((n, 70))
    if inner == "mg":
        ops = MultigridBlockOps(model, None, min_nodes=60)
    else:
        ops = BlockOps(model.dev, model.K_bsr("f64"), None)
    X = ops.solve(torch.as_tensor(B).cuda(), 1e-10).cpu().numpy()
    assert X.shape == (n, 70) and len(ops.iterations) == 1
    res = np.linalg.norm(B - K @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() < 1e-8, res.max()


@pytest.mark.parametrize("n,k,q", [(1, 1, 1), (1000, 37, 10), (70001, 200, 16), (5000, 16, 64), (3, 300, 8)])
def test_gemm_nn_matches_numpy(gpu, n, k, q):
    """fem_gemm_nn_f64 (the Lanczos block updates) against NumPy, including a
    strided A (a column slice of a wider Q), beta = 1 and alpha = -1."""
    import torch
    from fem_taichi_amd.solve_eigen import gemm_nn
    rng = np.random.default_rng(n + k + q)
    Qw = rng.standard_normal((n, k + 5))
    C = rng.standard_normal((k, q))
    Y0 = rng.standard_normal((n, q))
    Qd = torch.as_tensor(Qw).cuda()
    A = Qd[:, :k]                          # row stride k + 5
    Y = gemm_nn(A, torch.as_tensor(C).cuda())
    ref = Qw[:, :k] @ C
    assert np.abs(Y.cpu().numpy() - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * k
    Yd = torch.as_tensor(Y0).cuda()
    gemm_nn(A, torch.as_tensor(C).cuda(), Yd, -1.0, 1.0)
    ref2 = Y0 - ref
    assert np.abs(Yd.cpu().numpy() - ref2).max() <= 1e-13 * max(1.0, np.abs(ref2).max()) * k


@pytest.mark.parametrize("b", [1, 4, 10, 16, 64])
def test_chol_upper_inv_matches_scipy(gpu, b):
    import scipy.linalg as sl
    import torch
    from fem_taichi_amd.solve_eigen import chol_upper_inv
    rng = np.random.default_rng(b)
    G = rng.standard_normal((b + 3, b))
    A = G.T @ G + b * np.eye(b) * 1e-3
    R, Rinv = chol_upper_inv(torch.as_tensor(A).cuda())
    R, Rinv = R.cpu().numpy(), Rinv.cpu().numpy()
    Rref = sl.cholesky(A)                  # upper, A = R^T R (the reference's call)
    assert np.allclose(np.tril(R, -1), 0.0, atol=0.0)
    assert np.abs(R - Rref).max() <= 1e-12 * np.abs(Rref).max()
    assert np.abs(Rinv @ R - np.eye(b)).max() <= 1e-12 * b
    # not positive definite: NaN, no exception inside the kernel
    Bad = A.copy()
    Bad[b - 1, b - 1] = -1.0
    Rb, _ = chol_upper_inv(torch.as_tensor(Bad).cuda())
    assert np.isnan(Rb.cpu().numpy()[b - 1, b - 1])
