"""The reference's own CPU-runnable case end to end on the GPU: cantilever
20x20x50 elements (22,491 nodes, 67,473 DOF), jitter 5% (seed 0), z = 0 face
fixed, -z load of 1000 N on the top face (mesh.cantilever_20x20x50), through
the drop-in calls the reference's run script makes (Get_MatrixKL / M /
VectorF, Static_solver, Post.Get_Stress, Eigen_solver with 20 modes and
block 8).

* K (the reference's CSR layout), M and F against the oracle: 1e-12
  relative Frobenius, index arrays exact, penalty entries exact.
* Static solve, plain FP64 PCG and multigrid flexible CG, to an FP64
  residual of 1e-11 on the oracle's K and F; the top-face displacements
  equal the direct solution's (SuperLU, tests/golden/c0_direct.npz) to 1e-9
  (a residual of 1e-11 moves U by about 1e-12 on this mesh).
* Stress of the GPU solution against the oracle's Get_Stress of the same U:
  1e-10.
* Block Lanczos: the four lowest eigenvalues against ARPACK shift-invert on
  the oracle's K and M (the fixture) to 1e-8.
"""
import os

import numpy as np
import pytest

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "c0_direct.npz"))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def c0(gpu):
    from fem_taichi_amd import Model, Physics
    g = mesh.cantilever_20x20x50()
    print("\ncase c0 cantilever 20x20x50", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    F = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    return g, phys, model, pp, K, F


def test_c0_matrices(c0):
    g, phys, model, pp, K, F = c0
    Kg = model.Get_MatrixKL()
    assert np.array_equal(Kg.indptr, K.indptr) and np.array_equal(Kg.indices, K.indices)
    pen = K.data >= 1e19
    assert np.array_equal(Kg.data[pen], K.data[pen])
    assert rel(Kg.data[~pen], K.data[~pen]) < 1e-12
    M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    Mg = model.Get_MatrixM()
    assert np.array_equal(Mg.indptr, M.indptr) and np.array_equal(Mg.indices, M.indices)
    assert rel(Mg.data, M.data) < 1e-12
    assert rel(model.Get_VectorF(), F) < 1e-12


@pytest.mark.parametrize("method", ["pcg", "mg"])
def test_c0_static_matches_direct_solve(c0, method):
    from fem_taichi_amd import Post, Static_solver
    g, phys, model, pp, K, F = c0
    U = Static_solver(model, 0).Solve_cpu(method=method, tol=1e-11)
    res = O.static_residual(K, U, F)
    print(f"{method}: residual {res:.2e} (direct solve {float(GOLD['residual']):.2e})", flush=True)
    assert res < 1e-11
    assert np.isclose(np.linalg.norm(F), float(GOLD["F_norm"]), rtol=1e-12)
    assert rel(U.reshape(-1, 3)[GOLD["top_nodes"]], GOLD["U_top"]) < 1e-9
    assert np.isclose(np.linalg.norm(U), float(GOLD["U_norm"]), rtol=1e-9)
    S = Post(g.name, phys, g.Materials, U.reshape(-1, 3, 1)).Get_Stress()
    Sref = O.get_stress(g.Node, g.Elem, pp, g.EX, g.PO, U)
    assert rel(S[:, :6], Sref[:, :6]) < 1e-10
    assert rel(S[:, 6:9], Sref[:, 6:9]) < 1e-10


def test_c0_modes_match_arpack(c0):
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g, phys, model, pp, K, F = c0
    s = Eigen_solver(model, 0)
    s.Solve_cpu(num_eigen=20, num_block=8, seed=0)
    lam = s.eigenvalues.cpu().numpy()
    print("lanczos", lam[:4], "arpack", GOLD["eig"][:4], flush=True)
    assert np.allclose(lam[:4], GOLD["eig"][:4], rtol=1e-8, atol=0)
