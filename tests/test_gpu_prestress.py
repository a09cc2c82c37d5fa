"""Prestressed modal analysis on the GPU (stress stiffening and spin
softening; the reference lists them as missing, Readme.md:189) and the
reference's own prestressed path (the modal solve after a large-displacement
static solve uses its tangent KT, SOLVE_eigen.py:32).

* K_L + K_sigma(U) (fem_hex8_assemble_ks) and the spin softening
  (fem_spin_softening) against the oracle (element_KS, spin_blocks): 1e-12
  relative Frobenius, penalty entries exact, both layouts.
* Modal solves: the lowest eigenvalues against scipy's shift-invert eigsh on
  the oracle's matrices, 1e-8 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu

K_TOL = 1e-12
EIG_TOL = 1e-8


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _model(g, omegas):
    from fem_taichi_amd import Model, Physics
    phys = Physics.from_grid(g)
    return phys, Model(g.name, phys, g.Materials, Omegas=omegas)


@pytest.mark.parametrize("omega", [None, (0.0, 0.0, 100.0), (30.0, -50.0, 80.0)])
def test_KS_matches_oracle(gpu, omega):
    g = mesh.structured_block(5, 4, 6, 1.0, 0.10, 1, E_sigma=0.3)
    print(f"\ncase KS jit_5x4x6 omega {omega}", flush=True)
    phys, model = _model(g, (0, 0, 0))
    pp = O.pre_process(g.Node, g.Elem)
    U = np.random.default_rng(4).standard_normal(3 * phys.Num_Node) * 1e-3
    Vref = O.model_KS_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds, U, DE=g.DE, OM=omega)
    vals = model.assemble_KS(U, omega=omega).cpu().numpy().reshape(-1, 3, 3)
    pen = Vref >= 1e19
    assert np.array_equal(vals[pen], Vref[pen])
    assert rel(vals[~pen], Vref[~pen]) < K_TOL
    n = 3 * phys.Num_Node
    Kref = sp.bsr_matrix((Vref, phys.ReNode, phys.SmNode), shape=(n, n)).tocsr()
    Kcsr = model.Get_MatrixKS(U, omega)
    assert np.array_equal(Kcsr.indptr, Kref.indptr) and np.array_equal(Kcsr.indices, Kref.indices)
    assert rel(Kcsr.data[Kref.data < 1e19], Kref.data[Kref.data < 1e19]) < K_TOL


def test_KS_zero_displacement_is_K(gpu):
    g = mesh.structured_block(4, 3, 5, 1.0, 0.10, 2)
    phys, model = _model(g, (0, 0, 0))
    a = model.assemble_KS(np.zeros(3 * phys.Num_Node)).cpu().numpy()
    b = model.assemble_K("f64", "bsr").cpu().numpy()
    assert rel(a, b) < K_TOL


def test_prestressed_modal_of_rotating_blade(gpu):
    """A rotating twisted blade (omega = 100 rad/s about z), the prestress
    from its own centrifugal load: the lowest eigenvalues of
    (K_L + K_sigma(U) - K_spin, M) against eigsh on the oracle's matrices,
    U from the oracle's direct static solve."""
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g = mesh.twisted_blade(24, 4, 4)
    print("\ncase prestressed modal blade 24x4x4", flush=True)
    phys, model = _model(g, g.Omegas)
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    F = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    U = sla.spsolve(K.tocsc(), F)
    n = 3 * phys.Num_Node
    VS = O.model_KS_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds, U, DE=g.DE, OM=g.Omegas)
    KS = sp.bsr_matrix((VS, phys.ReNode, phys.SmNode), shape=(n, n)).tocsc()
    ref = np.sort(sla.eigsh(KS, k=4, M=M.tocsr(), sigma=0, which='LM', return_eigenvectors=False))
    ref0 = np.sort(sla.eigsh(K.tocsc(), k=4, M=M.tocsr(), sigma=0, which='LM', return_eigenvectors=False))
    s = Eigen_solver(model, 0)
    s.Solve_cpu(num_eigen=12, num_block=4, seed=1, prestress="static")
    assert s.operator == "KS"
    # no inner solve handed back a column stalled above the FP64 floor
    print("worst inner residual", s.ops.inner_rel_max, "fallback columns", s.ops.fallback_columns, flush=True)
    assert s.ops.inner_rel_max < 1e-7
    lam = s.eigenvalues.cpu().numpy()
    print("prestressed", lam[:4], "eigsh", ref, "unstressed eigsh", ref0, flush=True)
    assert np.allclose(lam[:3], ref[:3], rtol=EIG_TOL, atol=0)
    assert not np.allclose(ref[:3], ref0[:3], rtol=1e-6, atol=0)     # the prestress matters


def test_modal_after_newton_uses_tangent(gpu):
    """After the large-displacement static solve, the modal solve runs on its
    final tangent KT (the reference factors Model.KT): eigenvalues against
    eigsh(KT(U), M) from the oracle."""
    from fem_taichi_amd import Static_solver
    from fem_taichi_amd.solve_eigen import Eigen_solver
    g = mesh.structured_block(2, 2, 12, 1.0, 0.05, 1, total_load=1000.0)
    ids = g.Forces[0][0]
    g.Forces = [[ids, [150.0 / len(ids), 0.0, 0.0]]]
    print("\ncase modal after Newton, cantilever 2x2x12", flush=True)
    phys, model = _model(g, (0, 0, 0))
    U = Static_solver(model, 1).Solve_cpu(max_errs=1e-9, max_iters=30, tol=1e-12)
    pp = O.pre_process(g.Node, g.Elem)
    VT, _ = O.model_KN_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds, U)
    n = 3 * phys.Num_Node
    KT = sp.bsr_matrix((VT, phys.ReNode, phys.SmNode), shape=(n, n)).tocsc()
    M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
    ref = np.sort(sla.eigsh(KT, k=4, M=M.tocsr(), sigma=0, which='LM', return_eigenvectors=False))
    s = Eigen_solver(model, 0)
    s.Solve_cpu(num_eigen=12, num_block=4, seed=2)
    assert s.operator == "KT"
    print("worst inner residual", s.ops.inner_rel_max, "fallback columns", s.ops.fallback_columns, flush=True)
    assert s.ops.inner_rel_max < 1e-7
    lam = s.eigenvalues.cpu().numpy()
    print("KT modes", lam[:4], "eigsh", ref, flush=True)
    assert np.allclose(lam[:3], ref[:3], rtol=EIG_TOL, atol=0)
