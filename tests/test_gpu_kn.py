"""GPU parity of the large-displacement matrices (Get_MatrixKN / UnitK,
FEM_model_bsr.py:134-145, 476-598) against the oracle's element_KN, and the
Qtype=1 Newton solve (SOLVE_static.py:41-76).  Tolerance: the FP64 bar,
1e-12 relative Frobenius, penalty entries exact."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def cases():
    from types import SimpleNamespace
    from fixtures import STEEL, load_mesh2
    Node, Elem, Bounds, Forces = load_mesh2()
    m2 = SimpleNamespace(Node=Node, Elem=Elem, Bounds=Bounds, Forces=Forces, Materials=dict(STEEL),
                         EX=STEEL["EX"], PO=STEEL["PO"], DE=STEEL["DE"], Omegas=(0, 0, 100), name="ansys_mesh2",
                         node_dims=None)
    return {"jit_5x4x6_lognormal": mesh.structured_block(5, 4, 6, 1.0, 0.1, 3, E_sigma=0.3), "ansys_mesh2": m2}


@pytest.fixture(scope="module", params=list(cases().keys()))
def kn_setup(request, gpu):
    from fem_taichi_amd import Model, Physics
    g = cases()[request.param]
    print(f"\ncase {request.param}", flush=True)
    phys = Physics.from_grid(g)
    pp = O.pre_process(g.Node, g.Elem)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    scale = 0.02 * np.ptp(np.asarray(g.Node[-3:], dtype=np.float64), axis=1).min() / 10
    U = scale * np.random.default_rng(11).standard_normal(3 * phys.Num_Node)
    return g, phys, pp, model, U


def test_kn_bsr_and_csr_match_oracle(kn_setup):
    g, phys, pp, model, U = kn_setup
    model.LDinit(U)
    kt, km = model.assemble_KN("bsr")
    VT, VM = O.model_KN_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds, U)
    for v, ref in ((kt, VT), (km, VM)):
        v = v.cpu().numpy().reshape(-1, 3, 3)
        pen = ref >= 1e19
        assert np.array_equal(v[pen], ref[pen])
        assert rel(v[~pen], ref[~pen]) < 1e-12
    KT, KM = model.Get_MatrixKN()
    n = 3 * phys.Num_Node
    for C, ref in ((KT, VT), (KM, VM)):
        B = sp.bsr_matrix((ref, phys.ReNode, phys.SmNode), shape=(n, n)).tocsr()
        assert np.array_equal(C.indptr, B.indptr) and np.array_equal(C.indices, B.indices)
        free = np.abs(B.data) < 1e19
        assert rel(C.data[free], B.data[free]) < 1e-12


def test_kn_at_zero_displacement_is_linear_K(kn_setup):
    g, phys, pp, model, U = kn_setup
    model.LDinit(np.zeros(3 * phys.Num_Node))
    KT, KM = model.Get_MatrixKN()
    K = model.Get_MatrixKL()
    free = np.abs(K.data) < 1e19
    assert rel(KT.data[free], K.data[free]) < 1e-12 and rel(KM.data[free], K.data[free]) < 1e-12


def test_newton_large_displacement():
    """Qtype=1 on a 2 x 2 x 12 cantilever with a 150 N transverse tip load
    (1.3 % away from the linear solution): converges (13 steps on the CPU
    oracle with direct solves), and the oracle's own residual KM(U) U - F at
    the returned U is below the tolerance."""
    from fem_taichi_amd import Model, Physics, Static_solver
    g = mesh.structured_block(2, 2, 12, 1.0, 0.05, 1, total_load=1000.0)
    ids = g.Forces[0][0]
    g.Forces = [[ids, [150.0 / len(ids), 0.0, 0.0]]]
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0, 0, 0))
    s = Static_solver(model, 1)
    U = s.Solve_cpu(max_errs=1e-9, max_iters=30, tol=1e-12)
    assert s.history[-1] <= 1e-9 and len(s.history) > 2
    pp = O.pre_process(g.Node, g.Elem)
    VT, VM = O.model_KN_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds, U)
    n = 3 * phys.Num_Node
    KM = sp.bsr_matrix((VM, phys.ReNode, phys.SmNode), shape=(n, n)).tocsr()
    F = O.model_F(g.Node, g.Elem, pp, g.DE, (0, 0, 0), g.Forces)
    assert np.linalg.norm(KM @ U - F) / np.linalg.norm(F) < 1e-8
    # the linear solution differs: the load is in the nonlinear range
    Ul = Static_solver(model, 0).Solve_cpu(tol=1e-12)
    assert np.linalg.norm(U - Ul) / np.linalg.norm(Ul) > 1e-3
