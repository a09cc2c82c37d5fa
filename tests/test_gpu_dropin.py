"""Drop-in behaviour of the public classes on the GPU.

* A Read-shaped object (the reference's ``Read`` attributes only:
  READ_physics.py:44-61, 96-97, 114, 134 -- Node, Elem, Bounds, Forces and
  the SciPy pre_process outputs, no device state) handed straight to
  ``Model`` and ``Post``: the matrices, load and stress equal the oracle's,
  and Model and Post of one Read share one GPU pattern.
* Every returned CSR matrix owns its index arrays (the reference returns
  fresh to_numpy() copies, FEM_model_csr.py:118-127): pruning one matrix in
  place leaves the next one, and the Model's cache, unchanged.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class ReadLike:
    """What the reference's Read holds after its own (SciPy) pre_process."""

    def __init__(self, name, Node, Elem, Bounds, Forces):
        self.name, self.Node, self.Elem, self.Bounds, self.Forces = name, Node, Elem, Bounds, Forces
        pp = O.pre_process(Node, Elem)
        for k in ("Num_Node", "Num_Elem", "Num_MaxN", "Nind", "SmNode", "ReNode", "CoNode", "CoElem"):
            setattr(self, k, pp[k])


def test_read_object_into_model_and_post(gpu):
    from fixtures import STEEL, load_mesh2
    from fem_taichi_amd import Model, Post
    Node, Elem, Bounds, Forces = load_mesh2()
    print("\ncase Read-shaped Mesh/2.cdb into Model and Post", flush=True)
    rd = ReadLike("2", Node, Elem, Bounds, Forces)
    assert not hasattr(rd, "dev")
    model = Model("2", rd, dict(STEEL), Omegas=(0, 0, 100))
    pp = O.pre_process(Node, Elem)
    K = model.Get_MatrixKL()
    Kref = O.model_K_csr(Node, Elem, pp, STEEL["EX"], STEEL["PO"], Bounds)
    assert np.array_equal(K.indptr, Kref.indptr) and np.array_equal(K.indices, Kref.indices)
    pen = Kref.data >= 1e19
    assert np.array_equal(K.data[pen], Kref.data[pen]) and rel(K.data[~pen], Kref.data[~pen]) < 1e-12
    F = model.Get_VectorF()
    Fref = O.model_F(Node, Elem, pp, STEEL["DE"], (0, 0, 100), Forces)
    assert rel(F, Fref) < 1e-12
    for k in ("SmNode", "ReNode", "CoNode", "CoElem"):
        assert np.array_equal(getattr(model.Physics, k).astype(np.int64), getattr(rd, k).astype(np.int64)), k
    U = np.random.default_rng(3).standard_normal(3 * rd.Num_Node) * 1e-3
    post = Post("2", rd, dict(STEEL), U.reshape(-1, 3, 1))
    assert post.Physics is model.Physics            # one pattern for both
    S = post.Get_Stress()
    Sref = O.get_stress(Node, Elem, pp, STEEL["EX"], STEEL["PO"], U)
    assert rel(S, Sref) < 1e-10


def test_returned_matrices_own_their_indices(gpu):
    from fem_taichi_amd import Model, Physics, mesh
    g = mesh.structured_block(4, 3, 5, 1.0, 0.1, 2)
    print("\ncase independent CSR index arrays 4x3x5", flush=True)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=(0, 0, 0))
    A = model.Get_MatrixKL()
    B = model.Get_MatrixKL()
    ref_ptr, ref_idx, ref_nnz = B.indptr.copy(), B.indices.copy(), B.nnz
    A.data[::7] = 0.0
    A.eliminate_zeros()                              # edits A's indptr / indices in place
    assert A.nnz < ref_nnz
    assert B.nnz == ref_nnz and np.array_equal(B.indptr, ref_ptr) and np.array_equal(B.indices, ref_idx)
    C = model.Get_MatrixKL()
    assert np.array_equal(C.indptr, ref_ptr) and np.array_equal(C.indices, ref_idx)
    KT, KM = model.Get_MatrixKN()
    KT.data[:] = 0.0
    KT.eliminate_zeros()
    assert KM.nnz == ref_nnz and np.array_equal(KM.indices, ref_idx)
    assert (sp.csr_matrix(C) != B).nnz == 0
