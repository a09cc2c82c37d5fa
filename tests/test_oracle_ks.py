"""The oracle's prestressed stiffness (element_KS) and spin softening
(spin_blocks), CPU only.  Each check is independent of the restatement's own
formula: reductions to element_K, and closed forms on affine fields."""
import numpy as np

from oracle import fem_oracle as O
from fem_taichi_amd import mesh


def _elements(seed=0, jitter=0.1):
    g = mesh.structured_block(2, 2, 2, 1.0, jitter, seed)
    pp = O.pre_process(g.Node, g.Elem)
    coords = O.element_coords(g.Node, g.Elem, pp["Nind"])          # (E, 3, 8)
    return g, pp, coords


def test_zero_and_rigid_displacement_give_element_K():
    g, pp, C = _elements()
    K = O.element_K(C, 2.0e5, 0.3)
    assert np.allclose(O.element_KS(C, np.zeros_like(C), 2.0e5, 0.3), K, rtol=0, atol=1e-9 * np.abs(K).max())
    U = np.tile(np.array([0.3, -0.1, 0.2]), g.Num_Node)
    CU = O.element_coord_u(U, g.Elem, pp["Nind"])
    assert np.allclose(O.element_KS(C, CU, 2.0e5, 0.3), K, rtol=0, atol=1e-9 * np.abs(K).max())


def test_affine_fields_closed_form():
    """u = G x gives the uniform stress s = D sym(G); for v = H x,
    v.K_sigma.v = V tr(H s H^T) with V the element volume (sum of detJ)."""
    g, pp, C = _elements(seed=3)
    rng = np.random.default_rng(0)
    G = 1e-3 * rng.standard_normal((3, 3))
    H = rng.standard_normal((3, 3))
    EX, PO = 2.0e5, 0.3
    xyz = g.Node[1:4].T
    U = (xyz @ G.T).reshape(-1)
    CU = O.element_coord_u(U, g.Elem, pp["Nind"])
    Ks = O.element_KS(C, CU, EX, PO) - O.element_K(C, EX, PO)      # K_sigma only
    lam, mu = EX * PO / ((1 + PO) * (1 - 2 * PO)), EX / (2 * (1 + PO))
    eps = 0.5 * (G + G.T)
    s = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    for e in range(C.shape[0]):
        v = (H @ C[e]).T                                            # (8, 3) nodal values of H x
        q = np.einsum('jp,jkpq,kq->', v, Ks[e], v)
        V = sum(np.linalg.det(C[e] @ O.shape_derv(O.RMAT[gp])) for gp in range(8))
        assert np.isclose(q, V * np.trace(H @ s @ H.T), rtol=1e-10, atol=0)


def test_element_KS_symmetric():
    g, pp, C = _elements(seed=5)
    U = np.random.default_rng(1).standard_normal(3 * g.Num_Node) * 1e-3
    K = O.element_KS(C, O.element_coord_u(U, g.Elem, pp["Nind"]), 2.0e5, 0.3)
    Kd = np.transpose(K, (0, 1, 3, 2, 4)).reshape(K.shape[0], 24, 24)
    assert np.allclose(Kd, np.transpose(Kd, (0, 2, 1)), rtol=0, atol=1e-9 * np.abs(Kd).max())


def test_spin_blocks_closed_forms():
    g = mesh.structured_block(3, 2, 2, 1.0, 0.1, 2)
    pp = O.pre_process(g.Node, g.Elem)
    en = O.element_nodes(g.Elem, pp["Nind"])
    DE = 7.8e-9
    Me = O.element_M(O.element_coords(g.Node, g.Elem, pp["Nind"]), DE)
    mb = O.assemble_M_block(Me, en, pp["SmNode"], pp["ReNode"])
    w = 100.0
    S = O.spin_blocks(mb, (0.0, 0.0, w))
    assert np.all(S[:, 2, :] == 0) and np.all(S[:, :, 2] == 0)
    # v = unit translation along x: v.K_spin.v = -w^2 * total mass
    V = sum(np.linalg.det(C @ O.shape_derv(O.RMAT[gp])) for C in O.element_coords(g.Node, g.Elem, pp["Nind"])
            for gp in range(8))
    assert np.isclose(S[:, 0, 0].sum(), -w * w * DE * V, rtol=1e-12)
    # a general axis: w w^T drops out along w
    wv = np.array([0.3, -0.5, 0.8]) * 50
    S2 = O.spin_blocks(mb, wv)
    assert np.allclose(np.einsum('bpq,q->bp', S2, wv), 0.0, atol=1e-12 * np.abs(S2).max())
