"""Pins the CPU oracle (oracle/fem_oracle.py) with closed forms, invariants,
brute force and an independent restatement -- each check chosen so that a
plausible mistake (dropped term, wrong sign or index, transposed operand)
fails it.  CPU only."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

E0, NU, RHO = 200000.0, 0.3, 7.8e-9


def k24(Ke):
    """(E,8,8,3,3) -> (E,24,24) with DOF order 3*j + p."""
    return Ke.transpose(0, 1, 3, 2, 4).reshape(Ke.shape[0], 24, 24)


def box_coord(a=1.0, b=2.0, c=3.0, origin=(0.0, 0.0, 0.0)):
    """(1,3,8) coordinates of an axis-aligned box in the reference's node order."""
    s = O.NODE_SIGNS
    xyz = (0.5 * (s + 1)) * np.array([a, b, c]) + np.array(origin)
    return xyz.T[None]


# ---------------------------------------------------------------- pattern (a1)

@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 4, 5), (2, 7, 3)])
def test_pattern_closed_form(dims):
    g = mesh.structured_block(*dims)
    pp = O.pre_process(g.Node, g.Elem)
    assert pp["SmNode"][-1] == mesh.nnzb_structured(*dims)
    # interior nodes have 27 neighbours, corners 8
    assert pp["CoNode"].max() == (27 if min(dims) >= 2 else pp["CoNode"].max())
    assert pp["CoNode"].min() == 8


def brute_force_pattern(en, N):
    nbrs = [set() for _ in range(N)]
    elems = [set() for _ in range(N)]
    for e, nodes in enumerate(en):
        for a in nodes:
            elems[a].add(e)
            for b in nodes:
                nbrs[a].add(b)
    return nbrs, elems


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pattern_brute_force_gapped_ids(seed):
    g = mesh.structured_block(3, 2, 4, 1.0, 0.1, seed)
    Node, Elem = mesh.permute_ids(g, seed=seed)
    pp = O.pre_process(Node, Elem)
    en = O.element_nodes(Elem, pp["Nind"])
    nbrs, elems = brute_force_pattern(en, Node.shape[1])
    for i in range(Node.shape[1]):
        row = pp["ReNode"][pp["SmNode"][i]:pp["SmNode"][i + 1]]
        assert list(row) == sorted(nbrs[i])
        assert pp["CoElem"][i] == len(elems[i])
    # symmetric pattern with the diagonal present
    n = Node.shape[1]
    A = sp.csr_matrix((np.ones(len(pp["ReNode"])), pp["ReNode"], pp["SmNode"]), shape=(n, n))
    assert (A != A.T).nnz == 0
    assert np.all(A.diagonal() == 1)


# ---------------------------------------------------------------- element K (a5)

def gram_form_K(Coord, E, nu):
    """Independent restatement: K_jk[p,q] = sum_g detJ E (lh a_p b_q + mh a_q b_p + mh d_pq a.b)."""
    lh = nu / ((1 + nu) * (1 - 2 * nu))
    mh = 1 / (2 * (1 + nu))
    out = np.zeros((Coord.shape[0], 8, 8, 3, 3))
    for g in range(8):
        DNr = O.shape_derv(O.RMAT[g])
        J = Coord @ DNr
        DN = DNr[None] @ np.linalg.inv(J)
        det = np.linalg.det(J)
        ab = np.einsum('ejp,ekq->ejkpq', DN, DN)
        dot = np.einsum('ejp,ekp->ejk', DN, DN)
        out += (det * E)[:, None, None, None, None] * (
            lh * ab + mh * ab.transpose(0, 1, 2, 4, 3) + mh * dot[..., None, None] * np.eye(3))
    return out


def jittered_coords(n=5, seed=0):
    g = mesh.structured_block(n, 1, 1, 1.0, 0.15, seed)
    pp = O.pre_process(g.Node, g.Elem)
    return O.element_coords(g.Node, g.Elem, pp["Nind"])


def test_K_symmetric_and_rigid_modes():
    C = jittered_coords()
    K = k24(O.element_K(C, E0, NU))
    assert np.abs(K - K.transpose(0, 2, 1)).max() / np.abs(K).max() < 1e-14
    for e in range(C.shape[0]):
        x = C[e].T                                  # (8,3)
        modes = []
        for t in np.eye(3):
            modes.append(np.tile(t, 8))
        for axis in np.eye(3):
            modes.append(np.cross(axis, x).reshape(-1))   # small rotation
        for u in modes:
            assert np.linalg.norm(K[e] @ u) < 1e-9 * np.linalg.norm(K[e]) * np.linalg.norm(u)
        ev = np.linalg.eigvalsh(K[e])
        assert np.sum(np.abs(ev) < 1e-10 * ev[-1]) == 6


def test_K_matches_independent_gram_form():
    C = jittered_coords(seed=3)
    Ee = np.linspace(1e5, 3e5, C.shape[0])
    a = O.element_K(C, Ee, NU)
    b = gram_form_K(C, Ee, NU)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-14


def test_K_scaling_linearity_rotation():
    C = jittered_coords(seed=4)
    K = O.element_K(C, E0, NU)
    tol = 1e-12 * np.abs(K).max()
    assert np.allclose(O.element_K(2.5 * C, E0, NU), 2.5 * K, rtol=0, atol=2.5 * tol)   # K ~ length^1
    assert np.allclose(O.element_K(C, 3 * E0, NU), 3 * K, rtol=0, atol=3 * tol)
    R = sla.expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.5], [-0.2, 0.5, 0]]))
    Kr = O.element_K(np.einsum('ab,ebk->eak', R, C), E0, NU)
    assert np.allclose(Kr, np.einsum('pa,ejkab,qb->ejkpq', R, K, R), rtol=0, atol=1e-10 * np.abs(K).max())


def test_K_box_uniaxial_energy():
    """A box under uniform uniaxial strain eps_xx stores energy 0.5 * V * (lambda+2mu) eps^2."""
    a, b, c = 1.0, 2.0, 3.0
    C = box_coord(a, b, c)
    K = k24(O.element_K(C, E0, NU))[0]
    eps = 1e-3
    u = np.zeros((8, 3))
    u[:, 0] = eps * C[0, 0]
    energy = 0.5 * u.reshape(-1) @ K @ u.reshape(-1)
    lam = E0 * NU / ((1 + NU) * (1 - 2 * NU))
    mu = E0 / (2 * (1 + NU))
    assert energy == pytest.approx(0.5 * a * b * c * (lam + 2 * mu) * eps ** 2, rel=1e-12)


# ---------------------------------------------------------------- M (a6), F (a9)

def test_M_box_closed_form():
    a, b, c = 1.0, 2.0, 3.0
    M = O.element_M(box_coord(a, b, c), RHO)[0]
    s = O.NODE_SIGNS
    ref = np.array([[RHO * a * b * c / 216 * np.prod(np.where(s[i] == s[j], 2.0, 1.0)) for j in range(8)]
                    for i in range(8)])
    assert np.allclose(M, ref, rtol=1e-12, atol=0)


def test_M_total_is_rho_volume():
    C = jittered_coords(seed=5)
    M = O.element_M(C, RHO)
    # volume by the divergence-free Gauss rule = sum_g detJ (exact for trilinear maps)
    vol = sum(np.linalg.det(C @ O.shape_derv(O.RMAT[g])) for g in range(8))
    assert np.allclose(M.sum(axis=(1, 2)), RHO * vol, rtol=1e-12)


def test_F_centrifugal_affine_box():
    """Total centrifugal load of a box spinning about z: rho w^2 V (xc, yc, 0)."""
    a, b, c = 1.0, 2.0, 3.0
    origin = (4.0, -1.0, 2.0)
    Fe = O.element_F(box_coord(a, b, c, origin), RHO, np.array([0.0, 0.0, 100.0]))[0]
    V = a * b * c
    xc, yc = origin[0] + a / 2, origin[1] + b / 2
    assert np.allclose(Fe.sum(axis=0), RHO * 1e4 * V * np.array([xc, yc, 0.0]), rtol=1e-12, atol=1e-20)
    # general omega: f = |w|^2 R - (w.R) w
    w = np.array([10.0, -20.0, 30.0])
    Fe = O.element_F(box_coord(a, b, c, origin), RHO, w)[0]
    R = np.array(origin) + np.array([a, b, c]) / 2
    assert np.allclose(Fe.sum(axis=0), RHO * V * (w @ w * R - (w @ R) * w), rtol=1e-12)


def test_F_nodal_loads_linear():
    g = mesh.structured_block(2, 2, 2, 1.0, 0.05, 0)
    pp = O.pre_process(g.Node, g.Elem)
    f1 = O.model_F(g.Node, g.Elem, pp, RHO, (0, 0, 0), g.Forces)
    assert f1.reshape(-1, 3)[:, 2].sum() == pytest.approx(-1000.0, rel=1e-12)
    f2 = O.model_F(g.Node, g.Elem, pp, RHO, (0, 0, 0), [[g.Forces[0][0], [0, 0, -2000.0 / 9]]])
    assert np.allclose(f2, 2 * f1)
    fi = O.model_F(g.Node, g.Elem, pp, RHO, (0, 0, 0), g.Forces, int_loads=True)
    # FEM_model_csr.py:78 casts -1000/9 = -111.1 to int -> -111 on each of the 9 nodes
    assert fi.reshape(-1, 3)[:, 2].sum() == -111.0 * 9


# ---------------------------------------------------------------- layouts (a2, a3, a7, a8)

def test_csr_layout_matches_bsr_and_dense():
    g = mesh.structured_block(2, 3, 2, 1.0, 0.1, 2, E_sigma=0.2)
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    V = O.model_K_bsr_values(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    n = 3 * g.Num_Node
    B = sp.bsr_matrix((V, pp["ReNode"], pp["SmNode"]), shape=(n, n)).tocsr()
    assert np.array_equal(B.indptr, K.indptr) and np.array_equal(B.indices, K.indices)
    assert np.array_equal(B.data, K.data)
    # dense brute-force assembly
    en = O.element_nodes(g.Elem, pp["Nind"])
    Ke = k24(O.element_K(O.element_coords(g.Node, g.Elem, pp["Nind"]), g.EX, g.PO))
    D = np.zeros((n, n))
    for e in range(en.shape[0]):
        dofs = (3 * en[e][:, None] + np.arange(3)).reshape(-1)
        D[np.ix_(dofs, dofs)] += Ke[e]
    Kd = K.toarray()
    pen = Kd >= 1e19
    assert np.allclose(Kd[~pen], D[~pen], rtol=0, atol=1e-9 * np.abs(D).max())
    assert np.all(Kd[pen] == 1e20)
    fixed = np.asarray(g.Bounds[0][0])
    assert pen.sum() == 3 * len(fixed)


def test_rowk_closed_form():
    g = mesh.structured_block(2, 2, 3)
    pp = O.pre_process(g.Node, g.Elem)
    RowK = O.csr_index_K(pp["SmNode"], pp["CoNode"])
    assert np.array_equal(np.diff(RowK), np.repeat(3 * pp["CoNode"], 3))
    RowM = O.csr_index_M(pp["SmNode"], pp["CoNode"])
    assert np.array_equal(np.diff(RowM), np.repeat(pp["CoNode"], 3))


def test_penalty_sets_not_adds_and_skips_free_flags():
    g = mesh.structured_block(2, 2, 2, 1.0, 0.0, 0)
    pp = O.pre_process(g.Node, g.Elem)
    en = O.element_nodes(g.Elem, pp["Nind"])
    Ke = O.element_K(O.element_coords(g.Node, g.Elem, pp["Nind"]), E0, NU)
    V, _ = O.assemble_K_csr(Ke, en, pp["SmNode"], pp["ReNode"], pp["CoNode"])
    bounds = np.array([[0.0, 4.0], [0.0, -1.0], [5.0, 0.0], [-1.0, -1.0]])   # node IDs-1 = 0, 4
    V2 = O.modify_K_csr(V, bounds, pp["Nind"], pp["SmNode"], pp["ReNode"], pp["CoNode"])
    changed = np.nonzero(V2 != V)[0]
    assert len(changed) == 3 and np.all(V2[changed] == 1e20)
    V3 = O.modify_K_csr(V2, bounds, pp["Nind"], pp["SmNode"], pp["ReNode"], pp["CoNode"])
    assert np.array_equal(V2, V3)


# ---------------------------------------------------------------- stress (a13)

def test_stress_uniform_strain_patch():
    """A linear displacement field gives sigma = D eps at every node (pins D, B
    and the nodal averaging; blind to the extrapolation map by construction)."""
    g = mesh.structured_block(3, 3, 3, 1.0, 0.12, 9)
    pp = O.pre_process(g.Node, g.Elem)
    eps = np.array([[1e-3, 2e-4, -3e-4], [2e-4, -5e-4, 1e-4], [-3e-4, 1e-4, 4e-4]])
    x = g.Node[1:].T
    u = x @ eps.T
    S = O.get_stress(g.Node, g.Elem, pp, E0, NU, u.reshape(-1))
    lam = E0 * NU / ((1 + NU) * (1 - 2 * NU))
    mu = E0 / (2 * (1 + NU))
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    ref = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[1, 2], sig[0, 2]])
    assert np.allclose(S[:, :6], ref[None], rtol=1e-10, atol=1e-9)
    ev = np.linalg.eigvalsh(sig)[::-1]
    assert np.allclose(S[:, 6:9], ev[None], rtol=1e-10, atol=1e-9)
    vm = np.sqrt((ev[0] - ev[1]) ** 2 + (ev[1] - ev[2]) ** 2 + (ev[2] - ev[1]) ** 2)
    assert np.allclose(S[:, 9], vm, rtol=1e-10)


# Q3 by hand.  POST_process.py:177-181 gives node j the weight
# Ns[k] = shapefunc_GS(GMat[j])[k] on SVec[k] (the stress at Gauss point
# RMat[k]).  shapefunc_GS (:215-222) writes factor d of Ns[k] as
# (1 +- rtr[d]/GS) with the sign of the shape-node table, and GMat[j]
# (:310-317) lists the corners (-1,-1,-1), (-1,-1,+1), (-1,+1,-1), ... So
# Ns[k] = (1 + 1/GS)^m (1 - 1/GS)^(3-m) / 8, m = the number of axes on which
# GMat[j] and the k-th sign row of shapefunc_GS agree.  The m table below was
# counted by hand from those two source tables:
#   shapefunc_GS rows   k0 (+,+,-) k1 (-,+,-) k2 (-,-,-) k3 (+,-,-)
#                       k4 (+,+,+) k5 (-,+,+) k6 (-,-,+) k7 (+,-,+)
_Q3_AGREE = np.array([
    [1, 2, 3, 2, 0, 1, 2, 1],   # GMat[0] (-,-,-)
    [0, 1, 2, 1, 1, 2, 3, 2],   # GMat[1] (-,-,+)
    [2, 3, 2, 1, 1, 2, 1, 0],   # GMat[2] (-,+,-)
    [1, 2, 1, 0, 2, 3, 2, 1],   # GMat[3] (-,+,+)
    [2, 1, 2, 3, 1, 0, 1, 2],   # GMat[4] (+,-,-)
    [1, 0, 1, 2, 2, 1, 2, 3],   # GMat[5] (+,-,+)
    [3, 2, 1, 2, 2, 1, 0, 1],   # GMat[6] (+,+,-)
    [2, 1, 0, 1, 3, 2, 1, 2],   # GMat[7] (+,+,+)
])


def _q3_by_hand():
    a, b = 1.0 + 1.0 / 0.57735026918962, 1.0 - 1.0 / 0.57735026918962
    return a ** _Q3_AGREE * b ** (3 - _Q3_AGREE) / 8.0


def test_extrapolation_weights_quirk():
    """Q3: W[j] = shapefunc_GS(GMat[j]) equals the hand-counted table entry for
    entry, e.g. W[0, 2] = (1 + 1/GS)^3 / 8 (GMat[0] and row k2 agree on all
    three axes).  Swapping GMat for the shape-node corners, or evaluating
    shapefunc_GS in RMat order, changes rows and fails here."""
    W = O.extrapolation_weights()
    hand = _q3_by_hand()
    assert np.allclose(W, hand, rtol=1e-14, atol=0)
    gs = 0.57735026918962
    assert np.isclose(W[0, 2], (1 + 1 / gs) ** 3 / 8, rtol=1e-14)
    assert np.isclose(W[7, 4], (1 + 1 / gs) ** 3 / 8, rtol=1e-14)
    assert np.isclose(W[7, 2], (1 - 1 / gs) ** 3 / 8, rtol=1e-14)
    assert np.isclose(W[0, 0], (1 + 1 / gs) * (1 - 1 / gs) ** 2 / 8, rtol=1e-14)
    # the plausible "fixes" differ from what the source computes
    shape_corner = np.array([[0.125 * np.prod(1 + O.NODE_SIGNS[k] * O.NODE_SIGNS[j] / gs) for k in range(8)]
                             for j in range(8)])
    rmat_order = np.array([[0.125 * np.prod(1 + O.RMAT_SIGNS[k] * O.RMAT_SIGNS[j] / gs) for k in range(8)]
                           for j in range(8)])
    assert not np.allclose(hand, shape_corner) and not np.allclose(hand, rmat_order)
    assert np.allclose(hand.sum(axis=1), 1.0)


def test_extrapolation_one_element_probe():
    """Q3 through nodal_stress on one element: CoElem = 1 at every node, so
    node Elem[j] receives sum_k W[j, k] SVec[k], with SVec the Gauss-point
    stresses.  A stress field that is a unit pulse at one Gauss point k is
    reproduced as column k of the hand table (probed through the linear map
    SVec -> nodal stress: the displacement is chosen so SVec is known)."""
    g = mesh.structured_block(1, 1, 1, 1.0, 0.15, 4)
    pp = O.pre_process(g.Node, g.Elem)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(3 * 8) * 1e-3
    en = O.element_nodes(g.Elem, pp["Nind"])
    Coord = O.element_coords(g.Node, g.Elem, pp["Nind"])
    SVec = O.gauss_stress(Coord, E0, NU, u.reshape(8, 3)[en])        # (2, 8, 1, 3)
    S = O.nodal_stress(g.Node, g.Elem, pp, E0, NU, u)                 # (8, 6)
    hand = _q3_by_hand()
    sv = np.concatenate([SVec[0, :, 0, :], SVec[1, :, 0, :]], axis=1)  # (8 Gauss points, 6)
    expect = np.zeros((8, 6))
    expect[en[0]] = hand @ sv
    assert np.allclose(S, expect, rtol=1e-12, atol=1e-12 * np.abs(sv).max())


def test_extrapolation_unit_cube_by_hand():
    """Q3 end to end with Gauss-point stresses derived by hand.  Unit cube
    [0,1]^3, local node k at the corner with the signs of the shape-node
    table (so x = (1 + xi)/2, y = (1 + eta)/2, z = (1 + zeta)/2), and
    u = (x y z, 0, 0), which the trilinear space holds exactly.  Then
    eps_xx = y z, gamma_xy = x z, gamma_xz = x y, and at the Gauss point g of
    FEM_model_csr.py:408-415 (signs (-,-,-), (-,-,+), (-,+,-), ... : bit 2 of
    g is xi, bit 1 eta, bit 0 zeta) sigma = ((l+2m) yz, l yz, l yz, m xz, 0,
    m xy).  Node k must get sum_g hand[k, g] sigma_g: this pins the Gauss
    order of SVec, the shear order and D together with the Q3 table."""
    g = mesh.structured_block(1, 1, 1, 1.0, 0.0, 0)
    pp = O.pre_process(g.Node, g.Elem)
    x = g.Node[1:].T
    u = np.zeros((8, 3))
    u[:, 0] = x[:, 0] * x[:, 1] * x[:, 2]
    S = O.get_stress(g.Node, g.Elem, pp, E0, NU, u.reshape(-1))
    lam = E0 * NU / ((1 + NU) * (1 - 2 * NU))
    mu = E0 / (2 * (1 + NU))
    gs = 0.57735026918962
    sig = np.zeros((8, 6))
    for gp in range(8):
        xi = [gs if (gp >> b) & 1 else -gs for b in (2, 1, 0)]
        px, py, pz = [(1 + s) / 2 for s in xi]
        sig[gp] = [(lam + 2 * mu) * py * pz, lam * py * pz, lam * py * pz, mu * px * pz, 0.0, mu * px * py]
    en = O.element_nodes(g.Elem, pp["Nind"])[0]
    expect = np.zeros((8, 6))
    expect[en] = _q3_by_hand() @ sig
    assert np.allclose(S[:, :6], expect, rtol=1e-11, atol=1e-9 * np.abs(sig).max())


def test_von_mises_quirk_by_hand():
    """Q4 (POST_process.py:93-94): sqrt((p0-p1)^2 + (p1-p2)^2 + (p2-p1)^2)
    with no 1/2, on hand-computed principal triples: (3, 2, 0) -> sqrt(1 + 4 +
    4) = 3 (true von Mises: sqrt(7)); (s, 0, 0) -> s (true: s); (5, 1, -1) ->
    sqrt(16 + 4 + 4) = sqrt(24) (true: sqrt(28)).  The tensors are rotated
    so the eigen-solve is exercised; principal values come out descending."""
    c, s = np.cos(0.7), np.sin(0.7)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    c2, s2 = np.cos(-0.4), np.sin(-0.4)
    Rx = np.array([[1, 0, 0], [0, c2, -s2], [0, s2, c2]])
    R = Rz @ Rx
    cases = [((3.0, 2.0, 0.0), 3.0), ((7.5, 0.0, 0.0), 7.5), ((5.0, 1.0, -1.0), np.sqrt(24.0)),
             ((0.0, 2.0, 3.0), 3.0)]
    vec = []
    for p, _ in cases:
        T = R @ np.diag(p) @ R.T
        vec.append([T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[1, 2], T[0, 2]])
    principle, vm = O.principal_vm(np.array(vec))
    for i, (p, v) in enumerate(cases):
        assert np.allclose(principle[i], sorted(p, reverse=True), atol=1e-12)
        assert np.isclose(vm[i, 0], v, rtol=1e-12)
    assert not np.isclose(vm[0, 0], np.sqrt(7.0))


# ---------------------------------------------------------------- Lanczos (a11)

def test_lanczos_small_dense():
    rng = np.random.default_rng(0)
    n = 60
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    B = np.diag(rng.uniform(1, 2, n))
    lam, X, T = O.lanczos(A, B, lambda X: np.linalg.solve(A, X), num_eigen=30, num_block=6)
    ref = sla.eigh(A, B, eigvals_only=True)
    assert np.allclose(lam[:5], ref[:5], rtol=1e-8)
    assert np.allclose(X.T @ B @ X, np.eye(X.shape[1]), atol=1e-8)


def test_lanczos_reference_pencil():
    """The tridiagonal A of SOLVE_eigen.py:101-117 (diagonal i+2, off-diagonal
    -(i+1)/2; diagonally dominant, so SPD).  That smoke block's B (diagonal
    i+1, off-diagonal +(i+1)/2) is not positive definite (its leading minor of
    order 4 fails Cholesky), so only B's diagonal is kept; dense eigh is the
    known answer."""
    N = 300
    off = (np.arange(1, N) + 1) / 2.0
    A = np.diag(np.arange(N) + 2.0) - np.diag(off, 1) - np.diag(off, -1)
    B = np.diag(np.arange(N) + 1.0)
    ref = sla.eigh(A, B, eigvals_only=True)
    lam, X, _ = O.lanczos(A, B, lambda X: np.linalg.solve(A, X), num_eigen=20, num_block=5)
    assert np.allclose(lam[:3], ref[:3], rtol=1e-6)
    assert np.allclose(X.T @ B @ X, np.eye(X.shape[1]), atol=1e-8)


def test_lanczos_zero_padding_quirk():
    """Q6: when num_block does not divide 2*num_eigen the tail of T stays zero
    and those Ritz values come out as inf (sorted last)."""
    rng = np.random.default_rng(1)
    n = 40
    A = np.diag(rng.uniform(1, 10, n))
    B = np.eye(n)
    lam, X, T = O.lanczos(A, B, lambda X: np.linalg.solve(A, X), num_eigen=7, num_block=4)
    assert T.shape == (14, 14) and np.all(T[12:, :] == 0) and np.all(T[:, 12:] == 0)
    theta = np.linalg.eigvalsh(T)
    assert np.sum(theta == 0) == 2          # -> 1/0 = inf Ritz values, sorted past num_eigen
    assert np.all(np.isfinite(lam))


# ---------------------------------------------------------------- C restatement

def test_c_port_matches_numpy_oracle():
    from oracle import c_oracle
    try:
        c_oracle.lib()
    except ImportError:
        pytest.skip("oracle/libfem_oracle.so not built")
    g = mesh.structured_block(4, 3, 5, 1.0, 0.1, 1, E_sigma=0.3)
    pp = O.pre_process(g.Node, g.Elem)
    en = O.element_nodes(g.Elem, pp["Nind"])
    xyz = g.Node[1:].T
    val, col = c_oracle.assemble_K_csr(xyz, en, g.EX, 0, NU, pp["SmNode"], pp["CoNode"], pp["ReNode"])
    Ke = O.element_K(O.element_coords(g.Node, g.Elem, pp["Nind"]), g.EX, NU)
    vr, cr = O.assemble_K_csr(Ke, en, pp["SmNode"], pp["ReNode"], pp["CoNode"])
    assert np.linalg.norm(val - vr) / np.linalg.norm(vr) < 1e-14
    assert np.array_equal(col, cr)
    Mv = c_oracle.assemble_M_csr(xyz, en, RHO, pp["SmNode"], pp["CoNode"], pp["ReNode"])
    Mr, _ = O.assemble_M_csr(O.element_M(O.element_coords(g.Node, g.Elem, pp["Nind"]), RHO), en,
                             pp["SmNode"], pp["ReNode"], pp["CoNode"])
    assert np.linalg.norm(Mv - Mr) / np.linalg.norm(Mr) < 1e-14
