"""NumPy/SciPy restatement of the reference's hex8 FEM path (TEST ORACLE).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by the product package.

The functions below restate, from the reference source text, the CPU
algorithm of tianshao1992/FEM_taichi for the data-parallel path:

=====================  ==========================================  ==============================
oracle function        reference                                   pinned by (tests/)
=====================  ==========================================  ==============================
pre_process            READ_physics.py:42-61                       closed-form nnzb, brute force,
                                                                   2.cdb fixture invariants
csr_index_K / _M       FEM_model_csr.py:168-173, 196-201           integer closed forms, BSR->CSR
                                                                   identity (FEM_model_bsr.py:128)
element_K              FEM_model_csr.py:246-257, 287-330, 405-436  rigid-body null space, symmetry,
                                                                   scaling, independent closed form
element_M              FEM_model_csr.py:260-270, 333-365           box closed form, sum = rho*V
element_F              FEM_model_csr.py:214-229, 275-283           affine-box total force closed form
assemble_K_csr / _bsr  FEM_model_csr.py:150-165 / bsr:163-175      brute-force dense assembly
assemble_M_csr         FEM_model_csr.py:178-192                    brute-force dense assembly
modify_K               FEM_model_csr.py:233-242 / bsr:248-259      exact 1e20 set (not add)
vector_F               FEM_model_csr.py:141-146, 205-229           linearity, closed form
get_stress             POST_process.py:72-99, 172-223, 277-321     uniform-strain patch test (D,B);
                                                                   eigh; Q3: the extrapolation map
                                                                   against a hand-counted table and
                                                                   a unit cube with hand-derived
                                                                   Gauss-point stresses; Q4: hand
                                                                   principal triples
lanczos                SOLVE_eigen.py:45-89                        dense scipy eigh on small pencils,
                                                                   the SOLVE_eigen.py:101-117 pencil
element_KN             FEM_model_bsr.py:476-598 (UnitK, BN, S)     U = 0 gives element_K; rigid
                                                                   translation of U invariance;
                                                                   independent F-based closed form
element_KS             none (Readme.md:189 lists it missing):      U = 0 and rigid U give element_K;
                       linear prestress stiffening                 affine closed form
                                                                   v.K_sigma.v = V tr(H s H^T)
spin_blocks            none (Readme.md:189)                        zz blocks zero for w || z; a rigid
                                                                   x translation gives -w^2 rho V
=====================  ==========================================  ==============================

Known reference quirks restated on purpose (see the module docstrings of each
function): Q1 penalty *sets* 1e20; Q3 the Get_Stress extrapolation evaluates
``shapefunc_GS`` at RMat-ordered corners with shape-ordered Gauss weights;
Q4 the von Mises term ``(p2-p1)^2``; Q6 zero-padded T in block Lanczos; QN
the geometric stiffness of UnitK uses the element stress accumulated over the
Gauss points visited so far (S is an element field summed in the Gauss loop).

A patch test is blind to the extrapolation map of Q3 (constant fields pass
through any partition of unity), so Q3 is pinned by values counted by hand
from the two sign tables of POST_process.py:215-222 and :310-317, and on a
unit cube whose Gauss-point stresses are derived by hand; Q4 by principal
triples whose quirk value is computed by hand (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

# FEM_model_csr.py:93 -- the reference's truncated Gauss abscissa.
GS = 0.57735026918962

# FEM_model_csr.py:408-415 -- Gauss point order (xi, eta, zeta) signs.
RMAT_SIGNS = np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1],
                       [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]], dtype=np.float64)
RMAT = RMAT_SIGNS * GS

# FEM_model_csr.py:357-364 -- shape-function node order (signs of xi, eta, zeta).
NODE_SIGNS = np.array([[1, 1, -1], [-1, 1, -1], [-1, -1, -1], [1, -1, -1],
                       [1, 1, 1], [-1, 1, 1], [-1, -1, 1], [1, -1, 1]], dtype=np.float64)

# POST_process.py:310-317 -- "GMat" corners, in RMat sign order.
GMAT = RMAT_SIGNS.copy()


def shape_func(r):
    """FEM_model_csr.py:355-365: trilinear shape functions, (8,)."""
    r = np.asarray(r, dtype=np.float64)
    return 0.125 * np.prod(1.0 + NODE_SIGNS * r[None, :], axis=1)


def shape_derv(r):
    """FEM_model_csr.py:368-402: dN_k/dxi_d as an (8, 3) matrix."""
    r = np.asarray(r, dtype=np.float64)
    f = 1.0 + NODE_SIGNS * r[None, :]          # (8,3) factors
    d = np.empty((8, 3))
    d[:, 0] = 0.125 * NODE_SIGNS[:, 0] * f[:, 1] * f[:, 2]
    d[:, 1] = 0.125 * f[:, 0] * NODE_SIGNS[:, 1] * f[:, 2]
    d[:, 2] = 0.125 * f[:, 0] * f[:, 1] * NODE_SIGNS[:, 2]
    return d


def shapefunc_GS(r):
    """POST_process.py:212-223: shape functions with coordinates divided by GS."""
    r = np.asarray(r, dtype=np.float64) / GS
    return 0.125 * np.prod(1.0 + NODE_SIGNS * r[None, :], axis=1)


def dmat(EX, PO):
    """FEM_model_csr.py:418-432: the normal (D0) and shear (D1) 3x3 blocks.

    ``EX`` may be a scalar or a per-element array (extension: the reference
    has one scalar ``EX``); the result then carries a leading element axis.
    """
    EX = np.asarray(EX, dtype=np.float64)
    a = EX * (1 - PO) / (1 + PO) / (1 - 2 * PO)
    b = EX * PO / (1 + PO) / (1 - 2 * PO)
    c = EX / 2 / (1 + PO)
    one = np.ones(EX.shape + (3, 3))
    D0 = one * b[..., None, None]
    idx = np.arange(3)
    D0[..., idx, idx] = a[..., None]
    D1 = np.zeros(EX.shape + (3, 3))
    D1[..., idx, idx] = c[..., None]
    return D0, D1


# ----------------------------------------------------------------------------
# a1: sparsity pattern
# ----------------------------------------------------------------------------

def pre_process(Node, Elem):
    """READ_physics.py:42-61, restated with the same SciPy calls.

    ``Node``: (>=4, N) with row 0 = 1-based node IDs; ``Elem``: (>=8, E) whose
    last 8 rows are node ID - 1.  Returns a dict with Nind, SmNode, ReNode,
    CoNode, CoElem, Num_Node, Num_Elem, Num_MaxN.
    """
    Num_Node = Node.shape[-1]
    Num_Elem = Elem.shape[-1]
    Num_MaxN = int(Node[0, :].max())
    Nind = np.zeros((Num_MaxN,), dtype=np.int64)
    Nind[Node[0, :].astype(np.int64) - 1] = np.arange(Num_Node, dtype=np.int64)
    Row_Elem = np.arange(Num_Elem, dtype=np.int64).repeat(8)
    Col_Node = Nind[np.asarray(Elem[-8:]).T.flatten().astype(np.int64)]
    Val_Rela = np.ones_like(Row_Elem)
    Elem_Node = sp.coo_matrix((Val_Rela, (Row_Elem, Col_Node)),
                              shape=(Num_Elem, Num_Node)).tocsc()
    Node_Node = Elem_Node.T * Elem_Node
    Node_Node.sort_indices()
    return dict(Nind=Nind, SmNode=Node_Node.indptr, ReNode=Node_Node.indices,
                CoNode=Node_Node.indptr[1:] - Node_Node.indptr[:-1],
                CoElem=Elem_Node.indptr[1:] - Elem_Node.indptr[:-1],
                Num_Node=Num_Node, Num_Elem=Num_Elem, Num_MaxN=Num_MaxN)


def element_nodes(Elem, Nind):
    """(E, 8) node *indices* of each element: ``Nind[Elem[j, i]]`` (csr:155, 436)."""
    return Nind[np.asarray(Elem[-8:]).astype(np.int64)].T.copy()


def element_coords(Node, Elem, Nind):
    """FEM_model_csr.py:434-436: Coord[i][k, j] = Node[k, Nind[Elem[j, i]]] -> (E, 3, 8)."""
    en = element_nodes(Elem, Nind)
    xyz = np.asarray(Node[-3:], dtype=np.float64)          # (3, N)
    return np.transpose(xyz[:, en], (1, 0, 2)).copy()     # (E, 3, 8)


# ----------------------------------------------------------------------------
# a5 / a6 / a9: element matrices and vectors
# ----------------------------------------------------------------------------

def _local_B(DN):
    """FEM_model_csr.py:287-330: B1 (normal) and B2 (shear rows xy, yz, xz).

    DN: (..., 8, 3) physical gradients.  Returns B1, B2 of shape (..., 8, 3, 3).
    """
    sh = DN.shape[:-1] + (3, 3)
    B1 = np.zeros(sh)
    B2 = np.zeros(sh)
    for d in range(3):
        B1[..., d, d] = DN[..., d]
    B2[..., 0, 0] = DN[..., 1]; B2[..., 0, 1] = DN[..., 0]
    B2[..., 1, 1] = DN[..., 2]; B2[..., 1, 2] = DN[..., 1]
    B2[..., 2, 0] = DN[..., 2]; B2[..., 2, 2] = DN[..., 0]
    return B1, B2


def element_K(Coord, EX, PO):
    """FEM_model_csr.py:246-257 (UnitMatrixK): (E, 8, 8, 3, 3) element stiffness.

    K[j,k] = sum_g detJ_g (B1_j^T D0 B1_k + B2_j^T D1 B2_k), J = Coord @ DNr
    (no abs on detJ, unit Gauss weights).
    """
    Coord = np.asarray(Coord, dtype=np.float64)
    E = Coord.shape[0]
    D0, D1 = dmat(EX, PO)
    per_elem = D0.ndim == 3
    K = np.zeros((E, 8, 8, 3, 3))
    for g in range(8):
        DNr = shape_derv(RMAT[g])
        J = Coord @ DNr                                   # (E,3,3)
        Jinv = np.linalg.inv(J)
        det = np.linalg.det(J)
        DN = DNr[None] @ Jinv                             # (E,8,3)
        B1, B2 = _local_B(DN)
        if per_elem:
            t1 = np.einsum('ejrp,ers,ekst->ejkpt', B1, D0, B1, optimize=True)
            t2 = np.einsum('ejrp,ers,ekst->ejkpt', B2, D1, B2, optimize=True)
        else:
            t1 = np.einsum('ejrp,rs,ekst->ejkpt', B1, D0, B1, optimize=True)
            t2 = np.einsum('ejrp,rs,ekst->ejkpt', B2, D1, B2, optimize=True)
        K += det[:, None, None, None, None] * (t1 + t2)
    return K


def _local_B_bsr(DN):
    """FEM_model_bsr.py:303-347 (installLocalB): B1 (normal) and B2 with the
    shear rows in the BSR model's order yz, xz, xy.  (..., 8, 3, 3)."""
    sh = DN.shape[:-1] + (3, 3)
    B1 = np.zeros(sh)
    B2 = np.zeros(sh)
    for d in range(3):
        B1[..., d, d] = DN[..., d]
    B2[..., 0, 1] = DN[..., 2]; B2[..., 0, 2] = DN[..., 1]
    B2[..., 1, 0] = DN[..., 2]; B2[..., 1, 2] = DN[..., 0]
    B2[..., 2, 0] = DN[..., 1]; B2[..., 2, 1] = DN[..., 0]
    return B1, B2


def _local_BN(T, DN):
    """FEM_model_bsr.py:548-585 (InslocalBN): BN1[j][p, k] = T[k, p] DN[j, p];
    BN2 rows yz, xz, xy: T[k,2] DN[j,1] + T[k,1] DN[j,2], T[k,2] DN[j,0] +
    T[k,0] DN[j,2], T[k,1] DN[j,0] + T[k,0] DN[j,1].  T (E, 3, 3), DN (E, 8, 3)."""
    Tt = np.transpose(T, (0, 2, 1))[:, None]                        # Tt[e,0,p,k] = T[e,k,p]
    BN1 = DN[..., :, None] * Tt                                     # (E,8,3,3)
    BN2 = np.empty_like(BN1)
    BN2[..., 0, :] = T[:, None, :, 2] * DN[..., 1, None] + T[:, None, :, 1] * DN[..., 2, None]
    BN2[..., 1, :] = T[:, None, :, 2] * DN[..., 0, None] + T[:, None, :, 0] * DN[..., 2, None]
    BN2[..., 2, :] = T[:, None, :, 1] * DN[..., 0, None] + T[:, None, :, 0] * DN[..., 1, None]
    return BN1, BN2


def element_KN(Coord, CoordU, EX, PO):
    """FEM_model_bsr.py:476-505 (UnitK) after LDinit/initK (:588-598, :436-449):
    the large-displacement tangent KT and the matrix KM of the residual
    KM U - F, both (E, 8, 8, 3, 3).

    For each Gauss point g (RMat order): T = CoordU @ DN (displacement
    gradient), L_j = B_j + BN_j, for every node j the element stress
    accumulates S += D (B_j + BN_j / 2) u_j, then for every (j, k)
        KT[j,k] += detJ (L_j^T D L_k),  KM[j,k] += detJ (L_j^T D (B_k + BN_k / 2)),
    and, with s the 3x3 tensor of the stress accumulated so far (QN),
        KT[j,k] += detJ (DN_j^T s DN_k) I.
    ``EX`` may be a per-element array (extension; the reference has one EX)."""
    Coord = np.asarray(Coord, dtype=np.float64)
    CoordU = np.asarray(CoordU, dtype=np.float64)
    E = Coord.shape[0]
    D0, D1 = dmat(EX, PO)
    if D0.ndim == 2:
        D0 = np.broadcast_to(D0, (E, 3, 3))
        D1 = np.broadcast_to(D1, (E, 3, 3))
    KT = np.zeros((E, 8, 8, 3, 3))
    KM = np.zeros((E, 8, 8, 3, 3))
    S0 = np.zeros((E, 3))
    S1 = np.zeros((E, 3))                                            # yz, xz, xy
    u = np.transpose(CoordU, (0, 2, 1))                              # (E, 8, 3)
    for g in range(8):
        DNr = shape_derv(RMAT[g])
        J = Coord @ DNr
        det = np.linalg.det(J)
        DN = DNr[None] @ np.linalg.inv(J)                            # (E, 8, 3)
        T = CoordU @ DN                                              # (E, 3, 3)
        B1, B2 = _local_B_bsr(DN)
        BN1, BN2 = _local_BN(T, DN)
        H1, H2 = B1 + 0.5 * BN1, B2 + 0.5 * BN2
        S0 += np.einsum('eps,es->ep', D0, np.einsum('ejsk,ejk->es', H1, u))
        S1 += np.einsum('eps,es->ep', D1, np.einsum('ejsk,ejk->es', H2, u))
        L1, L2 = B1 + BN1, B2 + BN2
        lt1 = np.einsum('ejrp,ers->ejps', L1, D0)
        lt2 = np.einsum('ejrp,ers->ejps', L2, D1)
        AB = np.einsum('ejps,eksq->ejkpq', lt1, L1) + np.einsum('ejps,eksq->ejkpq', lt2, L2)
        AhB = np.einsum('ejps,eksq->ejkpq', lt1, H1) + np.einsum('ejps,eksq->ejkpq', lt2, H2)
        w = det[:, None, None, None, None]
        KT += w * AB
        KM += w * AhB
        s = np.zeros((E, 3, 3))                                      # MatrixS (:469-475)
        s[:, 0, 0], s[:, 1, 1], s[:, 2, 2] = S0[:, 0], S0[:, 1], S0[:, 2]
        s[:, 0, 1] = s[:, 1, 0] = S1[:, 2]
        s[:, 0, 2] = s[:, 2, 0] = S1[:, 1]
        s[:, 1, 2] = s[:, 2, 1] = S1[:, 0]
        geo = np.einsum('ejp,epq,ekq->ejk', DN, s, DN) * det[:, None, None]
        KT += geo[..., None, None] * np.eye(3)
    return KT, KM


def element_KS(Coord, CoordU, EX, PO):
    """Prestressed element stiffness of a linear displacement field (no
    counterpart in the reference, whose Readme.md:189 lists stress
    stiffening as missing): element_K + K_sigma, (E, 8, 8, 3, 3), with
        K_sigma[j, k] = sum_g detJ (DN_j^T s_g DN_k) I
    and s_g the Gauss point's own linear stress D B u (B of _local_B_bsr, the
    BSR model's strain rows; no accumulation over Gauss points)."""
    Coord = np.asarray(Coord, dtype=np.float64)
    CoordU = np.asarray(CoordU, dtype=np.float64)
    E = Coord.shape[0]
    D0, D1 = dmat(EX, PO)
    if D0.ndim == 2:
        D0 = np.broadcast_to(D0, (E, 3, 3))
        D1 = np.broadcast_to(D1, (E, 3, 3))
    K = element_K(Coord, EX, PO).copy()
    u = np.transpose(CoordU, (0, 2, 1))                              # (E, 8, 3)
    for g in range(8):
        DNr = shape_derv(RMAT[g])
        J = Coord @ DNr
        det = np.linalg.det(J)
        DN = DNr[None] @ np.linalg.inv(J)                            # (E, 8, 3)
        B1, B2 = _local_B_bsr(DN)
        S0 = np.einsum('eps,es->ep', D0, np.einsum('ejsk,ejk->es', B1, u))
        S1 = np.einsum('eps,es->ep', D1, np.einsum('ejsk,ejk->es', B2, u))     # yz, xz, xy
        s = np.zeros((E, 3, 3))
        s[:, 0, 0], s[:, 1, 1], s[:, 2, 2] = S0[:, 0], S0[:, 1], S0[:, 2]
        s[:, 0, 1] = s[:, 1, 0] = S1[:, 2]
        s[:, 0, 2] = s[:, 2, 0] = S1[:, 1]
        s[:, 1, 2] = s[:, 2, 1] = S1[:, 0]
        geo = np.einsum('ejp,epq,ekq->ejk', DN, s, DN) * det[:, None, None]
        K += geo[..., None, None] * np.eye(3)
    return K


def spin_blocks(Mb, OM):
    """Spin softening per node-pair block (no counterpart in the reference;
    Readme.md:189): -m_jk (|w|^2 I - w w^T) for the consistent-mass scalars
    Mb (nnzb,) -> (nnzb, 3, 3)."""
    w = np.asarray(OM, dtype=np.float64)
    W = np.dot(w, w) * np.eye(3) - np.outer(w, w)
    return -np.asarray(Mb, dtype=np.float64)[:, None, None] * W[None]


def element_coord_u(U, Elem, Nind):
    """FEM_model_bsr.py:443-446: CoordU[i][k, j] = U[Nind[Elem[j, i]]][k] -> (E, 3, 8)."""
    en = element_nodes(Elem, Nind)
    u = np.asarray(U, dtype=np.float64).reshape(-1, 3)
    return np.transpose(u[en], (0, 2, 1)).copy()


def element_M(Coord, DE):
    """FEM_model_csr.py:260-270 (UnitMatrixM): (E, 8, 8) consistent mass."""
    Coord = np.asarray(Coord, dtype=np.float64)
    M = np.zeros((Coord.shape[0], 8, 8))
    for g in range(8):
        Nr = shape_func(RMAT[g])
        J = Coord @ shape_derv(RMAT[g])
        det = DE * np.linalg.det(J)
        M += det[:, None, None] * np.outer(Nr, Nr)[None]
    return M


def unit_vector_F(R, OM):
    """FEM_model_csr.py:275-283: centrifugal acceleration term for positions R (...,3)."""
    w0, w1, w2 = OM
    F = np.empty_like(R)
    F[..., 0] = w1 * w1 * R[..., 0] + w2 * w2 * R[..., 0] - w0 * w1 * R[..., 1] - w0 * w2 * R[..., 2]
    F[..., 1] = w0 * w0 * R[..., 1] + w2 * w2 * R[..., 1] - w1 * w2 * R[..., 2] - w1 * w0 * R[..., 0]
    F[..., 2] = w1 * w1 * R[..., 2] + w0 * w0 * R[..., 2] - w2 * w0 * R[..., 0] - w2 * w1 * R[..., 1]
    return F


def element_F(Coord, DE, OM):
    """FEM_model_csr.py:214-229 (InstallVectorF): (E, 8, 3) element body force."""
    Coord = np.asarray(Coord, dtype=np.float64)
    Fe = np.zeros((Coord.shape[0], 8, 3))
    for g in range(8):
        Nr = shape_func(RMAT[g])
        DNr = shape_derv(RMAT[g])
        R = Coord @ Nr                                    # (E,3)
        J = Coord @ DNr
        F = unit_vector_F(R, OM)
        det = DE * np.linalg.det(J)
        Fe += (det[:, None, None] * Nr[None, :, None]) * F[:, None, :]
    return Fe


# ----------------------------------------------------------------------------
# a2 / a3 / a7 / a8: index layouts, slot search, scatter, penalty
# ----------------------------------------------------------------------------

def csr_index_K(SmNode, CoNode):
    """FEM_model_csr.py:168-173: RowK (3N+1)."""
    N = len(CoNode)
    RowK = np.zeros(3 * N + 1, dtype=np.int64)
    for p in range(3):
        RowK[3 * np.arange(N) + 1 + p] = CoNode * 3 * (p + 1) + SmNode[:-1] * 9
    return RowK


def csr_index_M(SmNode, CoNode):
    """FEM_model_csr.py:196-201: RowM (3N+1)."""
    N = len(CoNode)
    RowM = np.zeros(3 * N + 1, dtype=np.int64)
    for p in range(3):
        RowM[3 * np.arange(N) + 1 + p] = CoNode * (p + 1) + SmNode[:-1] * 3
    return RowM


def slot_map(en, SmNode, ReNode):
    """FEM_model_csr.py:158-160: m such that ReNode[SmNode[row] + m] == col.

    The reference searches linearly; ReNode rows are sorted, so the same ``m``
    is found with a global binary search on (row, col) keys.  Returns (E,8,8).
    """
    N = len(SmNode) - 1
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(SmNode))
    keys = rows * N + ReNode.astype(np.int64)
    r = en[:, :, None].astype(np.int64)
    c = en[:, None, :].astype(np.int64)
    pos = np.searchsorted(keys, r * N + c)
    assert np.all(keys[pos] == r * N + c)
    return pos - SmNode[r]


def assemble_K_csr(Ke, en, SmNode, ReNode, CoNode):
    """FEM_model_csr.py:150-165: ValK / ColK in the CSR layout of InstallIndexK.

    IndexK = 9*Sm[row] + 3*p*Co[row] + 3*m + q; ColK = 3*col + q.
    """
    nnzb = int(SmNode[-1])
    m = slot_map(en, SmNode, ReNode)                      # (E,8,8)
    row = np.broadcast_to(en[:, :, None], m.shape).astype(np.int64)
    col = np.broadcast_to(en[:, None, :], m.shape).astype(np.int64)
    ValK = np.zeros(9 * nnzb)
    ColK = np.zeros(9 * nnzb, dtype=np.int64)
    for p in range(3):
        for q in range(3):
            idx = 9 * SmNode[row] + 3 * p * CoNode[row] + 3 * m + q
            ValK += np.bincount(idx.ravel(), weights=Ke[..., p, q].ravel(), minlength=9 * nnzb)
            ColK[idx.ravel()] = (3 * col + q).ravel()
    return ValK, ColK


def assemble_K_bsr(Ke, en, SmNode, ReNode):
    """FEM_model_bsr.py:163-175: ValKT (nnzb, 3, 3), ValKT[Sm[row]+m] += K[j,k]."""
    nnzb = int(SmNode[-1])
    m = slot_map(en, SmNode, ReNode)
    row = np.broadcast_to(en[:, :, None], m.shape).astype(np.int64)
    slot = (SmNode[row] + m).ravel()
    V = np.zeros((nnzb, 3, 3))
    for p in range(3):
        for q in range(3):
            V[:, p, q] = np.bincount(slot, weights=Ke[..., p, q].ravel(), minlength=nnzb)
    return V


def assemble_M_block(Me, en, SmNode, ReNode):
    """The scalar per node-pair block that InstallMatrixM replicates 3x (csr:178-192)."""
    nnzb = int(SmNode[-1])
    m = slot_map(en, SmNode, ReNode)
    row = np.broadcast_to(en[:, :, None], m.shape).astype(np.int64)
    slot = (SmNode[row] + m).ravel()
    return np.bincount(slot, weights=Me.ravel(), minlength=nnzb)


def assemble_M_csr(Me, en, SmNode, ReNode, CoNode):
    """FEM_model_csr.py:178-192: ValM / ColM, IndexM = 3*Sm + p*Co + m, ColM = 3*col + p."""
    nnzb = int(SmNode[-1])
    m = slot_map(en, SmNode, ReNode)
    row = np.broadcast_to(en[:, :, None], m.shape).astype(np.int64)
    col = np.broadcast_to(en[:, None, :], m.shape).astype(np.int64)
    ValM = np.zeros(3 * nnzb)
    ColM = np.zeros(3 * nnzb, dtype=np.int64)
    for p in range(3):
        idx = 3 * SmNode[row] + p * CoNode[row] + m
        ValM += np.bincount(idx.ravel(), weights=Me.ravel(), minlength=3 * nnzb)
        ColM[idx.ravel()] = (3 * col + p).ravel()
    return ValM, ColM


def stage_bounds(Bounds):
    """FEM_model_csr.py:64-73: (4, nb) array [ID-1, c0, c1, c2]; Q7 default fixes Nind[0]."""
    tmp = []
    for ids, c in Bounds:
        ids = np.asarray(ids, dtype=np.float64)[None, :]
        cc = np.repeat(np.asarray(c, dtype=np.float64)[:, None], ids.shape[-1], axis=-1)
        tmp.append(np.concatenate((ids, cc), axis=0))
    if tmp:
        return np.concatenate(tmp, axis=-1)
    return np.zeros((4, 1))


def stage_forces(Forces, int_loads=False):
    """FEM_model_csr.py:75-84.  The reference casts load values to int (Q2,
    csr:78); ``int_loads=True`` restates that, the default keeps floats."""
    tmp = []
    for ids, v in Forces:
        ids = np.asarray(ids, dtype=np.float64)[None, :]
        vv = np.asarray(v, dtype=np.float64)
        if int_loads:
            vv = np.trunc(vv)
        vv = np.repeat(vv[:, None], ids.shape[-1], axis=-1)
        tmp.append(np.concatenate((ids, vv), axis=0))
    if tmp:
        return np.concatenate(tmp, axis=-1)
    return np.zeros((4, 1))


def modify_K_csr(ValK, bounds, Nind, SmNode, ReNode, CoNode):
    """FEM_model_csr.py:233-242: SET ValK at each constrained DOF's diagonal to 1e20 (Q1)."""
    ValK = ValK.copy()
    for i in range(bounds.shape[1]):
        idx = Nind[int(bounds[0, i])]
        s, e = SmNode[idx], SmNode[idx + 1]
        hit = np.nonzero(ReNode[s:e] == idx)[0]
        if len(hit) == 0:
            continue
        k = hit[0]
        for j in range(3):
            if bounds[j + 1, i] != -1:
                ValK[9 * SmNode[idx] + 3 * j * CoNode[idx] + k * 3 + j] = 1e20
    return ValK


def modify_K_bsr(V, bounds, Nind, SmNode, ReNode):
    """FEM_model_bsr.py:248-259: the same penalty on the BSR blocks."""
    V = V.copy()
    for i in range(bounds.shape[1]):
        idx = Nind[int(bounds[0, i])]
        s, e = SmNode[idx], SmNode[idx + 1]
        hit = np.nonzero(ReNode[s:e] == idx)[0]
        if len(hit) == 0:
            continue
        for j in range(3):
            if bounds[j + 1, i] != -1:
                V[s + hit[0], j, j] = 1e20
    return V


def vector_F(Fe, en, forces, Nind, N):
    """FEM_model_csr.py:141-146, 205-210, 214-229: F in DOF order 3*i + p."""
    ValF = np.zeros((3, N))
    for p in range(3):
        ValF[p] = np.bincount(en.ravel(), weights=Fe[..., p].ravel(), minlength=N)
    for i in range(forces.shape[1]):
        point = int(forces[0, i])
        ValF[:, Nind[point]] += forces[1:4, i]
    return ValF.T.flatten()


# ----------------------------------------------------------------------------
# whole-model convenience (mirrors Model.Get_MatrixKL / Get_MatrixM / Get_VectorF)
# ----------------------------------------------------------------------------

def model_K_csr(Node, Elem, pp, EX, PO, Bounds):
    """Model.Get_MatrixKL, FEM_model_csr.py:118-127."""
    en = element_nodes(Elem, pp['Nind'])
    Ke = element_K(element_coords(Node, Elem, pp['Nind']), EX, PO)
    ValK, ColK = assemble_K_csr(Ke, en, pp['SmNode'], pp['ReNode'], pp['CoNode'])
    ValK = modify_K_csr(ValK, stage_bounds(Bounds), pp['Nind'], pp['SmNode'], pp['ReNode'], pp['CoNode'])
    RowK = csr_index_K(pp['SmNode'], pp['CoNode'])
    n = 3 * pp['Num_Node']
    return sp.csr_matrix((ValK, ColK, RowK), shape=(n, n))


def model_K_bsr_values(Node, Elem, pp, EX, PO, Bounds):
    """FEM_model_bsr.py:123-131 values: (nnzb,3,3) after the penalty."""
    en = element_nodes(Elem, pp['Nind'])
    Ke = element_K(element_coords(Node, Elem, pp['Nind']), EX, PO)
    V = assemble_K_bsr(Ke, en, pp['SmNode'], pp['ReNode'])
    return modify_K_bsr(V, stage_bounds(Bounds), pp['Nind'], pp['SmNode'], pp['ReNode'])


def model_KN_bsr_values(Node, Elem, pp, EX, PO, Bounds, U):
    """Model.LDinit(U); Model.Get_MatrixKN() (FEM_model_bsr.py:134-145):
    (KT, KM) BSR values (nnzb, 3, 3), the penalty set on both (:248-259)."""
    en = element_nodes(Elem, pp['Nind'])
    KTe, KMe = element_KN(element_coords(Node, Elem, pp['Nind']), element_coord_u(U, Elem, pp['Nind']), EX, PO)
    b = stage_bounds(Bounds)
    out = []
    for Ke in (KTe, KMe):
        V = assemble_K_bsr(Ke, en, pp['SmNode'], pp['ReNode'])
        out.append(modify_K_bsr(V, b, pp['Nind'], pp['SmNode'], pp['ReNode']))
    return out[0], out[1]


def model_KS_bsr_values(Node, Elem, pp, EX, PO, Bounds, U, DE=None, OM=None):
    """K_L + K_sigma(U) (element_KS), minus the spin softening of OM when DE
    and OM are given, then the penalty: BSR values (nnzb, 3, 3)."""
    en = element_nodes(Elem, pp['Nind'])
    Ke = element_KS(element_coords(Node, Elem, pp['Nind']), element_coord_u(U, Elem, pp['Nind']), EX, PO)
    V = assemble_K_bsr(Ke, en, pp['SmNode'], pp['ReNode'])
    if OM is not None:
        Me = element_M(element_coords(Node, Elem, pp['Nind']), DE)
        V = V + spin_blocks(assemble_M_block(Me, en, pp['SmNode'], pp['ReNode']), OM)
    return modify_K_bsr(V, stage_bounds(Bounds), pp['Nind'], pp['SmNode'], pp['ReNode'])


def model_M_csr(Node, Elem, pp, DE):
    """Model.Get_MatrixM, FEM_model_csr.py:129-137."""
    en = element_nodes(Elem, pp['Nind'])
    Me = element_M(element_coords(Node, Elem, pp['Nind']), DE)
    ValM, ColM = assemble_M_csr(Me, en, pp['SmNode'], pp['ReNode'], pp['CoNode'])
    RowM = csr_index_M(pp['SmNode'], pp['CoNode'])
    n = 3 * pp['Num_Node']
    return sp.csr_matrix((ValM, ColM, RowM), shape=(n, n))


def model_F(Node, Elem, pp, DE, OM, Forces, int_loads=False):
    """Model.Get_VectorF, FEM_model_csr.py:141-146."""
    en = element_nodes(Elem, pp['Nind'])
    Fe = element_F(element_coords(Node, Elem, pp['Nind']), DE, np.asarray(OM, dtype=np.float64))
    return vector_F(Fe, en, stage_forces(Forces, int_loads), pp['Nind'], pp['Num_Node'])


# ----------------------------------------------------------------------------
# a13: stress recovery
# ----------------------------------------------------------------------------

def gauss_stress(Coord, EX, PO, Ue):
    """POST_process.py:190-209 (UnitVectorS): SVec (2, 8, E, 3).

    Coord (E,3,8); Ue (E,8,3) element nodal displacements.
    SVec[0,g] = sum_k D0 B1_k u_k ; SVec[1,g] = sum_k D1 B2_k u_k.
    """
    D0, D1 = dmat(EX, PO)
    per_elem = D0.ndim == 3
    E = Coord.shape[0]
    SVec = np.zeros((2, 8, E, 3))
    for g in range(8):
        DNr = shape_derv(RMAT[g])
        J = Coord @ DNr
        DN = DNr[None] @ np.linalg.inv(J)                 # (E,8,3)
        B1, B2 = _local_B(DN)                             # (E,8,3,3)
        e1 = np.einsum('ekpq,ekq->ep', B1, Ue)
        e2 = np.einsum('ekpq,ekq->ep', B2, Ue)
        if per_elem:
            SVec[0, g] = np.einsum('epq,eq->ep', D0, e1)
            SVec[1, g] = np.einsum('epq,eq->ep', D1, e2)
        else:
            SVec[0, g] = e1 @ D0.T
            SVec[1, g] = e2 @ D1.T
    return SVec


def extrapolation_weights():
    """POST_process.py:177-181: W[j, k] = shapefunc_GS(GMat[j])[k] (Q3 as written)."""
    return np.stack([shapefunc_GS(GMAT[j]) for j in range(8)])


def nodal_stress(Node, Elem, pp, EX, PO, U):
    """POST_process.py:172-187: the (N, 6) averaged nodal stress vector."""
    N = pp['Num_Node']
    Disp = np.asarray(U, dtype=np.float64).reshape(N, 3)
    en = element_nodes(Elem, pp['Nind'])
    Coord = element_coords(Node, Elem, pp['Nind'])
    SVec = gauss_stress(Coord, EX, PO, Disp[en])
    W = extrapolation_weights()
    Stress = np.zeros((2, N, 3))
    for j in range(8):
        contrib = np.einsum('k,skep->sep', W[j], SVec)     # (2,E,3)
        for s in range(2):
            for p in range(3):
                Stress[s, :, p] += np.bincount(en[:, j], weights=contrib[s, :, p], minlength=N)
    with np.errstate(divide='ignore', invalid='ignore'):
        Stress = Stress / pp['CoElem'][None, :, None]
    return Stress.transpose((1, 0, 2)).reshape((-1, 6))


def principal_vm(vector):
    """POST_process.py:80-95: principal stresses (descending) and the reference's
    von Mises term sqrt((p0-p1)^2 + (p1-p2)^2 + (p2-p1)^2) (Q4)."""
    n = vector.shape[0]
    tensor = np.zeros((n, 3, 3))
    tensor[:, 0, 0] = vector[:, 0]
    tensor[:, 1, 1] = vector[:, 1]
    tensor[:, 2, 2] = vector[:, 2]
    tensor[:, 0, 1] = tensor[:, 1, 0] = vector[:, 3]
    tensor[:, 1, 2] = tensor[:, 2, 1] = vector[:, 4]
    tensor[:, 0, 2] = tensor[:, 2, 0] = vector[:, 5]
    principle, _ = np.linalg.eigh(tensor)
    principle = principle[:, ::-1]
    vm = np.sqrt(np.square(principle[:, 0] - principle[:, 1]) + np.square(principle[:, 1] - principle[:, 2])
                 + np.square(principle[:, 2] - principle[:, 1]))[:, None]
    return principle, vm


def get_stress(Node, Elem, pp, EX, PO, U):
    """Post.Get_Stress, POST_process.py:72-99 -> (N, 10)."""
    vector = nodal_stress(Node, Elem, pp, EX, PO, U)
    principle, vm = principal_vm(vector)
    return np.concatenate((vector, principle, vm), axis=-1)


# ----------------------------------------------------------------------------
# a11: block Lanczos
# ----------------------------------------------------------------------------

def lanczos(A, B, solve, num_eigen=100, num_block=10, u0=None, rng=None):
    """SOLVE_eigen.py:45-89: shift-invert block Lanczos in the B inner product.

    ``solve(X)`` returns A^{-1} X for an (n, b) block.  ``u0`` is the (n, b)
    start block (the reference draws it unseeded with np.random.uniform,
    SOLVE_eigen.py:53); when None it is drawn from ``rng``.
    Returns (eigenvalues ascending [:num_eigen], vectors, T).
    """
    m = num_eigen * 2
    n = A.shape[0]
    b = int(num_block)
    Q = np.zeros((n, m))
    T = np.zeros((m, m))
    if u0 is None:
        rng = np.random.default_rng(0) if rng is None else rng
        u0 = rng.uniform(size=(n, b))
    u = np.array(u0, dtype=np.float64)
    v = sla.cholesky(u.T @ (B @ u))
    u = u @ sla.inv(v)
    Q[:, :b] = u
    steps = int(m / b)
    for i in range(steps):
        v = solve(B @ u)
        T[i * b:(i + 1) * b, i * b:(i + 1) * b] = u.T @ (B @ v)
        v = v - u @ T[i * b:(i + 1) * b, i * b:(i + 1) * b]
        if i >= steps - 1:
            break
        v = v - Q[:, :i * b + b] @ (Q[:, :i * b + b].T @ (B @ v))
        v = v - Q[:, :i * b + b] @ (Q[:, :i * b + b].T @ (B @ v))
        t = sla.cholesky(v.T @ (B @ v))
        T[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = t
        T[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = t.T
        u = v @ sla.inv(t)
        Q[:, (i + 1) * b:(i + 2) * b] = u
    Eval, Vect = sla.eigh(T)
    with np.errstate(divide='ignore'):
        Eval = np.ones_like(Eval) / Eval
    Q = Q @ Vect
    ind = np.argsort(Eval)
    return Eval[ind][:num_eigen], Q[:, ind][:, :num_eigen], T


def static_residual(K, U, F):
    """SOLVE_static.py:36-38: ||K U - F|| / ||F||."""
    return np.linalg.norm(K @ U - F) / np.linalg.norm(F)
