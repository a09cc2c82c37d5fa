"""Pins the oracle's large-displacement matrices element_KN (UnitK,
FEM_model_bsr.py:476-598): zero displacement gives the linear K, rigid
translation of U changes nothing, KT is symmetric, and an independent closed
form in the deformation gradient F = I + du/dx matches.  CPU only."""
import numpy as np
import pytest

from oracle import fem_oracle as O
from fem_taichi_amd import mesh


def coords(seed=0, n=(2, 2, 2)):
    g = mesh.structured_block(*n, 1.0, 0.15, seed)
    pp = O.pre_process(g.Node, g.Elem)
    return g, pp, O.element_coords(g.Node, g.Elem, pp["Nind"])


def closed_form(C, Cu, EX, PO):
    """KT, KM from F = I + T, Fh = I + T/2, E = (F^T F - I)/2 and the stress
    S_g = lambda tr(E) I + 2 mu E summed over the Gauss points visited so far."""
    lam = EX * PO / ((1 + PO) * (1 - 2 * PO))
    mu = EX / (2 * (1 + PO))
    Ecount = C.shape[0]
    KT = np.zeros((Ecount, 8, 8, 3, 3))
    KM = np.zeros_like(KT)
    s = np.zeros((Ecount, 3, 3))
    I = np.eye(3)
    for g in range(8):
        DNr = O.shape_derv(O.RMAT[g])
        J = C @ DNr
        det = np.linalg.det(J)
        a = DNr[None] @ np.linalg.inv(J)                      # (E, 8, 3)
        T = Cu @ a
        F = I + T
        Fh = I + 0.5 * T
        Eg = 0.5 * (np.transpose(F, (0, 2, 1)) @ F - I)
        s = s + lam * np.trace(Eg, axis1=1, axis2=2)[:, None, None] * I + 2 * mu * Eg
        A = np.einsum('emp,ejp->ejm', F, a)                   # F a_j
        H = np.einsum('emp,ejp->ejm', Fh, a)
        d = np.einsum('ejp,ekp->ejk', a, a)
        FFt = F @ np.transpose(F, (0, 2, 1))
        FFh = F @ np.transpose(Fh, (0, 2, 1))
        geo = np.einsum('ejp,epq,ekq->ejk', a, s, a)
        w = det[:, None, None, None, None]
        KT += w * (lam * np.einsum('ejm,ekn->ejkmn', A, A) + mu * np.einsum('ekm,ejn->ejkmn', A, A)
                   + mu * d[..., None, None] * FFt[:, None, None]) + (det[:, None, None] * geo)[..., None, None] * I
        KM += w * (lam * np.einsum('ejm,ekn->ejkmn', A, H) + mu * np.einsum('ekm,ejn->ejkmn', A, H)
                   + mu * d[..., None, None] * FFh[:, None, None])
    return KT, KM


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_zero_displacement_is_linear_K():
    g, pp, C = coords(1)
    KT, KM = O.element_KN(C, np.zeros_like(C), 2.0e5, 0.3)
    K = O.element_K(C, 2.0e5, 0.3)
    assert rel(KT, K) < 1e-14 and rel(KM, K) < 1e-14


def test_closed_form_and_symmetry():
    g, pp, C = coords(2)
    rng = np.random.default_rng(3)
    Cu = 0.05 * rng.standard_normal(C.shape)
    KT, KM = O.element_KN(C, Cu, 2.0e5, 0.3)
    KTc, KMc = closed_form(C, Cu, 2.0e5, 0.3)
    assert rel(KT, KTc) < 1e-13
    assert rel(KM, KMc) < 1e-13
    assert rel(KT, np.transpose(KT, (0, 2, 1, 4, 3))) < 1e-14          # KT[j,k] = KT[k,j]^T
    assert rel(KM, np.transpose(KM, (0, 2, 1, 4, 3))) > 1e-6           # KM is not symmetric


def test_rigid_translation_invariance_and_per_element_E():
    g, pp, C = coords(4)
    rng = np.random.default_rng(5)
    Cu = 0.03 * rng.standard_normal(C.shape)
    E = 2.0e5 * np.exp(0.3 * rng.standard_normal(C.shape[0]))
    KT, KM = O.element_KN(C, Cu, E, 0.3)
    KT2, KM2 = O.element_KN(C, Cu + np.array([0.3, -0.2, 0.1])[None, :, None], E, 0.3)
    assert rel(KT2, KT) < 1e-12 and rel(KM2, KM) < 1e-12
    # K is linear in E element by element
    KTs, _ = O.element_KN(C[:1], Cu[:1], E[0], 0.3)
    assert rel(KT[:1], KTs) < 1e-14


def test_model_level_penalty_on_both():
    g, pp, C = coords(6, (2, 1, 3))
    U = 0.01 * np.random.default_rng(7).standard_normal(3 * g.Num_Node)
    VT, VM = O.model_KN_bsr_values(g.Node, g.Elem, pp, 2.0e5, 0.3, g.Bounds, U)
    pen = VT >= 1e19
    assert pen.any() and np.array_equal(VT >= 1e19, VM >= 1e19)
    assert np.all(VT[pen] == 1e20) and np.all(VM[pen] == 1e20)
