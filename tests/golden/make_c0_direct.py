"""Writes c0_direct.npz: the direct-solver answers for the reference's own
CPU-runnable case (cantilever 20x20x50 elements, 22,491 nodes, 67,473 DOF,
jitter 5% seed 0, z = 0 face fixed, -z load of 1000 N on the top face,
Omegas 0): K, M and F from the oracle (oracle/fem_oracle.py), the static
solution by SuperLU (scipy.sparse.linalg.splu, standing in for the
reference's Pardiso, SOLVE_static.py:31-33) and the 20 lowest eigenvalues of
(K, M) by shift-invert ARPACK (scipy.sparse.linalg.eigsh, sigma = 0; the
reference's Lanczos targets the same modes, SOLVE_eigen.py:45-89).  Kept:
the top-face displacements, |U|, |F|, the FP64 residual of the direct
solution and the eigenvalues.  Run from the repository root (about 4
minutes of CPU):  python tests/golden/make_c0_direct.py"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as sla

sys.path.insert(0, os.getcwd())
from oracle import fem_oracle as O            # noqa: E402
from fem_taichi_amd import mesh               # noqa: E402

g = mesh.cantilever_20x20x50()
pp = O.pre_process(g.Node, g.Elem)
K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
M = O.model_M_csr(g.Node, g.Elem, pp, g.DE)
F = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
U = sla.splu(K.tocsc()).solve(F)
lam = np.sort(sla.eigsh(K.tocsc(), k=20, M=M.tocsr(), sigma=0, which="LM", return_eigenvectors=False))
z = np.asarray(g.Node[-1], dtype=np.float64)
top = np.nonzero(z >= z.max() - 0.5)[0]
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c0_direct.npz"),
                    top_nodes=top.astype(np.int64), U_top=U.reshape(-1, 3)[top], U_norm=np.linalg.norm(U),
                    F_norm=np.linalg.norm(F), residual=O.static_residual(K, U, F), eig=lam)
print("top nodes", len(top), "|U|", np.linalg.norm(U), "residual", O.static_residual(K, U, F), "eig", lam[:4])
