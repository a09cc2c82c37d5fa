"""Drop-in ``Static_solver`` (SOLVE_static.py:6-38, linear case) on the GPU.

``Static_solver(Model, 0).Solve_cpu()`` keeps the reference's name and
return value (U as an ndarray (3N,)).  The Pardiso factorisation is replaced
by one of the GPU forms of solvers.py, chosen by the extra keyword
``method`` ('pcg' FP64 block-Jacobi PCG, 'mixed' FP32-PCG inside FP64
refinement, 'learned' flexible PCG with the mesh-convolution
preconditioner).  ``tol`` is the FP64 relative residual target; the achieved
value is printed as the reference prints its Pardiso residual and kept in
``self.info``.  Large-displacement Newton (Qtype=1) is not part of this
package.
"""
from __future__ import annotations

from . import solvers


class Static_solver:
    def __init__(self, Model, Qtype=0):
        self.Qtype = Qtype
        self.Model = Model
        self.info = None

    def Solve_cpu(self, max_errs=1e-4, max_iters=20, method="pcg", tol=1e-8, verbose=False, **kw):
        if self.Qtype != 0:
            raise NotImplementedError("only the linear static case (Qtype=0) is implemented")
        M = self.Model
        F = M.F_device()
        K64 = solvers.BsrOperator(M.dev, M.K_bsr("f64"))
        if method == "pcg":
            U, info = solvers.solve_pcg64(K64, F, tol=tol, **kw)
        elif method == "mixed":
            K32 = solvers.BsrOperator(M.dev, M.K_bsr("f32"))
            U, info = solvers.solve_mixed(K64, K32, F, tol=tol, **kw)
        elif method == "learned":
            from .learned import solve_learned
            U, info = solve_learned(M, K64, F, tol=tol, **kw)
        else:
            raise ValueError(f"unknown method {method!r}")
        self.info = info
        self.U_device = U
        if verbose:
            print(f"\t\t\t{info.method}: {info.iterations} iterations, relative residual "
                  f"{info.rel_residual:.3e}, {info.seconds:.3f} s")
        return U.cpu().numpy()
