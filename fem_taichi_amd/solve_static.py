"""Drop-in ``Static_solver`` (SOLVE_static.py:6-38, linear case) on the GPU.

``Static_solver(Model, 0).Solve_cpu()`` keeps the reference's name and
return value (U as an ndarray (3N,)).  The Pardiso factorisation is replaced
by one of the GPU forms, chosen by the extra keyword ``method`` ('pcg' FP64
block-Jacobi PCG, 'mixed' FP32-PCG inside FP64 refinement, 'mg' FP64
flexible CG preconditioned by an FP32 geometric multigrid V-cycle on
structured grids (multigrid.py), 'learned' flexible PCG with the
mesh-convolution preconditioner).  ``tol`` is the FP64 relative residual target; the achieved
value is printed as the reference prints its Pardiso residual and kept in
``self.info``.

``Qtype=1`` is the reference's large-displacement Newton loop
(SOLVE_static.py:41-76): at each step LDinit(U), (KT, KM) = Get_MatrixKN(),
residual RF = KM U - F, stop when ||RF|| / ||F|| <= max_errs or after
max_iters steps.  The update solves KT dU = RF (FP64 block-Jacobi PCG on the
GPU) and sets U = U - dU.  The reference passes F instead of RF to its
solve (SOLVE_static.py:70), which from U = 0 steps to U = -K^-1 F and does
not converge; the residual form is used here and ``self.history`` keeps the
residual of every step.  The final tangent KT stays on the model
(``Model.KT_device``) for a following modal solve, which then uses it as
the reference's Eigen_solver uses ``Model.KT``.
"""
from __future__ import annotations

import torch

from . import solvers
from .runtime import to_host


class Static_solver:
    def __init__(self, Model, Qtype=0):
        self.Qtype = Qtype
        self.Model = Model
        self.info = None

    def Solve_cpu(self, max_errs=1e-4, max_iters=20, method="pcg", tol=1e-8, verbose=False, **kw):
        if self.Qtype == 1:
            return self._solve_newton(max_errs, max_iters, tol, verbose)
        if self.Qtype != 0:
            raise ValueError(f"unknown Qtype {self.Qtype!r}")
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
        elif method == "mg":
            from .multigrid import solve_mg
            U, info = solve_mg(M, K64, F, tol=tol, **kw)
        else:
            raise ValueError(f"unknown method {method!r}")
        self.info = info
        self.U_device = U
        if verbose:
            print(f"\t\t\t{info.method}: {info.iterations} iterations, relative residual "
                  f"{info.rel_residual:.3e}, {info.seconds:.3f} s")
        return to_host(U)

    def _solve_newton(self, max_errs, max_iters, tol, verbose):
        M = self.Model
        F = M.F_device()
        nf = float(torch.linalg.vector_norm(F))
        U = torch.zeros_like(F)
        self.history = []
        step = 0
        while True:
            M.LDinit(U)
            kt, km = M.assemble_KN("bsr")
            KT = solvers.BsrOperator(M.dev, kt)
            KM = solvers.BsrOperator(M.dev, km)
            RF = KM.matvec(U) - F
            err = float(torch.linalg.vector_norm(RF)) / nf if nf > 0 else 0.0
            self.history.append(err)
            if verbose:
                print(f"\t\t\tNewton step {step + 1}: relative residual {err:.3e}")
            if err <= max_errs or step == max_iters:
                break
            dU, info = solvers.solve_pcg64(KT, RF, tol=tol)
            U = U - dU
            step += 1
        self.U_device = U
        # the final tangent stays on the model: the modal solve uses it, as the
        # reference's Eigen_solver factors Model.KT (SOLVE_eigen.py:32)
        M.KT_device = kt
        self.info = {"steps": step, "residual": self.history[-1]}
        return to_host(U)
