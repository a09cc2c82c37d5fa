"""GPU static solvers on the device-resident BSR(3x3) stiffness.

These replace the Pardiso factor/solve of SOLVE_static.py:31-38 (which is
kept only as the CPU oracle).  Every form is judged by the FP64 relative
residual ||F - K U|| / ||F|| on the penalised K, as the reference prints it
(SOLVE_static.py:36-38):

* ``pcg``   -- FP64 block-Jacobi (3x3 node blocks) preconditioned CG;
* ``mixed`` -- FP64 iterative refinement around an FP32 block-Jacobi PCG on
  an FP32 copy of K: the residual is formed in FP64 each outer step, the
  correction is solved in FP32 to a loose relative tolerance;
* ``learned`` -- see learned.py (flexible PCG with the mesh-convolution
  preconditioner and block-Jacobi fallback).
"""
from __future__ import annotations

import dataclasses
import time

import torch

from . import _lib
from .runtime import device, stream_ptr


@dataclasses.dataclass
class SolveInfo:
    method: str
    iterations: int
    outer: int
    rel_residual: float          # FP64 ||F - K U|| / ||F||, computed afresh
    seconds: float
    history: list


class BsrOperator:
    """A device BSR matrix on a mesh's pattern (f64 or f32 values)."""

    def __init__(self, dev_mesh, vals: torch.Tensor):
        self.mesh = dev_mesh
        self.vals = vals
        self.dtype = vals.dtype
        self.cdt = _lib.DTYPE_F64 if vals.dtype == torch.float64 else _lib.DTYPE_F32
        self.n = 3 * dev_mesh.num_node

    def matvec(self, x: torch.Tensor, y: torch.Tensor | None = None):
        if y is None:
            y = torch.empty_like(x)
        _lib.call("fem_bsr_spmv", self.mesh.ref, self.vals.data_ptr(), self.cdt, x.data_ptr(), y.data_ptr(),
                  stream_ptr())
        return y

    def block_jacobi(self):
        N = self.mesh.num_node
        dinv = torch.empty(9 * N, dtype=self.dtype, device=device())
        flag = torch.zeros(1, dtype=torch.int32, device=device())
        _lib.call("fem_bj_setup", self.mesh.ref, self.vals.data_ptr(), self.cdt, dinv.data_ptr(),
                  flag.data_ptr(), stream_ptr())
        return dinv, flag


class PcgWorkspace:
    """Vectors, scalar state and reduction scratch of one PCG."""

    def __init__(self, N: int, dtype):
        d = device()
        n = 3 * N
        self.x = torch.zeros(n, dtype=dtype, device=d)
        self.r = torch.empty(n, dtype=dtype, device=d)
        self.z = torch.empty(n, dtype=dtype, device=d)
        self.p = torch.empty(n, dtype=dtype, device=d)
        self.q = torch.empty(n, dtype=dtype, device=d)
        self.state = torch.zeros(8, dtype=torch.float64, device=d)
        self.wb = _lib.lib().fem_pcg_work_bytes(N)
        self.work = torch.zeros(self.wb, dtype=torch.uint8, device=d)


def pcg(op: BsrOperator, b: torch.Tensor, dinv: torch.Tensor, tol: float, maxiter: int,
        ws: PcgWorkspace | None = None, check_every: int = 16, x0: torch.Tensor | None = None):
    """Block-Jacobi PCG on ``op`` (dtype of op) to ||r||/||b|| < tol.

    Convergence is read from the device every ``check_every`` iterations (the
    recursion's residual); returns (x, iterations, rel_recursive_residual).
    """
    N = op.mesh.num_node
    ws = ws or PcgWorkspace(N, op.dtype)
    if x0 is None:
        ws.x.zero_()
    else:
        ws.x.copy_(x0)
    st = stream_ptr()
    m = op.mesh.ref
    _lib.call("fem_pcg_start", m, op.vals.data_ptr(), dinv.data_ptr(), op.cdt, b.data_ptr(), ws.x.data_ptr(),
              ws.r.data_ptr(), ws.z.data_ptr(), ws.p.data_ptr(), ws.state.data_ptr(), ws.work.data_ptr(),
              ws.wb, st)
    it = 0
    rel = float("inf")
    while it < maxiter:
        s = ws.state[[4, 5]].cpu()
        rr, bb = float(s[0]), float(s[1])
        rel = (rr / bb) ** 0.5 if bb > 0 else 0.0
        if rel < tol or bb == 0:
            break
        for _ in range(min(check_every, maxiter - it)):
            _lib.call("fem_pcg_iterate", m, op.vals.data_ptr(), dinv.data_ptr(), op.cdt, ws.x.data_ptr(),
                      ws.r.data_ptr(), ws.z.data_ptr(), ws.p.data_ptr(), ws.q.data_ptr(), ws.state.data_ptr(),
                      ws.work.data_ptr(), ws.wb, st)
            it += 1
    else:
        s = ws.state[[4, 5]].cpu()
        rel = (float(s[0]) / float(s[1])) ** 0.5 if float(s[1]) > 0 else 0.0
    return ws.x, it, rel


def true_residual(K64: BsrOperator, u: torch.Tensor, F: torch.Tensor):
    r = F - K64.matvec(u)
    nf = torch.linalg.vector_norm(F).item()
    return r, (torch.linalg.vector_norm(r).item() / nf if nf > 0 else 0.0)


def solve_pcg64(K64: BsrOperator, F: torch.Tensor, tol=1e-8, maxiter=200000):
    t0 = time.perf_counter()
    dinv, flag = K64.block_jacobi()
    x, it, _ = pcg(K64, F, dinv, tol, maxiter)
    _, rel = true_residual(K64, x, F)
    torch.cuda.synchronize()
    return x.clone(), SolveInfo("pcg", it, 1, rel, time.perf_counter() - t0, [])


def solve_mixed(K64: BsrOperator, K32: BsrOperator, F: torch.Tensor, tol=1e-8, inner_tol=1e-4,
                max_outer=40, maxiter=200000):
    """FP64 iterative refinement around FP32 block-Jacobi PCG."""
    t0 = time.perf_counter()
    dinv32, _ = K32.block_jacobi()
    N = K64.mesh.num_node
    ws = PcgWorkspace(N, torch.float32)
    u = torch.zeros_like(F)
    total = 0
    hist = []
    rel = 1.0
    outer = 0
    for outer in range(1, max_outer + 1):
        r, rel = true_residual(K64, u, F)
        hist.append(rel)
        if rel < tol:
            outer -= 1
            break
        scale = torch.linalg.vector_norm(r)
        r32 = (r / scale).float()
        d32, it, _ = pcg(K32, r32, dinv32, inner_tol, maxiter, ws=ws)
        total += it
        u += d32.double() * scale
    else:
        _, rel = true_residual(K64, u, F)
    torch.cuda.synchronize()
    return u, SolveInfo("mixed", total, outer, rel, time.perf_counter() - t0, hist)
