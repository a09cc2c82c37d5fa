"""GPU static solvers on the device-resident BSR(3x3) stiffness.

These replace the Pardiso factor/solve of SOLVE_static.py:31-38 (the tests
check them against a direct sparse solve on the CPU).  Every form is measured by the FP64 relative
residual ||F - K U|| / ||F|| on the penalised K, as the reference prints it
(SOLVE_static.py:36-38):

* ``pcg``   -- FP64 block-Jacobi (3x3 node blocks) preconditioned CG;
* ``mixed`` -- block-Jacobi PCG whose SpMV streams an FP32 copy of K (half
  the matrix bytes) while vectors and reductions stay FP64; the residual is
  replaced by the FP64 residual b - K64 x each time the recursion has
  dropped 100x and before stopping (iterative refinement folded into one
  Krylov run, so the Krylov space is not thrown away at each refinement);
* ``learned`` -- see learned.py (flexible PCG with the mesh-convolution
  preconditioner and block-Jacobi fallback).
"""
from __future__ import annotations

import dataclasses
import math
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


def bsr_spmm(dev_mesh, vals, X, Y, nb):
    """Y = K X for nb column-interleaved FP64 vectors: the structured-grid
    16-column kernel (fem_bsr_spmm16_grid) when the mesh is a node grid of
    that numbering and nb == 16, else fem_bsr_spmm."""
    grid = getattr(dev_mesh, "grid_nodes", None)
    if nb == 16 and grid is not None and X.data_ptr() % 16 == 0:
        _lib.call("fem_bsr_spmm16_grid", dev_mesh.ref, grid[0], grid[1], grid[2], vals.data_ptr(), X.data_ptr(),
                  Y.data_ptr(), stream_ptr())
    else:
        _lib.call("fem_bsr_spmm", dev_mesh.ref, vals.data_ptr(), X.data_ptr(), Y.data_ptr(), nb, stream_ptr())
    return Y


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
        ws: PcgWorkspace | None = None, check_every: int = 16, x0: torch.Tensor | None = None,
        exact: BsrOperator | None = None, drop: float = 1e-2, trace: list | None = None,
        patience: int = 4, gain: float = 0.1, status: dict | None = None):
    """Block-Jacobi PCG to ||r||/||b|| < tol.

    ``op`` supplies the matrix used inside the iteration.  When ``exact`` (an
    FP64 operator) is given, the vectors are FP64 and ``op``'s FP32 values
    drive the SpMV (FEM_DTYPE_F32F64); at a convergence check, whenever the
    recursive residual norm has dropped by ``drop`` since the last
    replacement or claims convergence, it is replaced by the exact
    ``b - exact @ x`` (reliable updates, van der Vorst & Ye 2000), which keeps
    the gap between recursion and true residual small.  If the exact residual
    still exceeds the recursive one by more than 4x the search direction
    restarts (p = z).  The stopping test is thus on the FP64 residual.

    The FP32 matrix differs from the FP64 one by ~6e-8 per entry; when the
    conditioning of K makes that perturbation comparable to its smallest
    eigenvalues the FP64 residual stops falling however long the recursion
    runs, or the recursion diverges.  In mixed mode a replacement also
    happens when the recursive residual grows by 1/``drop`` or turns
    non-finite, and the iteration stops once ``patience`` consecutive
    replacements fail to reduce the best FP64 residual norm by the factor
    ``gain`` (a sound FP32 matrix gains ~``drop`` per replacement; 0.1 leaves
    a 10x margin), or it turns non-finite; ``status`` (a dict, if given) then
    receives "stalled": True and "best_x", the iterate with the smallest
    FP64 residual.
    Returns (x, iterations, relative residual).
    """
    N = op.mesh.num_node
    mixed = exact is not None
    code = _lib.DTYPE_F32F64 if mixed else op.cdt
    vdt = torch.float64 if mixed else op.dtype
    ws = ws or PcgWorkspace(N, vdt)
    if x0 is None:
        ws.x.zero_()
    else:
        ws.x.copy_(x0)
    st = stream_ptr()
    m = op.mesh.ref
    _lib.call("fem_pcg_start", m, op.vals.data_ptr(), dinv.data_ptr(), code, b.data_ptr(), ws.x.data_ptr(),
              ws.r.data_ptr(), ws.z.data_ptr(), ws.p.data_ptr(), ws.state.data_ptr(), ws.work.data_ptr(),
              ws.wb, st)

    def read():
        s = ws.state[4:6].cpu()
        return float(s[0]), float(s[1])

    def replace(rr_rec):
        exact.matvec(ws.x, ws.q)
        _lib.call("fem_pcg_replace_residual", m, dinv.data_ptr(), code, b.data_ptr(), ws.q.data_ptr(),
                  ws.r.data_ptr(), ws.z.data_ptr(), None, ws.state.data_ptr(), ws.work.data_ptr(), ws.wb, st)
        rr, bb = read()
        restart = rr > 4.0 * rr_rec
        if restart:
            ws.p.copy_(ws.z)
        if trace is not None:
            trace.append(("replace", it, rr_rec, rr, restart))
        return rr, bb

    it = 0
    rr, bb = read()
    rr_anchor = rr
    best_rr, best_x, stalls = rr, None, 0
    if status is not None:
        status["stalled"] = False
    while True:
        rel = (rr / bb) ** 0.5 if bb > 0 else 0.0
        # replace on claimed convergence, on a drop by ``drop`` in norm and on a
        # rise by 1/drop (an FP32 recursion that diverges never drops)
        if mixed and it > 0 and (rel < tol or rr < drop * drop * rr_anchor or not math.isfinite(rr)
                                 or rr * drop * drop > rr_anchor):
            rr, bb = replace(rr)
            rr_anchor = rr
            rel = (rr / bb) ** 0.5 if bb > 0 else 0.0
            if math.isfinite(rr) and rr < gain * gain * best_rr:
                best_rr, stalls = rr, 0
                if status is not None:
                    best_x = ws.x.clone()
            else:
                stalls += 1
            if not math.isfinite(rr) or (stalls >= patience and rel >= tol):
                if status is not None:
                    status["stalled"] = True
                    status["best_x"] = best_x
                    status["best_rel"] = (best_rr / bb) ** 0.5 if bb > 0 else 0.0
                break
        if rel < tol or bb == 0 or it >= maxiter:
            break
        for _ in range(min(check_every, maxiter - it)):
            _lib.call("fem_pcg_iterate", m, op.vals.data_ptr(), dinv.data_ptr(), code, ws.x.data_ptr(),
                      ws.r.data_ptr(), ws.z.data_ptr(), ws.p.data_ptr(), ws.q.data_ptr(), ws.state.data_ptr(),
                      ws.work.data_ptr(), ws.wb, st)
            it += 1
        rr, bb = read()
        if trace is not None:
            trace.append(("check", it, rr, bb))
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


def solve_mixed(K64: BsrOperator, K32: BsrOperator, F: torch.Tensor, tol=1e-8, drop=1e-2,
                maxiter=200000, patience=2):
    """PCG whose SpMV streams the FP32 copy of K while vectors, dot products
    and the block-Jacobi application stay FP64; the residual is replaced by
    the FP64 one on the FP64 K (reliable updates), so the stopping test is the
    FP64 relative residual.  If the FP64 residual stalls (the FP32 rounding
    of K is too large for K's conditioning) the solve finishes with FP64 PCG
    from the best iterate; ``history`` records both phases."""
    t0 = time.perf_counter()
    dinv32, _ = K32.block_jacobi()
    st = {}
    x, it, _ = pcg(K32, F, dinv32, tol, maxiter, exact=K64, drop=drop, patience=patience, status=st)
    x = x.clone()
    hist = [{"mixed_iterations": it, "stalled": st["stalled"]}]
    outer = 1
    if st["stalled"]:
        # the FP32 matrix cannot reach tol on this K: finish with FP64 PCG from
        # the best iterate (or from zero if none beat it)
        x0 = st["best_x"] if (st["best_x"] is not None and st["best_rel"] < 1.0) else torch.zeros_like(F)
        hist[0]["best_rel"] = st["best_rel"]
        dinv64, _ = K64.block_jacobi()
        x, it2, _ = pcg(K64, F, dinv64, tol, maxiter, x0=x0)
        x = x.clone()
        it += it2
        outer = 2
        hist.append({"fp64_iterations": it2})
    _, rel = true_residual(K64, x, F)
    torch.cuda.synchronize()
    return x, SolveInfo("mixed", it, outer, rel, time.perf_counter() - t0, hist)
