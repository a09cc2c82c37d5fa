"""Drop-in block Lanczos (SOLVE_eigen.py:9-89) on the GPU.

``Eigen_solver(Model, Qtype=0).Solve_cpu(num_eigen=100, num_block=10)``
returns (frequencies in Hz, mode shapes (3N, num_eigen)) like the reference;
``Block_lancozs(K, M, K_inv, num_eigen).Solve_cpu()`` (the Readme's API)
returns (eigenvalues, vectors) for scipy K and M.

The recurrence is the reference's (SOLVE_eigen.py:45-89): shift-invert
block Lanczos in the M inner product with an upper-Cholesky M-orthonormal
start block, two full reorthogonalisation passes per step, Rayleigh-Ritz on
the block tridiagonal T, lambda = 1/theta, ascending order, zero-padded T
when num_block does not divide 2*num_eigen.  Differences, each stated:
the inner solves K^-1 (M u) are iterative GPU solves to ``inner_tol``
instead of a Pardiso factorisation (FP64 flexible CG with the FP32 multigrid
V-cycle on structured grids, multi-RHS block-Jacobi PCG otherwise); the start block is drawn on
the device from a generator seeded with ``seed``, or passed in as ``u0`` (the reference's is
unseeded); the dense
b x b / m x m work (Cholesky, inverse, eigh of T) and the tall-by-small
products Q C run as FP64 library calls on the device; the inner products
X^T Y over the 3N rows (the M-Gram blocks and the reorthogonalisation
coefficients) are fem_gemm_tn_f64, a deterministic row-split reduction.
Sparse products (K and M) and the PCG are the fem355 HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .runtime import device, stream_ptr, to_host


def _pow2(n):
    p = 1
    while p < n:
        p *= 2
    return p


class PatternMesh:
    """A pattern-only fem_mesh descriptor (enough for SpMV/SpMM/PCG kernels)."""

    def __init__(self, indptr, indices):
        d = device()
        self.sm_node = torch.as_tensor(np.asarray(indptr, dtype=np.int32)).to(d)
        self.re_node = torch.as_tensor(np.asarray(indices, dtype=np.int32)).to(d)
        self.co_node = (self.sm_node[1:] - self.sm_node[:-1]).contiguous()
        self.num_node = int(len(indptr) - 1)
        self.nnzb = int(len(indices))
        m = _lib.FemMesh()
        m.num_node, m.num_elem, m.nnzb = self.num_node, 0, self.nnzb
        m.sm_node = self.sm_node.data_ptr()
        m.re_node = self.re_node.data_ptr() if self.nnzb else None
        m.co_node = self.co_node.data_ptr()
        self.c = m

    @property
    def ref(self):
        import ctypes
        return ctypes.byref(self.c)


MAX_BLOCK = 64      # columns per multi-RHS kernel call


def _in_pieces(ops, B, tol, maxiter):
    """ops.solve on column pieces of at most MAX_BLOCK; one iteration record
    (the pieces' sum) per call."""
    parts, total = [], 0
    for j in range(0, B.shape[1], MAX_BLOCK):
        parts.append(ops.solve(B[:, j:j + MAX_BLOCK].contiguous(), tol, maxiter))
        total += ops.iterations.pop()
    ops.iterations.append(total)
    return torch.cat(parts, dim=1)


class BlockOps:
    """K (BSR 3x3 f64 on a mesh pattern) and an M-apply, with block PCG."""

    def __init__(self, mesh, k_vals, m_apply):
        self.mesh = mesh
        self.k_vals = k_vals
        self.m_apply = m_apply
        self.n = 3 * mesh.num_node
        dinv = torch.empty(9 * mesh.num_node, dtype=torch.float64, device=device())
        flag = torch.zeros(1, dtype=torch.int32, device=device())
        _lib.call("fem_bj_setup", mesh.ref, k_vals.data_ptr(), _lib.DTYPE_F64, dinv.data_ptr(),
                  flag.data_ptr(), stream_ptr())
        self.dinv = dinv
        self.iterations = []
        self.device = device()

    def solve(self, B, tol, maxiter=100000, check_every=16):
        """X = K^-1 B for an (n, b) block by multi-RHS block-Jacobi PCG
        (blocks wider than the kernels' 64 columns in pieces)."""
        if B.shape[1] > MAX_BLOCK:
            return _in_pieces(self, B, tol, maxiter)
        n, b = B.shape
        nb = _pow2(b)
        d = device()
        Bp = torch.zeros(n, nb, dtype=torch.float64, device=d)
        Bp[:, :b] = B
        X = torch.zeros_like(Bp)
        R, Z, P, Q = (torch.empty_like(Bp) for _ in range(4))
        state = torch.zeros(8, nb, dtype=torch.float64, device=d)
        wb = _lib.lib().fem_mpcg_work_bytes(nb)
        work = torch.zeros(wb, dtype=torch.uint8, device=d)
        st = stream_ptr()
        m = self.mesh.ref
        _lib.call("fem_mpcg_start", m, self.k_vals.data_ptr(), self.dinv.data_ptr(), nb, Bp.data_ptr(),
                  X.data_ptr(), R.data_ptr(), Z.data_ptr(), P.data_ptr(), state.data_ptr(), work.data_ptr(),
                  wb, st)
        it = 0
        while it < maxiter:
            s = state[4:6, :b].cpu()
            bb = s[1].clamp_min(1e-300)
            rel = torch.sqrt(s[0] / bb).max().item()
            if rel < tol:
                break
            for _ in range(min(check_every, maxiter - it)):
                _lib.call("fem_mpcg_iterate", m, self.k_vals.data_ptr(), self.dinv.data_ptr(), nb, X.data_ptr(),
                          R.data_ptr(), Z.data_ptr(), P.data_ptr(), Q.data_ptr(), state.data_ptr(),
                          work.data_ptr(), wb, st)
                it += 1
        self.iterations.append(it)
        return X[:, :b].contiguous()


class MultigridBlockOps:
    """K (the model's FP64 BSR) and an M-apply, with the inner solves
    K^-1 B done by FP64 block flexible CG (all columns at once, per-column
    scalars) preconditioned by the FP32 multigrid V-cycle on all columns
    (structured node grids; multigrid.py).  The
    operator is the model's K or ``k_vals`` (BSR values on the same pattern,
    e.g. a prestressed K); the V-cycle is built from the model's K either
    way.  One V-cycle set-up serves every solve of the eigen-solve.  A column stops at
    ``tol`` or at the attainable FP64 accuracy (its true residual no longer
    improving); one that fails otherwise within ``maxiter`` is finished by
    FP64 block-Jacobi PCG."""

    def __init__(self, model, m_apply, min_nodes=1000, k_vals=None, patience=2):
        from . import multigrid, solvers
        from .learned import free_dof_mask
        self.model = model
        self.patience = patience      # true-residual checks without halving before a column is at its floor
        self.K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64") if k_vals is None else k_vals)
        self.pre = multigrid.MultigridPreconditioner(model, multigrid.hierarchy_for(model, min_nodes))
        self.dinv64, _ = self.K64.block_jacobi()
        self.free = free_dof_mask(model)
        self.m_apply = m_apply
        self.n = 3 * model.Num_Node
        self.device = device()
        self.iterations = []
        self.inner_rel_max = 0.0      # worst relative residual ||B - K X|| / ||B|| an inner solve returned
        self.fallback_columns = 0

    def solve(self, B, tol, maxiter=1000):
        from . import multigrid, solvers
        if B.shape[1] > MAX_BLOCK:
            return _in_pieces(self, B, tol, maxiter)
        X, info = multigrid.solve_block_fcg_mg(self.model, self.K64, B.contiguous(), self.pre, tol=tol,
                                               maxiter=maxiter, free=self.free, dinv64=self.dinv64,
                                               patience=self.patience)
        total = info["iterations"] * B.shape[1]
        rel = np.array(info["rel_residual"], dtype=np.float64)
        for j in np.nonzero(~(info["converged"] | info["stagnated"]))[0]:
            bj = B[:, j].contiguous()
            x, it2, _ = solvers.pcg(self.K64, bj, self.dinv64, tol, 200000, x0=X[:, j].contiguous().clone())
            X[:, j] = x
            total += it2
            self.fallback_columns += 1
            bnj = float(torch.linalg.vector_norm(bj))
            rel[j] = float(torch.linalg.vector_norm(bj - self.K64.matvec(x))) / bnj if bnj > 0 else 0.0
        if rel.size:
            self.inner_rel_max = max(self.inner_rel_max, float(np.max(rel)))
        self.iterations.append(total)
        return X


_TN_WORK = {}


def gram(X, Y):
    """X^T Y for tall FP64 device matrices X (n, p) and Y (n, q) whose rows are
    contiguous (fem_gemm_tn_f64; X's columns in pieces of at most 512)."""
    if X.dtype != torch.float64 or Y.dtype != torch.float64 or X.shape[0] != Y.shape[0]:
        raise ValueError("gram needs FP64 matrices with the same number of rows")
    if X.stride(1) != 1:
        X = X.contiguous()
    if Y.stride(1) != 1:
        Y = Y.contiguous()
    n, p = X.shape
    q = Y.shape[1]
    if q > 64:
        return torch.cat([gram(X, Y[:, j:j + 64]) for j in range(0, q, 64)], dim=1)
    if p > 512 or p * q > 8192:
        w = max(1, min(512, 8192 // q))
        return torch.cat([gram(X[:, i:i + w], Y) for i in range(0, p, w)], dim=0)
    if n == 0:
        return torch.zeros(p, q, dtype=torch.float64, device=X.device)
    lib = _lib.lib()
    wb = int(lib.fem_gemm_tn_work_bytes(p, q))
    work = _TN_WORK.get(X.device)
    if work is None or work.numel() < wb:
        work = torch.empty(max(wb, 1 << 20), dtype=torch.uint8, device=X.device)
        _TN_WORK[X.device] = work
    C = torch.empty(p, q, dtype=torch.float64, device=X.device)
    _lib.call("fem_gemm_tn_f64", X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0), n, p, q, C.data_ptr(),
              work.data_ptr(), work.numel(), stream_ptr())
    return C


def gemm_nn(A, C, Y=None, alpha=1.0, beta=0.0):
    """Y = alpha A C + beta Y (fem_gemm_nn_f64): A (n, k) with unit column
    stride (a column slice of Q is fine), C (k, q) small; a new Y when None."""
    n, k = A.shape
    q = C.shape[1]
    C = C.contiguous()
    if Y is None:
        Y = torch.empty(n, q, dtype=torch.float64, device=A.device)
        beta = 0.0
    if A.stride(1) != 1 or Y.stride(1) != 1 or A.dtype != torch.float64 or Y.dtype != torch.float64:
        raise ValueError("gemm_nn needs FP64 operands with unit column stride")
    _lib.call("fem_gemm_nn_f64", A.data_ptr(), A.stride(0), C.data_ptr(), C.stride(0), Y.data_ptr(), Y.stride(0),
              n, k, q, float(alpha), float(beta), stream_ptr())
    return Y


def chol_upper_inv(A):
    """(R, R^-1) with A = R^T R, R upper (scipy.linalg.cholesky(A), SOLVE_eigen.py:52, 75), on the
    device by fem_chol_upper_inv_f64; a matrix that is not positive definite gives NaN."""
    A = A.contiguous()
    R = torch.empty_like(A)
    Rinv = torch.empty_like(A)
    _lib.call("fem_chol_upper_inv_f64", A.data_ptr(), A.shape[0], R.data_ptr(), Rinv.data_ptr(), stream_ptr())
    return R, Rinv


def lanczos(ops, num_eigen=100, num_block=10, seed=0, inner_tol=1e-12, u0=None):
    """SOLVE_eigen.py:45-89.  ``ops`` provides ``n``, ``device``,
    ``m_apply(X)`` and ``solve(B, tol)``.  Returns (eigenvalues ascending
    [:num_eigen], vectors (n, num_eigen), T) as torch tensors.

    The start block is ``u0`` (n, b; an array or a device tensor) when
    given, else uniform [0, 1) draws from a device generator seeded with
    ``seed`` (the reference draws it unseeded with np.random.uniform,
    SOLVE_eigen.py:51); on the device because a host draw of a 7M x 16 block
    alone takes seconds."""
    d = ops.device
    m = num_eigen * 2
    n = ops.n
    b = int(num_block)
    Q = torch.zeros(n, m, dtype=torch.float64, device=d)
    T = torch.zeros(m, m, dtype=torch.float64, device=d)
    if u0 is None:
        gen = torch.Generator(device=d)
        gen.manual_seed(int(seed))
        u = torch.rand(n, b, dtype=torch.float64, device=d, generator=gen)
    elif isinstance(u0, torch.Tensor):
        u = u0.to(device=d, dtype=torch.float64).contiguous()
    else:
        u = torch.as_tensor(np.asarray(u0, dtype=np.float64)).to(d)
    _, vinv = chol_upper_inv(gram(u, ops.m_apply(u)))
    u = gemm_nn(u, vinv)
    Q[:, :b] = u
    steps = int(m / b)
    for i in range(steps):
        v = ops.solve(ops.m_apply(u), inner_tol)
        Tii = gram(u, ops.m_apply(v))
        T[i * b:(i + 1) * b, i * b:(i + 1) * b] = Tii
        gemm_nn(u, Tii, v, -1.0, 1.0)                         # v -= u T_ii
        if i >= steps - 1:
            break
        k = i * b + b
        Qk = Q[:, :k]
        gemm_nn(Qk, gram(Qk, ops.m_apply(v)), v, -1.0, 1.0)  # two passes of v -= Q (Q^T M v)
        gemm_nn(Qk, gram(Qk, ops.m_apply(v)), v, -1.0, 1.0)
        t, tinv = chol_upper_inv(gram(v, ops.m_apply(v)))
        T[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = t
        T[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = t.T
        u = gemm_nn(v, tinv)
        Q[:, (i + 1) * b:(i + 2) * b] = u
    # T's entries are ~1/lambda (1e-11 .. 1e-13 for steel in mm/t/s units);
    # the device eigensolver loses relative accuracy at that scale, so it runs
    # on T / max|T| and theta is scaled back.
    if not bool(torch.isfinite(T).all()):
        raise np.linalg.LinAlgError("block Lanczos: a block Gram matrix u^T M u is not positive definite")
    scale = T.abs().max()
    scale = torch.where(scale > 0, scale, torch.ones_like(scale))
    theta, V = torch.linalg.eigh(T / scale)
    theta = theta * scale
    lam = 1.0 / theta
    ind = torch.argsort(lam)[:num_eigen]
    Vs = V[:, ind]                                             # the returned Ritz pairs, ascending
    k = Vs.shape[1]
    X = torch.empty(n, k, dtype=torch.float64, device=d)
    for c0 in range(0, k, 64):                                 # Ritz vectors Q V
        gemm_nn(Q, Vs[:, c0:c0 + 64], X[:, c0:c0 + 64])
    return lam[ind], X, T


class Eigen_solver:
    """SOLVE_eigen.py:9-43: modal analysis of a Model (linear K and M)."""

    def __init__(self, Model, Qtype=0):
        self.Qtype = Qtype
        self.Model = Model
        self.ops = None

    def Solve_cpu(self, num_eigen=100, num_block=10, seed=0, inner_tol=1e-12, verbose=False, inner="auto",
                  prestress=None, omega=None, u0=None):
        """``inner``: "mg" (multigrid flexible CG, structured grids), "pcg"
        (multi-RHS block-Jacobi PCG) or "auto" (mg when the Physics has
        node_dims).

        The stiffness operator is K_L; or, after a large-displacement static
        solve (``Static_solver(Model, 1)``), its final tangent KT, as the
        reference factors ``Model.KT`` (SOLVE_eigen.py:32); or, when
        ``prestress`` is a displacement U (3N) or "static" (U from the linear
        static solve of the model's own load), the prestressed K_L +
        K_sigma(U) minus the spin softening of ``omega`` (default: the
        model's Omegas) -- stress stiffening and spin softening, which the
        reference lists as missing (Readme.md:189)."""
        Mo = self.Model
        if prestress is not None:
            if isinstance(prestress, str):
                if prestress != "static":
                    raise ValueError(f"prestress must be a displacement or 'static', got {prestress!r}")
                from .solve_static import Static_solver
                structured = getattr(Mo.Physics, "node_dims", None) is not None
                prestress = Static_solver(Mo, 0).Solve_cpu(method="mg" if structured else "pcg", tol=1e-10)
            w = np.asarray(Mo.OM if omega is None else omega, dtype=np.float64)
            kv = Mo.assemble_KS(prestress, omega=w if np.any(w) else None)
            self.operator = "KS"
        elif getattr(Mo, "KT_device", None) is not None:
            kv = Mo.KT_device
            self.operator = "KT"
        else:
            kv = Mo.K_bsr("f64")
            self.operator = "KL"
        mb = Mo.M_blocks()
        dev_mesh = Mo.dev
        if inner == "auto":
            inner = "mg" if getattr(Mo.Physics, "node_dims", None) is not None else "pcg"
        if inner not in ("mg", "pcg"):
            raise ValueError(f"unknown inner solver {inner!r}")

        grid = getattr(dev_mesh, "grid_nodes", None)

        def m_apply(X):
            X = X.contiguous()
            Y = torch.empty(X.shape, dtype=torch.float64, device=X.device)
            if X.shape[1] == 16 and grid is not None and X.data_ptr() % 16 == 0:
                _lib.call("fem_mass_spmm16_grid", dev_mesh.ref, grid[0], grid[1], grid[2], mb.data_ptr(),
                          X.data_ptr(), Y.data_ptr(), stream_ptr())
            else:
                _lib.call("fem_mass_spmm", dev_mesh.ref, mb.data_ptr(), X.data_ptr(), Y.data_ptr(), X.shape[1],
                          stream_ptr())
            return Y

        self.ops = MultigridBlockOps(Mo, m_apply, k_vals=kv) if inner == "mg" else BlockOps(dev_mesh, kv, m_apply)
        lam, X, T = lanczos(self.ops, num_eigen, num_block, seed=seed, inner_tol=inner_tol, u0=u0)
        self.eigenvalues = lam
        freq = torch.sqrt(lam) / 2 / np.pi
        if verbose:
            for i in range(num_eigen):
                print('\t\t\tmode {:3d} frequency {:.3e}'.format(i + 1, float(freq[i])))
        return freq.cpu().numpy(), to_host(X)


class Block_lancozs:
    """The Readme's API (Readme.md:100-108): Block_lancozs(K, M, K_inv,
    num_eigen).Solve_cpu() -> (eigenvalues, vectors) for scipy K and M with
    3x3 node blocks.  K_inv is accepted for signature compatibility; the
    inner solves are GPU block PCG."""

    def __init__(self, K, M, K_inv=None, num_eigen=100, num_block=10):
        self.K, self.M, self.K_inv = K, M, K_inv
        self.num_eigen, self.num_block = num_eigen, num_block

    def Solve_cpu(self, seed=0, inner_tol=1e-12, u0=None):
        Kb = self.K.tobsr(blocksize=(3, 3))
        Kb.sort_indices()
        Mb = self.M.tobsr(blocksize=(3, 3))
        Mb.sort_indices()
        d = device()
        kmesh = PatternMesh(Kb.indptr, Kb.indices)
        mmesh = PatternMesh(Mb.indptr, Mb.indices)
        kv = torch.as_tensor(np.ascontiguousarray(Kb.data, dtype=np.float64).reshape(-1)).to(d)
        mv = torch.as_tensor(np.ascontiguousarray(Mb.data, dtype=np.float64).reshape(-1)).to(d)
        self._keep = (kmesh, mmesh, kv, mv)

        def m_apply(X):
            X = X.contiguous()
            Y = torch.empty(X.shape, dtype=torch.float64, device=X.device)
            _lib.call("fem_bsr_spmm", mmesh.ref, mv.data_ptr(), X.data_ptr(), Y.data_ptr(), X.shape[1],
                      stream_ptr())
            return Y

        ops = BlockOps(kmesh, kv, m_apply)
        lam, X, _ = lanczos(ops, self.num_eigen, self.num_block, seed=seed, inner_tol=inner_tol, u0=u0)
        return lam.cpu().numpy(), to_host(X)
