"""Geometric multigrid preconditioner for the static solve on structured grids.

No counterpart in the reference, whose static solve is a Pardiso
factorisation (SOLVE_static.py:31-38); this is one of the GPU forms that
replace it, measured like the others by the FP64 residual ||F - K U|| / ||F|| on
the penalised K.

Hierarchy (symbolic, once per mesh, cached on the ``Physics``): the node grid
``node_dims`` (x fastest) is coarsened by keeping every second index along
each axis plus the last one, until at most ``min_nodes`` nodes remain or an
axis has fewer than 3 nodes.  Each coarse level is a hex8 mesh of its own
(the selected nodes at their fine coordinates, elements between neighbouring
selected nodes) with its own GPU pattern, row groups and task lists, so its
stiffness is assembled by the same kernels as the fine one
(re-discretisation).  A coarse element's modulus is the mean of the fine
moduli it covers; a coarse DOF is penalised when the coinciding fine DOF is.

Numeric setup (per solve): FP32 K with the penalty on every level (FP64 on
the coarsest), the 3x3 block-Jacobi inverses, a power-iteration estimate of
lambda_max(Dinv K) per level giving the damping omega = 1 / theta of the
degree-1 Chebyshev (damped block-Jacobi) smoother for the interval
[lambda_max / 30, 1.1 lambda_max], and the dense inverse of the coarsest K.

V-cycle (FP32, all on the GPU): per level x = omega Dinv b, r = b - K x,
b_c = P^T r, recurse, x += P x_c, x <- x + omega Dinv (b - K x); the
coarsest level applies its dense inverse.  The outer iteration is FP64
flexible CG (FCG(1)) on the free DOFs: the penalised rows (diagonal 1e20)
carry almost no energy, so a Krylov iteration that includes them reduces
their residual only by rounding luck and, once its residual is replaced by
the true one, stalls on them; they are kept out of the iteration and set at
the end by one block-Jacobi step on their own rows (exact to rounding, since
their diagonal dominates), and the full FP64 residual is checked before
returning.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _lib
from .runtime import device, stream_ptr

CHEB_RATIO = 30.0      # smoothing interval [lambda_max / ratio, 1.1 lambda_max]
CHEB_SAFETY = 1.1
EIG_ITERS = 15


def axis_select(n):
    """Indices kept along an axis of n nodes: 0, 2, 4, ... and n - 1."""
    sel = np.arange(0, n, 2)
    if sel[-1] != n - 1:
        sel = np.append(sel, n - 1)
    return sel


def axis_tables(n, sel):
    """left (n,), wright (n,), span (2 nc,) of one axis (see fem_grid_transfer)."""
    nc = len(sel)
    f = np.arange(n)
    left = np.searchsorted(sel, f, side="right") - 1
    left = np.minimum(left, nc - 1)
    nxt = np.minimum(left + 1, nc - 1)
    gap = sel[nxt] - sel[left]
    wright = np.where(gap > 0, (f - sel[left]) / np.maximum(gap, 1), 0.0)
    first = np.where(np.arange(nc) > 0, np.concatenate([[0], sel[:-1] + 1]), 0)
    last = np.where(np.arange(nc) < nc - 1, np.concatenate([sel[1:] - 1, [n - 1]]), n - 1)
    span = np.stack([first, last], axis=1).reshape(-1)
    return left.astype(np.int32), wright.astype(np.float32), span.astype(np.int32)


def interpolation_matrix(dims_f, sels):
    """Dense-free scipy form of P (nodes), for tests: kron(Pz, Py, Px)."""
    import scipy.sparse as sp
    mats = []
    for n, sel in zip(dims_f, sels):
        left, wright, _ = axis_tables(n, sel)
        nc = len(sel)
        rows = np.concatenate([np.arange(n), np.arange(n)])
        cols = np.concatenate([left, np.minimum(left + 1, nc - 1)])
        vals = np.concatenate([1.0 - wright, wright])
        keep = vals != 0
        mats.append(sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(n, nc)))
    return sp.kron(mats[2], sp.kron(mats[1], mats[0])).tocsr()


def coarse_modulus(EX, dims_f, sels):
    """Mean of the fine element moduli over each coarse element (scalar stays scalar)."""
    ex = np.asarray(EX, dtype=np.float64)
    if ex.ndim == 0:
        return float(ex)
    ne = [d - 1 for d in dims_f]
    e3 = ex.reshape(ne[2], ne[1], ne[0])
    for ax, sel in ((2, sels[0]), (1, sels[1]), (0, sels[2])):
        starts = sel[:-1]
        cnt = np.diff(sel).astype(np.float64)
        e3 = np.add.reduceat(e3, starts, axis=ax)
        shape = [1, 1, 1]
        shape[ax] = -1
        e3 = e3 / cnt.reshape(shape)
    return np.ascontiguousarray(e3.reshape(-1))


class Transfer:
    """Device tables + the fem_grid_transfer descriptor between two grids."""

    def __init__(self, dims_f, sels):
        tabs = [axis_tables(n, s) for n, s in zip(dims_f, sels)]
        d = device()
        self.left = torch.as_tensor(np.concatenate([t[0] for t in tabs])).to(d)
        self.wright = torch.as_tensor(np.concatenate([t[1] for t in tabs])).to(d)
        self.span = torch.as_tensor(np.concatenate([t[2] for t in tabs])).to(d)
        c = _lib.GridTransfer()
        for a in range(3):
            c.nf[a] = int(dims_f[a])
            c.nc[a] = int(len(sels[a]))
        c.left, c.wright, c.span = self.left.data_ptr(), self.wright.data_ptr(), self.span.data_ptr()
        self.c = c
        self.dims_f = tuple(int(x) for x in dims_f)
        self.dims_c = tuple(len(s) for s in sels)

    @property
    def ref(self):
        import ctypes
        return ctypes.byref(self.c)


def _fixed_dofs(model):
    """(N, 3) bool: DOFs the penalty fixes (bound flag != -1)."""
    N = model.Num_Node
    fixed = np.zeros((N, 3), dtype=bool)
    bn = model.bound_node.cpu().numpy().astype(np.int64)
    fl = model.bound_flag.cpu().numpy().reshape(-1, 3)
    for j in range(3):
        fixed[bn[fl[:, j] != -1], j] = True
    return fixed


def _bounds_from_fixed(fixed):
    """Bounds list (reference format) penalising exactly the DOFs of ``fixed``."""
    code = fixed[:, 0] * 1 + fixed[:, 1] * 2 + fixed[:, 2] * 4
    out = []
    for c in range(1, 8):
        ids = np.nonzero(code == c)[0]
        if len(ids):
            out.append([ids.astype(np.int64), [0.0 if (c >> j) & 1 else -1.0 for j in range(3)]])
    return out


class Hierarchy:
    """The coarse grids of a structured mesh (symbolic; no matrix values)."""

    def __init__(self, phys, fixed, min_nodes=1000, max_levels=12):
        from .physics import Physics
        from .mesh import grid_connectivity
        if getattr(phys, "node_dims", None) is None:
            raise ValueError("multigrid needs a structured node grid (Physics.node_dims)")
        t0 = time.perf_counter()
        dims = tuple(int(x) for x in phys.node_dims)
        if int(np.prod(dims)) != phys.Num_Node:
            raise ValueError("node_dims do not match the node count")
        coords = np.ascontiguousarray(np.asarray(phys.Node[-3:], dtype=np.float64).T)
        self.dims = [dims]
        self.sels = []
        self.transfers = []
        self.phys = [phys]
        self.node_maps = []             # coarse node -> node of the finer level
        self.fixed = [fixed]
        while len(self.dims) < max_levels:
            d = self.dims[-1]
            if int(np.prod(d)) <= min_nodes or min(d) < 3:
                break
            sels = [axis_select(n) for n in d]
            dc = tuple(len(s) for s in sels)
            idx = (sels[0][None, None, :] + d[0] * (sels[1][None, :, None] + d[1] * sels[2][:, None, None])).ravel()
            coords = coords[idx]
            fixed_c = self.fixed[-1][idx]
            Nc = coords.shape[0]
            Node = np.concatenate([np.arange(1, Nc + 1, dtype=np.float64)[None, :], coords.T], axis=0)
            Elem = grid_connectivity(dc[0] - 1, dc[1] - 1, dc[2] - 1).T.copy()
            pc = Physics(Node, Elem, _bounds_from_fixed(fixed_c), [], name=f"{phys.name}_mg{len(self.dims)}",
                         node_dims=dc)
            self.sels.append(sels)
            self.transfers.append(Transfer(d, sels))
            self.node_maps.append(idx)
            self.dims.append(dc)
            self.phys.append(pc)
            self.fixed.append(fixed_c)
        self.seconds = time.perf_counter() - t0

    @property
    def num_levels(self):
        return len(self.dims)


def hierarchy_for(model, min_nodes=1000):
    """The cached hierarchy of the model's mesh (built on first use)."""
    phys = model.Physics
    key = ("mg_hierarchy", min_nodes)
    cache = phys.__dict__.setdefault("_mg_cache", {})
    if key not in cache:
        cache[key] = Hierarchy(phys, _fixed_dofs(model), min_nodes=min_nodes)
    return cache[key]


_SIDE_STREAMS = {}


def _side_stream(dev):
    """One second stream per device (default priority) for one-time setup
    work that can run beside the main stream's."""
    key = str(dev)
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return _SIDE_STREAMS[key]


class MultigridPreconditioner:
    """FP32 V-cycle z = MG(r) for FP64 r (numeric setup in the constructor)."""

    def __init__(self, model, hier: Hierarchy, seed=0):
        from .model import Model
        from .solvers import BsrOperator
        t0 = time.perf_counter()
        d = device()
        self.hier = hier
        L = hier.num_levels
        self.meshes, self.k32, self.dinv, self.omega = [], [], [], []
        EX = model.EX
        mats = {"EX": EX, "PO": model.PO, "DE": model.DE}
        models = []
        for l in range(L):
            if l == 0:
                m = model
            else:
                mats = {"EX": coarse_modulus(mats["EX"], hier.dims[l - 1], hier.sels[l - 1]), "PO": model.PO,
                        "DE": model.DE}
                m = Model(hier.phys[l].name, hier.phys[l], mats, Omegas=(0.0, 0.0, 0.0))
            models.append(m)
            self.meshes.append(m.dev)
        # The coarsest level's dense inverse (rocSOLVER's Cholesky: many small,
        # latency-bound kernels) runs on a second stream beside the finer
        # levels' assembly and power steps (bandwidth-bound); both are one-time
        # numeric setup.  Its status is read only after the join.
        main = torch.cuda.current_stream()
        self.k_coarse = models[-1].assemble_K("f64", "bsr")
        side = _side_stream(d)
        side.wait_stream(main)
        with torch.cuda.stream(side):
            coarse_inv, info = self._dense_inverse(models[-1], self.k_coarse)
        for l in range(L - 1):
            m = models[l]
            k32 = m.K_bsr("f32") if l == 0 else m.assemble_K("f32", "bsr")
            dinv, _ = BsrOperator(m.dev, k32).block_jacobi()
            self.k32.append(k32)
            self.dinv.append(dinv)
            self.omega.append(1.0 / self._theta(m.dev, k32, dinv, seed + l))
        main.wait_stream(side)
        coarse_inv.record_stream(main)
        if int(info) != 0:
            raise _lib.FemError(_lib.FEM_E_NUMERIC, "multigrid: the coarsest K is not positive definite")
        self.coarse_inv = coarse_inv
        n = [3 * int(np.prod(dm)) for dm in hier.dims]
        self.b = [torch.empty(k, dtype=torch.float32, device=d) for k in n]
        self.x = [torch.empty(k, dtype=torch.float32, device=d) for k in n]
        self.t = [torch.empty(k, dtype=torch.float32, device=d) for k in n]
        self.applications = 0
        # damped block-Jacobi sweeps before and after the coarse correction:
        # one for a single vector (the static solve: two take fewer
        # iterations, no less time); for the 16-column inner solves of block
        # Lanczos two on the finest level and three on the coarser ones
        # (sweeps_multi[min(level, 1)]), 21% less time than one everywhere on
        # the 500 x 60 x 75 blade (more sweeps, or a Chebyshev sequence of
        # damping factors, take no less)
        self.sweeps = 1
        self.sweeps_multi = (2, 3)
        self._multi = {}
        torch.cuda.synchronize()
        self.setup_seconds = time.perf_counter() - t0

    @staticmethod
    def _theta(dev, k32, dinv, seed):
        """Centre of the Chebyshev interval from lambda_max(Dinv K) by power iteration."""
        st = stream_ptr()
        n = 3 * dev.num_node
        g = torch.Generator(device=device()).manual_seed(seed)
        x = torch.randn(n, generator=g, dtype=torch.float32, device=device())
        x /= torch.linalg.vector_norm(x)
        zero = torch.zeros_like(x)
        r = torch.empty_like(x)
        y = torch.empty_like(x)
        lam = torch.zeros((), dtype=torch.float32, device=x.device)
        for _ in range(EIG_ITERS):
            _lib.call("fem_mg_smooth", dev.ref, k32.data_ptr(), _lib.LAYOUT_BSR, dinv.data_ptr(), 1.0, x.data_ptr(),
                      zero.data_ptr(), r.data_ptr(), 0, st)      # r = -K x
            _lib.call("fem_mg_smooth", dev.ref, None, _lib.LAYOUT_BSR, dinv.data_ptr(), -1.0, None, r.data_ptr(),
                      y.data_ptr(), 2, st)                       # y = Dinv K x
            lam = torch.linalg.vector_norm(y)
            x = y / lam
        lmax = float(lam)
        if not (math.isfinite(lmax) and lmax > 0):
            raise _lib.FemError(_lib.FEM_E_NUMERIC, f"multigrid: lambda_max(Dinv K) estimate {lmax}")
        hi, lo = CHEB_SAFETY * lmax, lmax / CHEB_RATIO
        return 0.5 * (hi + lo)

    @staticmethod
    def _dense_inverse(m, vals64):
        """FP32 copy of the inverse of the coarsest (penalised) K, by FP64
        Cholesky, and the factorisation's status (device; nonzero: not
        positive definite), queued on the current stream without waiting."""
        N = m.Num_Node
        d = device()
        co = m.dev["co_node"].long()
        rows = torch.repeat_interleave(torch.arange(N, device=d), co)
        cols = m.dev["re_node"].long()
        r3 = (3 * rows[:, None, None] + torch.arange(3, device=d)[None, :, None]).expand(-1, 3, 3)
        c3 = (3 * cols[:, None, None] + torch.arange(3, device=d)[None, None, :]).expand(-1, 3, 3)
        A = torch.zeros(3 * N, 3 * N, dtype=torch.float64, device=d)
        A[r3.reshape(-1), c3.reshape(-1)] = vals64.view(-1)
        Lc, info = torch.linalg.cholesky_ex(A)
        inv = torch.cholesky_inverse(Lc)
        return inv.float().contiguous(), info

    def vcycle(self, b0: torch.Tensor) -> torch.Tensor:
        """x = MG(b0) in FP32 (b0 FP32 on the finest level)."""
        st = stream_ptr()
        L = self.hier.num_levels
        self.b[0] = b0
        for l in range(L - 1):
            dev = self.meshes[l]
            k, dv, w = self.k32[l].data_ptr(), self.dinv[l].data_ptr(), self.omega[l]
            _lib.call("fem_mg_smooth", dev.ref, None, _lib.LAYOUT_BSR, dv, w, None, self.b[l].data_ptr(),
                      self.x[l].data_ptr(), 2, st)
            for _ in range(self.sweeps - 1):
                _lib.call("fem_mg_smooth", dev.ref, k, _lib.LAYOUT_BSR, dv, w, self.x[l].data_ptr(),
                          self.b[l].data_ptr(), self.t[l].data_ptr(), 1, st)
                self.x[l], self.t[l] = self.t[l], self.x[l]
            _lib.call("fem_mg_smooth", dev.ref, k, _lib.LAYOUT_BSR, dv, w, self.x[l].data_ptr(), self.b[l].data_ptr(),
                      self.t[l].data_ptr(), 0, st)
            _lib.call("fem_mg_restrict", self.hier.transfers[l].ref, self.t[l].data_ptr(), self.b[l + 1].data_ptr(), 1,
                      st)
        n = self.b[L - 1].numel()
        _lib.call("fem_dense_gemv_f32", self.coarse_inv.data_ptr(), n, self.b[L - 1].data_ptr(),
                  self.x[L - 1].data_ptr(), st)
        for l in range(L - 2, -1, -1):
            dev = self.meshes[l]
            k, dv, w = self.k32[l].data_ptr(), self.dinv[l].data_ptr(), self.omega[l]
            _lib.call("fem_mg_prolong_add", self.hier.transfers[l].ref, self.x[l + 1].data_ptr(), self.x[l].data_ptr(),
                      1, st)
            for _ in range(self.sweeps):
                _lib.call("fem_mg_smooth", dev.ref, k, _lib.LAYOUT_BSR, dv, w, self.x[l].data_ptr(),
                          self.b[l].data_ptr(), self.t[l].data_ptr(), 1, st)
                self.x[l], self.t[l] = self.t[l], self.x[l]
        return self.x[0]

    def __call__(self, r: torch.Tensor) -> torch.Tensor:
        self.applications += 1
        return self.vcycle(r.float().contiguous()).double()

    def _sweeps_at(self, level):
        sw = self.sweeps_multi
        return sw[min(level, len(sw) - 1)] if isinstance(sw, (list, tuple)) else sw

    def vcycle_multi(self, B0: torch.Tensor) -> torch.Tensor:
        """The V-cycle on nb columns at once: B0 (3N, nb) FP32, row-major
        (node-interleaved, columns fastest); returns X (3N, nb) FP32."""
        st = stream_ptr()
        L = self.hier.num_levels
        nb = int(B0.shape[1])
        key = ("multi", nb)
        if key not in self._multi:
            d = device()
            n = [3 * int(np.prod(dm)) for dm in self.hier.dims]
            self._multi[key] = tuple([torch.empty(k, nb, dtype=torch.float32, device=d) for k in n] for _ in range(3))
        b, x, t = self._multi[key]
        b[0] = B0.contiguous()
        for l in range(L - 1):
            dev = self.meshes[l]
            k, dv, w = self.k32[l].data_ptr(), self.dinv[l].data_ptr(), self.omega[l]
            _lib.call("fem_mg_smooth_multi", dev.ref, None, dv, w, None, b[l].data_ptr(), x[l].data_ptr(), nb, 2, st)
            for _ in range(self._sweeps_at(l) - 1):
                _smooth_multi(dev, k, dv, w, x[l], b[l], t[l], nb, 1, st)
                x[l], t[l] = t[l], x[l]
            _smooth_multi(dev, k, dv, w, x[l], b[l], t[l], nb, 0, st)
            _lib.call("fem_mg_restrict", self.hier.transfers[l].ref, t[l].data_ptr(), b[l + 1].data_ptr(), nb, st)
        _lib.call("fem_dense_gemm_nb_f32", self.coarse_inv.data_ptr(), b[L - 1].shape[0], b[L - 1].data_ptr(),
                  x[L - 1].data_ptr(), nb, st)
        for l in range(L - 2, -1, -1):
            dev = self.meshes[l]
            k, dv, w = self.k32[l].data_ptr(), self.dinv[l].data_ptr(), self.omega[l]
            _lib.call("fem_mg_prolong_add", self.hier.transfers[l].ref, x[l + 1].data_ptr(), x[l].data_ptr(), nb, st)
            for _ in range(self._sweeps_at(l)):
                _smooth_multi(dev, k, dv, w, x[l], b[l], t[l], nb, 1, st)
                x[l], t[l] = t[l], x[l]
        return x[0]


def _smooth_multi(dev, k, dv, w, x, b, out, nb, mode, st):
    """A residual (mode 0) or update (mode 1) pass of the multi-column V-cycle:
    the structured-grid 16-column kernel (fem_mg_smooth16_grid) on node grids
    of that numbering, else fem_mg_smooth_multi."""
    grid = getattr(dev, "grid_nodes", None)
    if nb == 16 and grid is not None and x.data_ptr() % 16 == 0:
        _lib.call("fem_mg_smooth16_grid", dev.ref, grid[0], grid[1], grid[2], k, dv, w, x.data_ptr(), b.data_ptr(),
                  out.data_ptr(), mode, st)
    else:
        _lib.call("fem_mg_smooth_multi", dev.ref, k, dv, w, x.data_ptr(), b.data_ptr(), out.data_ptr(), nb, mode, st)


class FcgSteps:
    """Device state of the fused FCG(1) vector steps (fem_fcg_direction /
    fem_fcg_update) for nb column-interleaved vectors of n rows: the
    dot-product partials, the per-column scalars carried from one step to
    the next, and the squared residual norms ``rr`` (device, nb)."""

    def __init__(self, n, nb, dev):
        self.n, self.nb = int(n), int(nb)
        self.wb = int(_lib.lib().fem_fcg_work_bytes(self.nb))
        self.work = torch.empty(self.wb, dtype=torch.uint8, device=dev)
        self.rr = torch.empty(self.nb, dtype=torch.float64, device=dev)

    def _check(self, *ts):
        for t in ts:
            if t is not None and (t.numel() != self.n * self.nb or not t.is_contiguous()):
                raise ValueError(f"FCG vectors must be contiguous with {self.n} x {self.nb} entries")

    def direction(self, Z, mask, Qp, P, first):
        """P = mask Z - beta P (first: P = mask Z); Z FP32 or FP64."""
        if Z.dtype not in (torch.float32, torch.float64) or P.dtype != torch.float64:
            raise ValueError("FCG direction needs an FP32 or FP64 Z and an FP64 P")
        Z = Z.contiguous()
        self._check(Z, Qp, P)
        zdt = _lib.DTYPE_F32 if Z.dtype == torch.float32 else _lib.DTYPE_F64
        _lib.call("fem_fcg_direction", self.n, self.nb, Z.data_ptr(), zdt, _ptr(mask), _ptr(Qp), P.data_ptr(),
                  1 if first else 0, self.work.data_ptr(), self.wb, stream_ptr())

    def update(self, P, Q, mask, active, X, R, R32):
        """alpha = p.r / p.q (0 on inactive columns), X += alpha P,
        R -= alpha mask Q, R32 = (float) R; returns rr = |R|^2 per column."""
        self._check(P, Q, X, R, R32)
        _lib.call("fem_fcg_update", self.n, self.nb, P.data_ptr(), Q.data_ptr(), _ptr(mask), _ptr(active),
                  X.data_ptr(), R.data_ptr(), _ptr(R32), self.rr.data_ptr(), self.work.data_ptr(), self.wb,
                  stream_ptr())
        return self.rr

    def direction_pr(self, Z, mask, Qp, R, P, first):
        """direction, also leaving the partials of p.r (R: the current
        residual) for update_split."""
        if Z.dtype not in (torch.float32, torch.float64) or P.dtype != torch.float64:
            raise ValueError("FCG direction needs an FP32 or FP64 Z and an FP64 P")
        Z = Z.contiguous()
        self._check(Z, Qp, R, P)
        zdt = _lib.DTYPE_F32 if Z.dtype == torch.float32 else _lib.DTYPE_F64
        _lib.call("fem_fcg_direction_pr", self.n, self.nb, Z.data_ptr(), zdt, _ptr(mask), _ptr(Qp), R.data_ptr(),
                  P.data_ptr(), 1 if first else 0, self.work.data_ptr(), self.wb, stream_ptr())

    def update_split(self, P, Q, mask, active, pq_part, pq_blocks, X, R, R32):
        """update with p.q from pq_blocks partials pq_part (nb per block, as
        fem_bsr_spmm16_grid_pq leaves them) and p.r from direction_pr."""
        self._check(P, Q, X, R, R32)
        if pq_part.dtype != torch.float64 or pq_part.numel() < pq_blocks * self.nb:
            raise ValueError("pq_part must hold pq_blocks x nb doubles")
        _lib.call("fem_fcg_update_split", self.n, self.nb, P.data_ptr(), Q.data_ptr(), _ptr(mask), _ptr(active),
                  pq_part.data_ptr(), int(pq_blocks), X.data_ptr(), R.data_ptr(), _ptr(R32), self.rr.data_ptr(),
                  self.work.data_ptr(), self.wb, stream_ptr())
        return self.rr

    def residual(self, B, Y, mask, sel, R, R32):
        """t = (B - Y) mask (Y None: B mask); R = t and R32 = (float) t on the
        columns with sel[c] != 0 (sel None: all; R, R32 may be None).
        Returns (|t_c|, |B_c|) per column as numpy arrays (one pass)."""
        self._check(B, Y, R, R32)
        out = torch.empty(2 * self.nb, dtype=torch.float64, device=B.device)
        _lib.call("fem_fcg_residual", self.n, self.nb, B.data_ptr(), _ptr(Y), _ptr(mask), _ptr(sel), _ptr(R),
                  _ptr(R32), out.data_ptr(), self.work.data_ptr(), self.wb, stream_ptr())
        sq = np.sqrt(np.maximum(out.cpu().numpy(), 0.0))
        return sq[:self.nb], sq[self.nb:]


def _ptr(t):
    return None if t is None else t.data_ptr()


def solve_fcg_mg(model, K64, F, precond, tol=1e-8, maxiter=1000, check_every=1, drop=1e-2, free=None,
                 dinv64=None, patience=3, restart=False):
    """FP64 flexible CG (FCG(1)) with a nonlinear or inexact preconditioner,
    run on the free DOFs only.

    The penalised rows (diagonal 1e20) are kept out of the Krylov iteration:
    residuals, preconditioned residuals and K p are masked to the free DOFs,
    so the iteration solves K_ff x_f = F_f (the coupling K_fp x_p is ~1e-15
    relative, x_p being ~F / 1e20).  When the free residual is below
    tol ||F|| (checked against the true F - K x), the penalised DOFs are set
    by one block-Jacobi step on their rows, which is exact to rounding since
    their diagonal dominates, and the full FP64 residual is computed.  The
    recursive residual is replaced by the true one whenever it has dropped by
    ``drop`` (and the search direction restarts when ``restart``).  FP64
    cannot push the true residual below about kappa(K) * 1e-16: when
    ``patience`` replacements in a row fail to halve the best true residual,
    the iteration stops with ``stagnated`` set (the attainable accuracy, not
    a divergence).  ``free`` (the free-DOF mask) and ``dinv64`` (K64's
    block-Jacobi inverse) may be passed in when many systems share K.
    Returns (x, info dict)."""
    from .learned import free_dof_mask
    if free is None:
        free = free_dof_mask(model)
    fixed = 1.0 - free
    if dinv64 is None:
        dinv64, _ = K64.block_jacobi()
    d3 = dinv64.view(-1, 3, 3)
    bn = float(torch.linalg.vector_norm(F))
    x = torch.zeros_like(F)
    if bn == 0.0:
        return x, {"iterations": 0, "rel_residual": 0.0, "converged": True, "trace": []}
    n = F.numel()
    r = F * free
    anchor = float(torch.linalg.vector_norm(r))
    # the multigrid V-cycle runs on the FP32 copy of r that the update step
    # writes; any other preconditioner sees the FP64 r
    fast = isinstance(precond, MultigridPreconditioner)
    r32 = r.float() if fast else None
    steps = FcgSteps(n, 1, F.device)
    p = torch.empty_like(F)
    q, q_prev = torch.empty_like(F), torch.empty_like(F)
    first = True
    it = 0
    converged = stagnated = False
    best_true, stalls = math.inf, 0
    trace = []
    while it < maxiter:
        if fast:
            precond.applications += 1
            z = precond.vcycle(r32)
        else:
            z = precond(r)
        steps.direction(z, free, None if first else q_prev, p, first)
        K64.matvec(p, q)
        rr = steps.update(p, q, free, None, x, r, r32)
        q, q_prev = q_prev, q
        first = False
        it += 1
        if it % check_every:
            continue
        rf = math.sqrt(float(rr[0]))
        if not math.isfinite(rf):
            break
        trace.append((it, rf / bn))
        if rf < drop * anchor or rf < tol * bn:
            r.copy_((F - K64.matvec(x)) * free)
            if fast:
                r32.copy_(r)
            anchor = float(torch.linalg.vector_norm(r))
            if anchor < tol * bn:
                converged = True
                break
            if anchor < 0.5 * best_true:
                best_true, stalls = anchor, 0
            else:
                stalls += 1
                if stalls >= patience:
                    stagnated = True
                    break
            if restart:
                first = True
    # the penalised DOFs from their own rows
    rp = (F - K64.matvec(x)) * fixed
    x.add_(torch.bmm(d3, rp.view(-1, 3, 1)).view(-1) * fixed)
    rel = float(torch.linalg.vector_norm(F - K64.matvec(x))) / bn
    converged = converged and rel < tol
    return x, {"iterations": it, "rel_residual": rel, "converged": converged, "stagnated": stagnated,
               "trace": trace}


def solve_block_fcg_mg(model, K64, B, precond, tol=1e-8, maxiter=1000, free=None, dinv64=None, patience=3,
                       drop=3e-3, check_after=50, floor=1e-7, check_delay=6):
    """solve_fcg_mg for the nb columns of B (3N, nb) at once: per-column
    FCG(1) scalars, one multi-column V-cycle (precond.vcycle_multi) on the
    FP32 copy of the residual, one FP64 SpMM (fem_bsr_spmm, or on a
    structured grid with nb = 16 fem_bsr_spmm16_grid_pq, which also leaves
    p.q) and the fused vector steps (FcgSteps) per iteration.  Each column stops on its own
    (tolerance, or the attainable accuracy as in solve_fcg_mg) and is frozen
    from then on.  A column asks for a true-residual check (one product of
    all columns, which replaces the recursive residuals of the columns that
    asked) when its recursive residual has dropped by ``drop`` since its last
    one, and the check runs once a column has asked for ``check_delay``
    iterations, or at once when a column's recursive residual is below
    ``tol``; the columns that ask meanwhile share it.  On the 16-column inner
    solves of block Lanczos on the 500 x 60 x 75 blade a delay of 6 takes 67
    checks instead of 119 and 1.5% fewer iterations (2, 4 or 20: 86, 78 or
    59 checks), the inner solves 6% less time.  A drop of 3e-3 takes a third
    fewer checks than 1e-2 with no more iterations (1e-3 takes more; both
    measured without the delay).  A column whose recursive
    residual has not dropped by ``drop`` within ``check_after`` iterations of
    its last true-residual check gets one anyway, so a residual that levels off above ``tol`` still
    meets the stagnation test.  A column counts as at its attainable FP64
    accuracy (``stagnated``) only when ``patience`` checks in a row fail to
    halve its best true residual AND that residual is below ``floor`` times
    its right-hand side; a column that stalls above ``floor`` stays active,
    so it runs to ``maxiter`` and is neither converged nor stagnated (the
    caller's fallback finishes it).  Returns (X, info) with per-column
    ``converged``, ``stagnated`` and ``rel_residual``."""
    from . import solvers
    from .learned import free_dof_mask
    n, nb = B.shape
    if free is None:
        free = free_dof_mask(model)
    fixed = 1.0 - free
    if dinv64 is None:
        dinv64, _ = K64.block_jacobi()
    d3 = dinv64.view(-1, 3, 3)
    Y = torch.empty_like(B)

    def spmm(X):
        return solvers.bsr_spmm(K64.mesh, K64.vals, X, Y, nb)

    B = B.contiguous()
    steps = FcgSteps(n, nb, B.device)
    X = torch.zeros_like(B)
    R = torch.empty_like(B)
    R32 = torch.empty(B.shape, dtype=torch.float32, device=B.device)
    anchor, bn = steps.residual(B, None, free, None, R, R32)      # R = B masked to the free DOFs
    active = bn > 0
    converged = ~active
    stagnated = np.zeros(nb, dtype=bool)
    best = np.full(nb, np.inf)
    stalls = np.zeros(nb, dtype=int)
    last_check = np.zeros(nb, dtype=int)
    ask_since = np.full(nb, -1)        # iteration a column began asking (-1: not asking)
    act_d = torch.as_tensor(active.astype(np.int32), device=B.device)
    P = torch.empty_like(B)
    Q, Qp = torch.empty_like(B), torch.empty_like(B)
    # on a structured grid with 16 columns the product also leaves p.q
    # (fem_bsr_spmm16_grid_pq), and p.r comes with the direction: no extra
    # pass over P, Q and R
    grid = getattr(K64.mesh, "grid_nodes", None)
    fused = nb == 16 and grid is not None
    if fused:
        pq_blocks = int(_lib.lib().fem_spmm16_grid_partials(*grid))
        pq_part = torch.empty(pq_blocks * nb, dtype=torch.float64, device=B.device)
    first = True
    it = 0
    while it < maxiter and active.any():
        Z = precond.vcycle_multi(R32)
        if fused:
            steps.direction_pr(Z, free, None if first else Qp, R, P, first)
            _lib.call("fem_bsr_spmm16_grid_pq", K64.mesh.ref, grid[0], grid[1], grid[2], K64.vals.data_ptr(),
                      P.data_ptr(), Q.data_ptr(), pq_part.data_ptr(), pq_part.numel(), stream_ptr())
            rr = steps.update_split(P, Q, free, act_d, pq_part, pq_blocks, X, R, R32)
        else:
            steps.direction(Z, free, None if first else Qp, P, first)
            solvers.bsr_spmm(K64.mesh, K64.vals, P, Q, nb)
            rr = steps.update(P, Q, free, act_d, X, R, R32)
        Q, Qp = Qp, Q
        first = False
        it += 1
        rf = np.sqrt(rr.cpu().numpy())
        if not np.all(np.isfinite(rf[active])):
            break
        asked = active & ((rf < drop * anchor) | (it - last_check >= check_after))
        conv = active & (rf < tol * bn)
        ask_since[asked & (ask_since < 0)] = it
        ask_since[~asked] = -1
        due = conv.any() or bool((asked & (it - ask_since >= check_delay)).any())
        need = (asked | conv) if due else np.zeros(nb, dtype=bool)
        if need.any():
            ask_since[need] = -1
            # true residuals of the columns that asked for one replace their
            # recursive ones (one product of all columns, one fused pass)
            cols = np.nonzero(need)[0]
            sel = torch.as_tensor(need.astype(np.int32)).to(B.device)
            tn, _ = steps.residual(B, spmm(X), free, sel, R, R32)
            for c in cols:
                anchor[c] = tn[c]
                last_check[c] = it
                if tn[c] < tol * bn[c]:
                    converged[c] = True
                    active[c] = False
                elif tn[c] < 0.5 * best[c]:
                    best[c], stalls[c] = tn[c], 0
                else:
                    stalls[c] += 1
                    if stalls[c] >= patience and tn[c] < floor * bn[c]:
                        stagnated[c] = True
                        active[c] = False
            act_d.copy_(torch.as_tensor(active.astype(np.int32)))
    # the penalised DOFs from their own rows (the nodes that have one)
    fn = torch.nonzero(fixed.view(-1, 3).amax(dim=1) > 0).flatten()
    if fn.numel():
        spmm(X)
        fm = fixed.view(-1, 3)[fn][:, :, None]
        Rp = (B.view(-1, 3, nb)[fn] - Y.view(-1, 3, nb)[fn]) * fm
        X.view(-1, 3, nb).index_add_(0, fn, (d3[fn][:, :, :, None] * Rp[:, None, :, :]).sum(dim=2) * fm)
    tn, _ = steps.residual(B, spmm(X), None, None, None, None)
    res = tn / np.where(bn > 0, bn, 1.0)
    converged = converged & (res < tol)
    return X, {"iterations": it, "rel_residual": res, "converged": converged, "stagnated": stagnated}


def solve_mg(model, K64, F, tol=1e-8, maxiter=1000, min_nodes=1000, fallback=True, **kw):
    """Static solve: FP64 flexible CG preconditioned by the FP32 multigrid
    V-cycle; falls back to FP64 block-Jacobi PCG from the best iterate when
    it does not reach ``tol`` within ``maxiter``."""
    from . import solvers
    t0 = time.perf_counter()
    hier = hierarchy_for(model, min_nodes=min_nodes)
    t1 = time.perf_counter()
    pre = MultigridPreconditioner(model, hier)
    x, info = solve_fcg_mg(model, K64, F, pre, tol=tol, maxiter=maxiter, **kw)
    info["levels"] = [tuple(dm) for dm in hier.dims]
    info["hierarchy_seconds"] = hier.seconds          # symbolic, once per mesh
    info["hierarchy_built_now"] = t1 - t0 > 0.5 * hier.seconds
    info["setup_seconds"] = pre.setup_seconds
    info["fell_back"] = False
    if not info["converged"] and fallback:
        info["fell_back"] = True
        dinv64, _ = K64.block_jacobi()
        x, it2, _ = solvers.pcg(K64, F, dinv64, tol, 200000, x0=x.clone())
        x = x.clone()
        info["iterations"] += it2
        info["rel_residual"] = float(torch.linalg.vector_norm(F - K64.matvec(x)) / torch.linalg.vector_norm(F))
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t1
    return x, solvers.SolveInfo("mg", info["iterations"], 1, info["rel_residual"], seconds, [info])
