"""Learned preconditioner: a 3-layer 3x3x3 mesh-convolution network on the
structured node grid, used inside flexible PCG with block-Jacobi fallback.

The network (no counterpart in the reference) maps a residual field to an
error correction: input channels (r_x, r_y, r_z scaled, log E of the node's
elements), 64 hidden channels, ReLU, output 3 channels.  Its forward and
backward passes run in the hand-written BF16 MFMA kernels of meshconv.hip
(FP32 accumulation); master weights stay FP32 and are cast to BF16 for
every pass.  The preconditioner is z = Dinv r + s(r) * net(r / s(r)) / E_ref
(block-Jacobi plus the learned correction, s = RMS of r), applied inside
truncated flexible CG (FCG(m); SpMV on the FP32 copy of K with FP64 vectors,
or on the FP64 K; FP64 residual replacement).  If the FP64 residual stops
improving over ``patience`` replacements, or turns non-finite, the solve
falls back to FP64 block-Jacobi PCG from its best iterate.

Training data are the residual/error pairs of plain block-Jacobi PCG on
seeded slabs: the residuals r_k at selected iterations k of a PCG run and
the errors e_k = x* - x_k (x* from a tight FP64 solve).  The network is
trained on the preconditioned residual, loss = ||r - K z(r)||^2 / ||r||^2
(which equals ||K (e - z)||^2 / ||K e||^2, the error of the preconditioner
measured through K): the smooth, global part of e that no local operator can
represent contributes little to it, so the loss measures what a 7^3-node
receptive field can fix.  The learned correction is zero on penalised
(fixed) DOFs, where block-Jacobi is already exact.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np
import torch

from . import _lib
from .runtime import device, stream_ptr

E_REF = 200000.0
CIN0 = 32          # padded input channels (4 used)
HID = 64
COUT = 16          # padded output channels (3 used)


def _bf(t):
    return t.detach().to(torch.bfloat16).contiguous()


@dataclasses.dataclass
class ConvWork:
    nx: int
    ny: int
    nz: int

    @property
    def P(self):
        return self.nx * self.ny * self.nz


class MeshConvNet:
    """3 layers of 3x3x3 convolution: 4 -> 64 -> 64 -> 3 (ReLU between)."""

    def __init__(self, seed=0, hidden=HID):
        if hidden != HID:
            raise ValueError(f"the mesh-convolution kernels are built for {HID} hidden channels, not {hidden}")
        g = torch.Generator().manual_seed(seed)
        d = device()
        def init(cout, cin, fan_in, gain):
            w = torch.randn(27, cout, cin, generator=g) * (gain / math.sqrt(fan_in))
            return w.to(d)
        self.w = [init(HID, CIN0, 27 * 4, math.sqrt(2.0)), init(HID, HID, 27 * HID, math.sqrt(2.0)),
                  init(COUT, HID, 27 * HID, 0.1)]
        # zero the padded channels
        self.w[0][:, :, 4:] = 0
        self.w[2][:, 3:, :] = 0
        self.b = [torch.zeros(HID, device=d), torch.zeros(HID, device=d), torch.zeros(COUT, device=d)]
        self.cin = [CIN0, HID, HID]
        self.cout = [HID, HID, COUT]

    def parameters(self):
        return self.w + self.b

    def state_dict(self):
        return {f"w{i}": w.detach().cpu() for i, w in enumerate(self.w)} | \
               {f"b{i}": b.detach().cpu() for i, b in enumerate(self.b)}

    def load_state_dict(self, sd):
        d = device()
        for i in range(3):
            self.w[i] = sd[f"w{i}"].to(d).float()
            self.b[i] = sd[f"b{i}"].to(d).float()
        self._infer = None

    def _infer_weights(self):
        """BF16 weights of the inference pass (the first layer packed for
        fem_meshconv_fwd_c4_bf16), rebuilt when a weight tensor changes."""
        key = tuple((id(w), w._version) for w in self.w)
        cache = getattr(self, "_infer", None)
        if cache is None or cache[0] != key:
            wb = [_bf(w) for w in self.w]
            wp = torch.empty(4, HID, 32, dtype=torch.bfloat16, device=wb[0].device)
            _lib.call("fem_meshconv_pack_c4_bf16", wb[0].data_ptr(), CIN0, HID, wp.data_ptr(), stream_ptr())
            self._infer = cache = (key, wp, wb[1], wb[2])
        return cache[1:]

    def forward_infer(self, x4, grid: ConvWork):
        """x4: (P, 4) bf16 input -> (P, 16) f32 output (3 channels used); the
        first layer runs packed (4 MFMA k-steps per tile instead of 27)."""
        st = stream_ptr()
        P = grid.P
        d = device()
        wp, w1, w2 = self._infer_weights()
        a1 = torch.empty(P, HID, dtype=torch.bfloat16, device=d)
        _lib.call("fem_meshconv_fwd_c4_bf16", x4.data_ptr(), wp.data_ptr(), self.b[0].data_ptr(), HID, grid.nx,
                  grid.ny, grid.nz, a1.data_ptr(), None, 1, st)
        a2 = torch.empty(P, HID, dtype=torch.bfloat16, device=d)
        _lib.call("fem_meshconv_fwd_bf16", a1.data_ptr(), HID, w1.data_ptr(), self.b[1].data_ptr(), HID, grid.nx,
                  grid.ny, grid.nz, a2.data_ptr(), None, None, 1, st)
        y = torch.empty(P, COUT, dtype=torch.float32, device=d)
        _lib.call("fem_meshconv_fwd_bf16", a2.data_ptr(), HID, w2.data_ptr(), self.b[2].data_ptr(), COUT, grid.nx,
                  grid.ny, grid.nz, None, None, y.data_ptr(), 0, st)
        return y

    # ------------------------------------------------------------- passes
    def forward(self, x0, grid: ConvWork, keep=False):
        """x0: (P, 32) bf16 input -> (P, 16) f32 output (3 channels used)."""
        st = stream_ptr()
        P = grid.P
        d = device()
        wb = [_bf(w) for w in self.w]
        acts = [x0]
        pres = []
        h = x0
        for li in range(2):
            a = torch.empty(P, HID, dtype=torch.bfloat16, device=d)
            z = torch.empty(P, HID, dtype=torch.bfloat16, device=d) if keep else None
            _lib.call("fem_meshconv_fwd_bf16", h.data_ptr(), self.cin[li], wb[li].data_ptr(),
                      self.b[li].data_ptr(), self.cout[li], grid.nx, grid.ny, grid.nz, a.data_ptr(),
                      z.data_ptr() if z is not None else None, None, 1, st)
            acts.append(a)
            pres.append(z)
            h = a
        y = torch.empty(P, COUT, dtype=torch.float32, device=d)
        _lib.call("fem_meshconv_fwd_bf16", h.data_ptr(), self.cin[2], wb[2].data_ptr(), self.b[2].data_ptr(),
                  COUT, grid.nx, grid.ny, grid.nz, None, None, y.data_ptr(), 0, st)
        if keep:
            self._cache = (wb, acts, pres, grid)
        return y

    def backward(self, dy):
        """dy: (P, 16) f32 gradient of the output -> grads for w (27,co,ci) and b."""
        wb, acts, pres, grid = self._cache
        st = stream_ptr()
        d = device()
        P = grid.P
        chunk = 2048
        gw = [None] * 3
        gb = [None] * 3

        def wgrad(x, cin, dyb, cout):
            out = torch.empty(27, cout, cin, dtype=torch.float32, device=d)
            nbytes = _lib.lib().fem_meshconv_wgrad_work_bytes(cin, cout, grid.nx, grid.ny, grid.nz, chunk)
            work = torch.empty(nbytes, dtype=torch.uint8, device=d)
            _lib.call("fem_meshconv_wgrad_bf16", x.data_ptr(), cin, dyb.data_ptr(), cout, grid.nx, grid.ny, grid.nz,
                      chunk, out.data_ptr(), work.data_ptr(), nbytes, st)
            return out

        def colsum(a, is_bf16, C):
            out = torch.empty(C, dtype=torch.float32, device=d)
            nbytes = _lib.lib().fem_colsum_work_bytes(P, C, 1024)
            work = torch.empty(nbytes, dtype=torch.uint8, device=d)
            _lib.call("fem_colsum", a.data_ptr(), int(is_bf16), P, C, 1024, out.data_ptr(), work.data_ptr(), nbytes, st)
            return out

        def flip(w, cout, cin, cout_t, cin_t):
            wt = torch.empty(27, cout_t, cin_t, dtype=torch.bfloat16, device=d)
            _lib.call("fem_meshconv_flip_bf16", w.data_ptr(), cout, cin, wt.data_ptr(), cout_t, cin_t, st)
            return wt

        dy = dy.contiguous()
        dy16 = torch.empty(P, COUT, dtype=torch.bfloat16, device=d)
        _lib.call("fem_to_bf16", dy.data_ptr(), P, COUT, COUT, dy16.data_ptr(), st)
        dy32 = torch.empty(P, 32, dtype=torch.bfloat16, device=d)
        _lib.call("fem_to_bf16", dy.data_ptr(), P, COUT, 32, dy32.data_ptr(), st)
        gw[2] = wgrad(acts[2], HID, dy16, COUT)
        gb[2] = colsum(dy, False, COUT)
        # da2 = conv^T(dy) ; dz2 = da2 * (z2 > 0)
        w3t = flip(wb[2], COUT, HID, HID, 32)
        da = torch.empty(P, HID, dtype=torch.bfloat16, device=d)
        _lib.call("fem_meshconv_fwd_bf16", dy32.data_ptr(), 32, w3t.data_ptr(), None, HID, grid.nx, grid.ny,
                  grid.nz, None, da.data_ptr(), None, 0, st)
        dz = torch.empty_like(da)
        _lib.call("fem_relu_bwd_bf16", da.data_ptr(), pres[1].data_ptr(), da.numel(), dz.data_ptr(), st)
        gw[1] = wgrad(acts[1], HID, dz, HID)
        gb[1] = colsum(dz, True, HID)
        w2t = flip(wb[1], HID, HID, HID, HID)
        da1 = torch.empty(P, HID, dtype=torch.bfloat16, device=d)
        _lib.call("fem_meshconv_fwd_bf16", dz.data_ptr(), HID, w2t.data_ptr(), None, HID, grid.nx, grid.ny, grid.nz,
                  None, da1.data_ptr(), None, 0, st)
        dz1 = torch.empty_like(da1)
        _lib.call("fem_relu_bwd_bf16", da1.data_ptr(), pres[0].data_ptr(), da1.numel(), dz1.data_ptr(), st)
        gw[0] = wgrad(acts[0], CIN0, dz1, HID)
        gb[0] = colsum(dz1, True, HID)
        return gw, gb


# ---------------------------------------------------------------------------
# preconditioner and flexible CG
# ---------------------------------------------------------------------------

def node_log_modulus(phys, EX):
    """Per-node mean of log(E_e / E_REF) over the node's elements, (N,) f32."""
    ex = np.asarray(EX, dtype=np.float64)
    N, E = phys.Num_Node, phys.Num_Elem
    le = np.full(E, np.log(float(ex) / E_REF)) if ex.ndim == 0 else np.log(ex / E_REF)
    en = phys.elem_nodes.reshape(-1)
    s = np.bincount(en, weights=np.repeat(le, 8), minlength=N)
    c = np.maximum(np.bincount(en, minlength=N), 1)
    return torch.as_tensor((s / c).astype(np.float32)).to(device())


def free_dof_mask(model):
    """(3N,) f64: 0 on DOFs the penalty fixes (flag != -1), 1 elsewhere."""
    m = torch.ones(3 * model.Num_Node, dtype=torch.float64, device=device())
    bn = model.bound_node.long()
    fl = model.bound_flag.view(-1, 3)
    for j in range(3):
        sel = bn[fl[:, j] != -1]
        m[3 * sel + j] = 0.0
    return m


class LearnedPreconditioner:
    """z = Dinv r + s * net(r / s, log E) / E_REF with s = RMS(r) (block-Jacobi
    base), or, with a ``base`` preconditioner B (the multigrid V-cycle):
    z = B r + tau c, c = s * net(r_a / s, log E) / E_REF for the residual
    r_a = r - K B r the base leaves, and tau = c.r_a / c.K c, the step that
    minimises the K-norm of the error along c (so the learned term can only
    lower it)."""

    def __init__(self, net: MeshConvNet, node_dims, efeat, dinv64, free_mask=None, use_net=True, base=None, K=None,
                 min_gain=None, gain_patience=3):
        self.net = net
        self.grid = ConvWork(*node_dims)
        self.efeat = efeat
        self.dinv = dinv64.view(-1, 3, 3)
        self.mask = free_mask        # (3N,) 1 on free DOFs, 0 on penalised ones
        self.use_net = use_net
        self.base = base
        self.K = K
        if base is not None and K is None:
            raise ValueError("a base preconditioner needs the operator K for its residual")
        self.applications = 0
        self.taus = []
        self.gains = []
        self.min_gain = min_gain
        self.gain_patience = gain_patience
        self.low = 0
        self.net_off = False
        self._pending = None           # device gain of the last application, read at the next one

    def features(self, r):
        P = self.grid.P
        s = torch.sqrt(torch.mean(r * r)).clamp_min(1e-300)
        x0 = torch.zeros(P, CIN0, dtype=torch.bfloat16, device=r.device)
        x0[:, :3] = (r.view(P, 3) / s).to(torch.bfloat16)
        x0[:, 3] = self.efeat.to(torch.bfloat16)
        return x0, s

    def bj(self, r):
        return torch.bmm(self.dinv, r.view(-1, 3, 1)).view(-1)

    def __call__(self, r):
        self.applications += 1
        if self.base is None:
            z = self.bj(r)
            if not self.use_net:
                return z
            return z + self.correction(r)
        self._read_gain()
        z = self.base(r)
        if not self.use_net or self.net_off:
            return z
        ra = r - self.K.matvec(z)
        c = self.correction(ra)
        kc = self.K.matvec(c)
        cra = torch.dot(c, ra)
        tau = torch.nan_to_num(cra / torch.dot(c, kc), nan=0.0, posinf=0.0, neginf=0.0)
        self.taus.append(tau)
        if self.min_gain is not None:
            # the error energy the correction removes, tau (c.r_a), against the
            # energy r.z of the V-cycle's own step; it is read at the next
            # application (no host wait here), and after ``gain_patience``
            # applications in a row below ``min_gain`` the network is switched
            # off for the rest of the solve (the V-cycle alone from then on)
            self._pending = torch.nan_to_num(tau * cra / torch.dot(r, z), nan=0.0)
        return z + tau * c

    def _read_gain(self):
        if self._pending is None:
            return
        gain = float(self._pending)
        self._pending = None
        self.gains.append(gain)
        self.low = self.low + 1 if gain < self.min_gain else 0
        if self.low >= self.gain_patience:
            self.net_off = True

    def correction(self, r, keep=False):
        """The masked correction c = s * net(r / s, log E) / E_REF, s = RMS(r).
        keep=False (inference): the packed first layer and fused input and
        output kernels; keep=True (training): the padded pass that keeps the
        activations for the backward kernels."""
        if keep:
            x0, s = self.features(r)
            y = self.net.forward(x0, self.grid, keep=True)
            c = y[:, :3].reshape(-1).double() * (s / E_REF)
            if self.mask is not None:
                c = c * self.mask
            return c
        st = stream_ptr()
        P = self.grid.P
        s = torch.sqrt(torch.mean(r * r)).clamp_min(1e-300)
        x4 = torch.empty(P, 4, dtype=torch.bfloat16, device=r.device)
        _lib.call("fem_meshconv_features_c4", r.data_ptr(), P, s.data_ptr(), self.efeat.data_ptr(), x4.data_ptr(), st)
        y = self.net.forward_infer(x4, self.grid)
        c = torch.empty(3 * P, dtype=torch.float64, device=r.device)
        _lib.call("fem_meshconv_correction", y.data_ptr(), COUT, P, s.data_ptr(), E_REF,
                  self.mask.data_ptr() if self.mask is not None else None, c.data_ptr(), st)
        return c


def solve_fcg(K64, K32, F, precond, tol=1e-8, maxiter=100000, check_every=8, drop=1e-2, patience=4,
              stagnation=100, mmax=4, trace=None):
    """Truncated flexible CG (FCG(m), Notay 2000): each new direction is
    M-orthogonalised (in the K inner product) against the last ``mmax``
    directions, which keeps CG-like convergence with a nonlinear
    preconditioner.  The SpMV streams ``K32`` (FP32 values, FP64 vectors) when
    it is given, else the FP64 ``K64``.  The recursive residual is replaced by
    the FP64 ``F - K64 x`` whenever it has dropped by ``drop`` or claims
    convergence (reliable updates).  The flexible iteration is abandoned for
    FP64 block-Jacobi PCG (from its best iterate, or from zero; with its own
    ``maxiter``) when ``patience`` consecutive replacements fail to halve the
    best FP64 residual, when the recursive residual sets no new low (by 10%)
    in ``stagnation`` consecutive checks (None: no such test; CG-type
    residual norms can plateau for long), when it grows 1e4-fold above its
    lowest value, or when it turns non-finite.
    Returns (x, info dict)."""
    from . import solvers
    st = stream_ptr()
    if K32 is not None:
        mesh_ref = K32.mesh.ref

        def spmv(x, y):
            _lib.call("fem_bsr_spmv", mesh_ref, K32.vals.data_ptr(), _lib.DTYPE_F32F64, x.data_ptr(),
                      y.data_ptr(), st)
            return y
    else:
        def spmv(x, y):
            return K64.matvec(x, y)

    x = torch.zeros_like(F)
    r = F.clone()
    bb = float(torch.dot(F, F))
    hist = []          # (p_j, q_j = K p_j, p_j . q_j)
    it = 0
    rr_anchor = bb
    best_rr, best_x, stalls = bb, None, 0
    low_rr, flat = bb, 0
    rel = 1.0 if bb > 0 else 0.0
    while it < maxiter and rel >= tol:
        z = precond(r)
        p = z
        for pj, qj, pqj in hist:
            p = p - (torch.dot(z, qj) / pqj) * pj
        q = spmv(p.contiguous(), torch.empty_like(F))
        pq = torch.dot(p, q)
        alpha = torch.dot(p, r) / pq
        x.add_(alpha * p)
        r.sub_(alpha * q)
        hist.append((p, q, pq))
        if len(hist) > mmax:
            hist.pop(0)
        it += 1
        if it % check_every:
            continue
        rr = float(torch.dot(r, r))
        if not math.isfinite(rr):
            break
        if rr < 0.81 * low_rr:
            low_rr, flat = rr, 0
        else:
            flat += 1
            if stagnation is not None and flat >= stagnation:
                break
        if rr > 1e8 * low_rr:           # the recursion diverges
            break
        rel = (rr / bb) ** 0.5 if bb > 0 else 0.0
        if rel < tol or rr < drop * drop * rr_anchor:
            r = F - K64.matvec(x)
            rr = float(torch.dot(r, r))
            rr_anchor = rr
            rel = (rr / bb) ** 0.5
            if trace is not None:
                trace.append((it, rel))
            if not math.isfinite(rr):
                break
            if rr < 0.25 * best_rr:
                best_rr, best_x, stalls = rr, x.clone(), 0
            else:
                stalls += 1
                if stalls >= patience:
                    break
    fell_back = False
    rel = float(torch.linalg.vector_norm(F - K64.matvec(x)) / math.sqrt(bb)) if bb > 0 else 0.0
    if not (rel < tol):
        fell_back = True
        if not (math.isfinite(rel) and rel < 1.0):
            x = best_x if best_x is not None else torch.zeros_like(F)
        dinv64, _ = K64.block_jacobi()
        x, it2, _ = solvers.pcg(K64, F, dinv64, tol, maxiter, x0=x.clone())
        x = x.clone()
        it += it2
        rel = float(torch.linalg.vector_norm(F - K64.matvec(x)) / math.sqrt(bb))
    return x, {"iterations": it, "rel_residual": rel, "fell_back": fell_back,
               "applications": getattr(precond, "applications", None)}


def solve_learned(model, K64, F, tol=1e-8, net=None, weights=None, maxiter=100000, use_net=True,
                  matrix="f32", base="mg", min_gain=0.02, gain_patience=3, **kw):
    """Static solve with the learned preconditioner (structured meshes).

    base "mg": flexible CG (multigrid.solve_fcg_mg) preconditioned by the FP32
    multigrid V-cycle plus the learned correction of the residual it leaves
    (weights of the "mg" role); when that does not reach ``tol`` in
    ``maxiter`` iterations, FP64 block-Jacobi PCG finishes from its iterate.
    The correction is dropped for the rest of the solve once it has removed
    less than ``min_gain`` of the V-cycle step's energy in ``gain_patience``
    applications in a row (min_gain None: never).
    base "bj": truncated flexible CG (solve_fcg) with block-Jacobi plus the
    learned correction (weights of the "bj" role); ``matrix`` "f32" streams
    the FP32 copy of K in that iteration, "f64" the FP64 one."""
    from . import solvers
    phys = model.Physics
    if getattr(phys, "node_dims", None) is None:
        raise ValueError("the learned preconditioner needs a structured node grid (Physics.node_dims)")
    if base not in ("mg", "bj"):
        raise ValueError(f"unknown base {base!r}")
    t0 = time.perf_counter()
    if net is None:
        net = MeshConvNet()
        if weights is None:
            weights = default_weights_path(base)
        net.load_state_dict(torch.load(weights, weights_only=True))
    dinv64, _ = K64.block_jacobi()
    efeat = node_log_modulus(phys, model.EX)
    if base == "mg":
        from . import multigrid
        hier = multigrid.hierarchy_for(model)
        t0 = time.perf_counter()                # the symbolic hierarchy is per-mesh preprocessing
        mg = multigrid.MultigridPreconditioner(model, hier)
        pre = LearnedPreconditioner(net, phys.node_dims, efeat, dinv64, free_mask=free_dof_mask(model),
                                    use_net=use_net, base=mg, K=K64, min_gain=min_gain, gain_patience=gain_patience)
        x, info = multigrid.solve_fcg_mg(model, K64, F, pre, tol=tol, maxiter=min(maxiter, 1000), **kw)
        pre._read_gain()
        info["fell_back"] = False
        info["applications"] = pre.applications
        info["net_applications"] = len(pre.taus)
        info["taus"] = [float(v) for v in pre.taus[:64]]
        info["gains"] = [round(v, 5) for v in pre.gains[:64]]
        if not info["converged"]:
            info["fell_back"] = True
            x, it2, _ = solvers.pcg(K64, F, dinv64, tol, maxiter, x0=x.clone())
            x = x.clone()
            info["iterations"] += it2
            info["rel_residual"] = float(torch.linalg.vector_norm(F - K64.matvec(x)) /
                                         torch.linalg.vector_norm(F))
        torch.cuda.synchronize()
        return x, solvers.SolveInfo("learned_mg", info["iterations"], 1, info["rel_residual"],
                                    time.perf_counter() - t0, [info])
    K32 = solvers.BsrOperator(model.dev, model.K_bsr("f32")) if matrix == "f32" else None
    pre = LearnedPreconditioner(net, phys.node_dims, efeat, dinv64, free_mask=free_dof_mask(model), use_net=use_net)
    x, info = solve_fcg(K64, K32, F, pre, tol=tol, maxiter=maxiter, **kw)
    torch.cuda.synchronize()
    return x, solvers.SolveInfo("learned", info["iterations"], 1, info["rel_residual"],
                                time.perf_counter() - t0, [info])


def default_weights_path(role="bj"):
    """Shipped weights: trained beside block-Jacobi ("bj") or after the multigrid V-cycle ("mg")."""
    import os
    name = {"bj": "meshconv_slab.pt", "mg": "meshconv_mg.pt"}[role]
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", name)


# ---------------------------------------------------------------------------
# training data and training
# ---------------------------------------------------------------------------

def pcg_trajectory(K64, b, dinv64, snaps, maxiter):
    """Plain FP64 block-Jacobi PCG from 0, returning {k: (x_k, r_k)} at the
    requested iteration counts."""
    from . import solvers
    N = K64.mesh.num_node
    ws = solvers.PcgWorkspace(N, torch.float64)
    m = K64.mesh.ref
    st = stream_ptr()
    _lib.call("fem_pcg_start", m, K64.vals.data_ptr(), dinv64.data_ptr(), _lib.DTYPE_F64, b.data_ptr(),
              ws.x.data_ptr(), ws.r.data_ptr(), ws.z.data_ptr(), ws.p.data_ptr(), ws.state.data_ptr(),
              ws.work.data_ptr(), ws.wb, st)
    out = {}
    want = sorted(set(snaps))
    for it in range(0, maxiter + 1):
        if it in want:
            out[it] = (ws.x.clone(), ws.r.clone())
            if it == want[-1]:
                break
        _lib.call("fem_pcg_iterate", m, K64.vals.data_ptr(), dinv64.data_ptr(), _lib.DTYPE_F64, ws.x.data_ptr(),
                  ws.r.data_ptr(), ws.z.data_ptr(), ws.p.data_ptr(), ws.q.data_ptr(), ws.state.data_ptr(),
                  ws.work.data_ptr(), ws.wb, st)
    return out


class TrainingSlab:
    """One training slab on the device: K (FP64 BSR), the preconditioner
    pieces and residual snapshots.

    role "bj": the residuals r_k of plain block-Jacobi PCG runs at the
    iterations ``snaps``; the network learns the correction added to Dinv r.
    role "mg": the residuals r_a = r_k - K MG(r_k) that the multigrid V-cycle
    leaves at the iterations ``snaps`` of multigrid flexible CG runs; the
    network learns the correction added to MG(r_k)."""

    def __init__(self, grid_mesh, snaps=None, rhs_seeds=(0,), with_physical=True, role="bj"):
        from . import solvers
        from .model import Model
        from .physics import Physics
        if role not in ("bj", "mg"):
            raise ValueError(f"unknown role {role!r}")
        self.role = role
        self.phys = Physics.from_grid(grid_mesh)
        self.model = Model(grid_mesh.name, self.phys, grid_mesh.Materials, Omegas=grid_mesh.Omegas)
        self.K64 = solvers.BsrOperator(self.model.dev, self.model.K_bsr("f64"))
        dinv64, _ = self.K64.block_jacobi()
        self.pre = LearnedPreconditioner(None, self.phys.node_dims, node_log_modulus(self.phys, self.model.EX),
                                         dinv64, free_mask=free_dof_mask(self.model), use_net=True)
        rhs = [self.model.F_device()] if with_physical else []
        n = 3 * self.phys.Num_Node
        for sd in rhs_seeds:
            g = torch.Generator(device="cpu").manual_seed(int(sd))
            rhs.append(torch.randn(n, generator=g, dtype=torch.float64).to(device()))
        self.residuals = []
        if role == "bj":
            snaps = snaps or (0, 3, 10, 30, 100, 300, 1000)
            for b in rhs:
                traj = pcg_trajectory(self.K64, b, dinv64, snaps, max(snaps))
                self.residuals += [rk for _, (_, rk) in sorted(traj.items())]
            return
        from . import multigrid
        snaps = set(snaps or (0, 1, 2, 4, 7, 11, 16))
        mg = multigrid.MultigridPreconditioner(self.model, multigrid.hierarchy_for(self.model))
        count = [0]

        def recording(r):
            z = mg(r)
            if count[0] in snaps:
                self.residuals.append((r - self.K64.matvec(z)).contiguous())
            count[0] += 1
            return z

        for b in rhs:
            count[0] = 0
            multigrid.solve_fcg_mg(self.model, self.K64, b, recording, tol=1e-12, maxiter=max(snaps) + 1)
        # the energy e.K.e = e.r_a of each residual's error e = K^-1 r_a (tight multigrid solves)
        self.energies = []
        for ra in self.residuals:
            e, info = multigrid.solve_fcg_mg(self.model, self.K64, ra, mg, tol=1e-10, maxiter=200)
            self.energies.append(float(torch.dot(e, ra)))

    def loss_and_grad(self, net, r, energy=None):
        """The loss and dloss/d(net output) (P, 16).

        role "bj": loss = ||r - K z||^2 / ||r||^2 with z = Dinv r + c.
        role "mg": loss = min over tau of ||e - tau c||_K^2 / ||e||_K^2
        = 1 - (c.r)^2 / ((c.K c) (e.r)) for the error e = K^-1 r of the
        residual r the V-cycle leaves (only the scalar e.r = ``energy`` is
        needed): the share of the error's energy that the correction, taken
        with the step the preconditioner uses, leaves behind.  Scale-free in
        c, and measured in the norm flexible CG minimises."""
        self.pre.net = net
        x0, s = self.pre.features(r)
        y = net.forward(x0, self.pre.grid, keep=True)
        c = y[:, :3].reshape(-1).double() * (s / E_REF) * self.pre.mask
        if self.role == "bj":
            z = self.pre.bj(r) + c
            res = r - self.K64.matvec(z.contiguous())
            rn = torch.dot(r, r)
            loss = torch.dot(res, res) / rn
            g = self.K64.matvec(res.contiguous()) * (-2.0 / rn)          # dloss/dz (K symmetric)
        else:
            if energy is None or not energy > 0:
                raise ValueError("role 'mg' needs the residual's error energy e.r > 0")
            kc = self.K64.matvec(c.contiguous())
            a = torch.dot(c, r)
            b = torch.dot(c, kc).clamp_min(1e-300)
            loss = 1.0 - a * a / (b * energy)
            g = (r - (a / b) * kc) * (-2.0 * a / (b * energy))            # dloss/dc
        dy = torch.zeros_like(y)
        dy[:, :3] = (g * self.pre.mask * (s / E_REF)).view(-1, 3).float()
        return loss, dy

    def samples(self):
        """(residual, energy) pairs (energy None for role "bj")."""
        if self.role == "bj":
            return [(r, None) for r in self.residuals]
        return list(zip(self.residuals, self.energies))

    def mean_loss(self, net):
        return float(np.mean([float(self.loss_and_grad(net, r, en)[0]) for r, en in self.samples()]))


def train(net, slabs, steps=400, lr=1e-3, seed=0, log_every=50, log=None):
    """Adam on the preconditioned-residual loss, gradients from the BF16
    backward kernels.  slabs: list of TrainingSlab."""
    for w in net.w + net.b:
        w.requires_grad_(True)
    opt = torch.optim.Adam(net.w + net.b, lr=lr)
    rng = np.random.default_rng(seed)
    flat = [(sl, r, en) for sl in slabs for r, en in sl.samples()]
    hist = []
    for step in range(steps):
        sl, r, en = flat[rng.integers(len(flat))]
        loss, dy = sl.loss_and_grad(net, r, en)
        gw, gb = net.backward(dy)
        gw[0][:, :, 4:] = 0
        gw[2][:, 3:, :] = 0
        gb[2][3:] = 0
        for p_, g_ in zip(net.w + net.b, gw + gb):
            p_.grad = g_
        opt.step()
        hist.append(float(loss))
        if log is not None and (step % log_every == 0 or step == steps - 1):
            log(f"step {step} loss {np.mean(hist[-log_every:]):.4f}")
    for w in net.w + net.b:
        w.requires_grad_(False)
    return hist


def bj_loss(slab):
    """The same loss with the learned correction switched off (block-Jacobi;
    1 for role "mg", where the residuals are already the base's)."""
    if slab.role == "mg":
        return 1.0
    vals = []
    for r in slab.residuals:
        res = r - slab.K64.matvec(slab.pre.bj(r).contiguous())
        vals.append(float(torch.dot(res, res) / torch.dot(r, r)))
    return float(np.mean(vals))
