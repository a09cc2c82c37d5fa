"""CPU restatement of the learned preconditioner's network (TEST
INFRASTRUCTURE ONLY): torch.nn.functional.conv3d in FP64 on the BF16-rounded
weights and input, used to check the HIP forward pass and, through torch
autograd, the HIP backward kernels.

With ``emulate_bf16=True`` the restatement rounds to BF16 exactly where the
HIP pipeline stores BF16: the hidden activations (forward) and the gradients
that flow back into them and out of the last layer (backward); everything
else is FP64.  The HIP result then differs only by FP32 accumulation order.
"""
from __future__ import annotations

import torch

CIN0, HID, COUT = 32, 64, 16


def _bf16(t):
    return t.to(torch.bfloat16).to(t.dtype)


class _RoundBoth(torch.autograd.Function):
    """Round to BF16 forward and round the incoming gradient backward."""

    @staticmethod
    def forward(ctx, x):
        return _bf16(x)

    @staticmethod
    def backward(ctx, g):
        return _bf16(g)


class _RoundGrad(torch.autograd.Function):
    """Identity forward, BF16-rounded gradient backward."""

    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return _bf16(g)


def torch_reference_forward(net, x0, grid, dtype=torch.float64, emulate_bf16=False):
    """Returns (y (P, 16), weight leaves, bias leaves); call backward on y."""
    import torch.nn.functional as Fn

    def conv(h, w, cin, cout):
        # h (P, cin) -> (1, cin, nz, ny, nx); weight (cout, cin, 3, 3, 3) from w[tap][co][ci]
        hh = h.t().reshape(1, cin, grid.nz, grid.ny, grid.nx)
        ww = w.reshape(3, 3, 3, cout, cin).permute(3, 4, 0, 1, 2)
        y = Fn.conv3d(hh, ww, None, padding=1)
        return y.reshape(cout, -1).t()

    ws = [w.detach().to(torch.bfloat16).to(dtype).cpu().requires_grad_(True) for w in net.w]
    bs = [b.detach().to(dtype).cpu().requires_grad_(True) for b in net.b]
    h = x0.detach().to(dtype).cpu()
    a = h
    for li, (cin, cout) in enumerate(((CIN0, HID), (HID, HID))):
        z = conv(a, ws[li], cin, cout) + bs[li]
        a = torch.relu(z)
        if emulate_bf16:
            a = _RoundBoth.apply(a)
    c = conv(a, ws[2], HID, COUT)
    if emulate_bf16:
        c = _RoundGrad.apply(c)
    y = c + bs[2]
    return y, ws, bs
