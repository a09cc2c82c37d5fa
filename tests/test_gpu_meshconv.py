"""BF16 mesh-convolution kernels (forward, backward-data, weight and bias
gradients) against the FP64 CPU restatement in oracle/meshconv_ref.py on the
same BF16-rounded weights and input.  Tolerances: 1e-3 relative Frobenius
against the restatement that rounds to BF16 where the kernels do (FP32
accumulation order only), 1e-1 against the exact FP64 network (BF16 storage
of hidden activations and backpropagated gradients)."""
import numpy as np
import pytest
import torch

from oracle import meshconv_ref as R

pytestmark = pytest.mark.gpu
EMU_TOL = 1e-3      # vs the restatement rounding to BF16 at the same points
TOL = 1e-1          # vs the exact FP64 network


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()


@pytest.mark.parametrize("dims", [(9, 7, 5), (16, 4, 3)])
def test_forward_and_gradients(gpu, dims):
    from fem_taichi_amd.learned import MeshConvNet, ConvWork
    nx, ny, nz = dims
    print(f"\ncase meshconv grid {dims}", flush=True)
    g = ConvWork(nx, ny, nz)
    net = MeshConvNet(seed=1)
    for b in net.b:
        b.normal_(0, 0.1)
    net.b[2][3:] = 0
    gen = torch.Generator().manual_seed(2)
    x = torch.zeros(g.P, 32)
    x[:, :4] = torch.randn(g.P, 4, generator=gen)
    x0 = x.to(torch.bfloat16).cuda()
    y = net.forward(x0, g, keep=True)
    dy = torch.zeros(g.P, 16)
    dy[:, :3] = torch.randn(g.P, 3, generator=gen)
    gw, gb = net.backward(dy.cuda())
    # (a) against the restatement that rounds where the kernels round:
    #     only FP32 accumulation order differs
    ye, wse, bse = R.torch_reference_forward(net, x0, g, emulate_bf16=True)
    (ye * dy.double()).sum().backward()
    errs = [rel(y[:, :3], ye[:, :3])] + [rel(gw[i], wse[i].grad) for i in range(3)] + \
           [rel(gb[i], bse[i].grad) for i in range(3)]
    print("emulated-bf16 rel errors", errs, flush=True)
    assert max(errs) < EMU_TOL
    # (b) against the exact FP64 network: BF16 storage error only
    yr, ws, bs = R.torch_reference_forward(net, x0, g)
    (yr * dy.double()).sum().backward()
    errs = [rel(y[:, :3], yr[:, :3])] + [rel(gw[i], ws[i].grad) for i in range(3)] + \
           [rel(gb[i], bs[i].grad) for i in range(3)]
    print("exact-fp64 rel errors", errs, flush=True)
    assert max(errs) < TOL
    assert y[:, 3:].abs().max().item() == 0.0


def test_deterministic(gpu):
    from fem_taichi_amd.learned import MeshConvNet, ConvWork
    g = ConvWork(12, 10, 8)
    net = MeshConvNet(seed=3)
    x = torch.zeros(g.P, 32)
    x[:, :4] = torch.randn(g.P, 4, generator=torch.Generator().manual_seed(4))
    x0 = x.to(torch.bfloat16).cuda()
    y1 = net.forward(x0, g, keep=True)
    dy = torch.randn(g.P, 16, generator=torch.Generator().manual_seed(5)).cuda()
    g1 = net.backward(dy)
    y2 = net.forward(x0, g, keep=True)
    g2 = net.backward(dy)
    assert torch.equal(y1, y2)
    for a, b in zip(g1[0] + g1[1], g2[0] + g2[1]):
        assert torch.equal(a, b)
