"""Learned preconditioner: training with the BF16 kernels lowers the
preconditioned residual below block-Jacobi's on a held-out slab, and the
learned solve reaches the FP64 residual target (falling back to
block-Jacobi PCG when the flexible iteration stalls)."""
import numpy as np
import pytest
import torch

from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(gpu):
    from fem_taichi_amd import learned as L
    print("\ncase learned: 3 slabs 14x14x12, 150 steps", flush=True)
    slabs = [L.TrainingSlab(mesh.training_slab(1000 + i, 14, 14, 12), rhs_seeds=(i,)) for i in range(3)]
    net = L.MeshConvNet(seed=0)
    hist = L.train(net, slabs, steps=150, lr=2e-3)
    return L, net, hist


def test_training_beats_block_jacobi(trained):
    L, net, hist = trained
    held = L.TrainingSlab(mesh.training_slab(2000, 14, 14, 12), rhs_seeds=(7,))
    bj = L.bj_loss(held)
    learned = float(np.mean([float(held.loss_and_grad(net, r)[0]) for r in held.residuals]))
    print(f"held-out preconditioned residual: block-Jacobi {bj:.3f}, learned {learned:.3f}", flush=True)
    assert np.mean(hist[-20:]) < np.mean(hist[:20])
    assert learned < 0.8 * bj


def test_learned_solve_reaches_tolerance(trained):
    from fem_taichi_amd import Model, Physics, solvers
    L, net, _ = trained
    g = mesh.training_slab(3000, 14, 14, 12)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    F = model.F_device()
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    x, info = L.solve_learned(model, K64, F, tol=1e-8, net=net)
    r = F - K64.matvec(x)
    assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(F)) < 1e-8
    assert info.rel_residual < 1e-8
