"""Learned preconditioner, both roles of the network.

Block-Jacobi role: training with the BF16 kernels lowers the preconditioned
residual below block-Jacobi's on a held-out slab, and the learned flexible
PCG reaches the FP64 residual target (falling back to block-Jacobi PCG when
the flexible iteration stalls).

Multigrid role: training lowers the error energy the V-cycle leaves, the
learned multigrid solve reaches the target, and the ablation against the
same solve without the network gives the network's gain in iterations."""
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
    learned = held.mean_loss(net)
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
    x, info = L.solve_learned(model, K64, F, tol=1e-8, net=net, base="bj")
    r = F - K64.matvec(x)
    assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(F)) < 1e-8
    assert info.rel_residual < 1e-8


@pytest.fixture(scope="module")
def trained_mg(gpu):
    from fem_taichi_amd import learned as L
    print("\ncase learned (multigrid role): 3 slabs 20x20x18, 150 steps", flush=True)
    slabs = [L.TrainingSlab(mesh.training_slab(1000 + i, 20, 20, 18), rhs_seeds=(i,), role="mg") for i in range(3)]
    net = L.MeshConvNet(seed=0)
    hist = L.train(net, slabs, steps=150, lr=2e-3)
    return L, net, hist


def test_mg_role_training_lowers_error_energy(trained_mg):
    """After the V-cycle, the network's correction (with the energy-optimal
    step) removes a share of the error energy on a held-out slab: the loss
    1 - (c.r)^2 / (c.K c e.r) falls below the 1 of no correction."""
    L, net, hist = trained_mg
    held = L.TrainingSlab(mesh.training_slab(2000, 20, 20, 18), rhs_seeds=(7,), role="mg")
    learned = held.mean_loss(net)
    print(f"held-out error energy left after the correction: {learned:.3f} (1 without the network)", flush=True)
    assert np.mean(hist[-20:]) < np.mean(hist[:20])
    assert 0.0 <= learned < 0.97


def test_learned_mg_ablation(trained_mg):
    """The learned form against the same multigrid flexible CG without the
    network: both reach the FP64 residual target; the iteration counts are
    the network's gain (small: the V-cycle already removes what a 7^3-node
    receptive field can see)."""
    from fem_taichi_amd import Model, Physics, solvers
    L, net, _ = trained_mg
    g = mesh.training_slab(3000, 20, 20, 18)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    F = model.F_device()
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    runs = {}
    for use_net in (True, False):
        x, info = L.solve_learned(model, K64, F, tol=1e-8, net=net, base="mg", use_net=use_net, min_gain=None)
        r = F - K64.matvec(x)
        assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(F)) < 1e-8
        assert not info.history[0]["fell_back"]
        runs[use_net] = info.iterations
    print(f"multigrid FCG iterations: with the network {runs[True]}, without {runs[False]}", flush=True)
    assert runs[True] <= runs[False] + 2


def test_shipped_mg_weights_solve_slab(gpu):
    """The shipped multigrid-role weights inside the learned solve of the
    60x60x55 slab (lognormal E): FP64 residual below 1e-8 without fallback."""
    from fem_taichi_amd import Model, Physics, solvers
    from fem_taichi_amd import learned as L
    g = mesh.slab_60x60x55()
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    F = model.F_device()
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    x, info = L.solve_learned(model, K64, F, tol=1e-8)
    assert info.method == "learned_mg" and not info.history[0]["fell_back"]
    r = F - K64.matvec(x)
    assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(F)) < 1e-8
    assert info.iterations < 40


def test_learned_mg_switches_network_off(gpu):
    """The learned form drops the network after gain_patience applications in
    a row whose energy gain is below min_gain (1.0 here, so always): exactly
    two network applications, then the V-cycle alone, to the FP64 target."""
    from fem_taichi_amd import Model, Physics, solvers
    from fem_taichi_amd import learned as L
    g = mesh.twisted_blade(40, 8, 8)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    F = model.F_device()
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    x, info = L.solve_learned(model, K64, F, tol=1e-8, base="mg", min_gain=1.0, gain_patience=2)
    h = info.history[0]
    assert h["net_applications"] == 2 and len(h["gains"]) == 2 and not h["fell_back"]
    assert float(torch.linalg.vector_norm(F - K64.matvec(x)) / torch.linalg.vector_norm(F)) < 1e-8
