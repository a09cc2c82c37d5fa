"""Train the learned preconditioner's mesh-convolution network.

    python -m fem_taichi_amd.train_learned --slabs 16 --dims 60x60x55 --steps 600 \\
        --out fem_taichi_amd/weights/meshconv_slab.pt

Slabs are ``mesh.training_slab(seed)`` (jitter 10 %, lognormal E with sigma
0.5) for seeds 1000, 1001, ...; each contributes the residual snapshots of
block-Jacobi PCG runs on its own load and on one random right-hand side.
The loss is the preconditioned residual ||r - K z(r)||^2 / ||r||^2; the
script prints it next to the block-Jacobi value on a held-out slab.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import learned as L
from . import mesh


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--slabs", type=int, default=16)
    ap.add_argument("--dims", default="60x60x55")
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=L.default_weights_path())
    a = ap.parse_args()
    dims = tuple(int(v) for v in a.dims.split("x"))
    t0 = time.perf_counter()
    slabs = []
    for i in range(a.slabs):
        slabs.append(L.TrainingSlab(mesh.training_slab(1000 + i, *dims), rhs_seeds=(i,)))
        print(f"slab {i}: {len(slabs[-1].residuals)} residuals, {time.perf_counter() - t0:.1f} s", flush=True)
    net = L.MeshConvNet(seed=a.seed)
    hist = L.train(net, slabs, steps=a.steps, lr=a.lr, seed=a.seed, log=lambda s: print(s, flush=True))
    held = L.TrainingSlab(mesh.training_slab(5000, *dims), rhs_seeds=(99,))
    bj = L.bj_loss(held)
    learned = float(np.mean([float(held.loss_and_grad(net, r)[0]) for r in held.residuals]))
    torch.save(net.state_dict(), a.out)
    summary = {"slabs": a.slabs, "dims": dims, "steps": a.steps, "final_train_loss": float(np.mean(hist[-50:])),
               "heldout_bj_loss": bj, "heldout_learned_loss": learned, "seconds": time.perf_counter() - t0}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
