"""Train the learned preconditioner's mesh-convolution network.

    python -m fem_taichi_amd.train_learned --role mg --slabs 16 --dims 60x60x55 --steps 600

Slabs are ``mesh.training_slab(seed)`` (jitter 10 %, lognormal E with sigma
0.5) for seeds 1000, 1001, ...; each contributes the residual snapshots of
solver runs on its own load and on one random right-hand side: block-Jacobi
PCG residuals (role "bj") or the residuals the multigrid V-cycle leaves
inside multigrid flexible CG (role "mg").  The loss is the preconditioned
residual ||r - K z(r)||^2 / ||r||^2 (role "bj") or the energy norm of the
error left after the correction relative to the error's (role "mg"); the
script prints it next to the value without the network on a held-out slab.
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
    ap.add_argument("--role", choices=("bj", "mg"), default="mg")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    out = a.out or L.default_weights_path(a.role)
    dims = tuple(int(v) for v in a.dims.split("x"))
    t0 = time.perf_counter()
    slabs = []
    for i in range(a.slabs):
        slabs.append(L.TrainingSlab(mesh.training_slab(1000 + i, *dims), rhs_seeds=(i,), role=a.role))
        print(f"slab {i}: {len(slabs[-1].residuals)} residuals, {time.perf_counter() - t0:.1f} s", flush=True)
    net = L.MeshConvNet(seed=a.seed)
    hist = L.train(net, slabs, steps=a.steps, lr=a.lr, seed=a.seed, log=lambda s: print(s, flush=True))
    held = L.TrainingSlab(mesh.training_slab(5000, *dims), rhs_seeds=(99,), role=a.role)
    bj = L.bj_loss(held)
    learned = held.mean_loss(net)
    torch.save(net.state_dict(), out)
    summary = {"role": a.role, "slabs": a.slabs, "dims": dims, "steps": a.steps, "final_train_loss": float(np.mean(hist[-50:])),
               "heldout_bj_loss": bj, "heldout_learned_loss": learned, "seconds": time.perf_counter() - t0}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
