"""Host control logic of the flexible (learned) solve: convergence, stagnation
and divergence detection, FP64 fallback.  CPU only: the operator, the
preconditioner and the fallback PCG are CPU stand-ins with the same
interfaces as the device ones."""
import numpy as np
import pytest
import torch

from fem_taichi_amd import learned as L
from fem_taichi_amd import solvers


class DenseOp:
    def __init__(self, A):
        self.A = torch.as_tensor(A, dtype=torch.float64)

    def matvec(self, x, y=None):
        out = self.A @ x
        if y is not None:
            y.copy_(out)
            return y
        return out

    def block_jacobi(self):
        return torch.diag(self.A).reciprocal(), None


def cpu_pcg(op, b, dinv, tol, maxiter, x0=None, **kw):
    x = x0.clone() if x0 is not None else torch.zeros_like(b)
    r = b - op.matvec(x)
    z = dinv * r
    p = z.clone()
    rz = float(r @ z)
    nb = float(torch.linalg.vector_norm(b))
    it = 0
    while it < maxiter and float(torch.linalg.vector_norm(r)) / nb >= tol:
        q = op.matvec(p)
        a = rz / float(p @ q)
        x += a * p
        r -= a * q
        z = dinv * r
        rz2 = float(r @ z)
        p = z + (rz2 / rz) * p
        rz = rz2
        it += 1
    return x, it, float(torch.linalg.vector_norm(r)) / nb


@pytest.fixture
def system(monkeypatch):
    monkeypatch.setattr(L, "stream_ptr", lambda: 0)
    monkeypatch.setattr(solvers, "pcg", cpu_pcg)
    rng = np.random.default_rng(0)
    n = 120
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.logspace(0, 4, n)) @ Q.T
    b = torch.as_tensor(rng.standard_normal(n))
    return DenseOp(A), b, A


def rel(A, x, b):
    return float(np.linalg.norm(b.numpy() - A @ x.numpy()) / np.linalg.norm(b.numpy()))


def test_good_preconditioner_converges_without_fallback(system):
    op, b, A = system
    Ainv = torch.as_tensor(np.linalg.inv(A))
    x, info = L.solve_fcg(op, None, b, lambda r: Ainv @ r, tol=1e-10, check_every=1)
    assert not info["fell_back"] and info["iterations"] <= 3
    assert rel(A, x, b) < 1e-10


def test_useless_preconditioner_falls_back_and_converges(system):
    op, b, A = system
    g = torch.Generator().manual_seed(1)
    x, info = L.solve_fcg(op, None, b, lambda r: torch.randn(r.shape, generator=g, dtype=r.dtype),
                          tol=1e-8, maxiter=5000)
    assert info["fell_back"]
    assert info["iterations"] < 5000 + 1000      # the stall is caught long before maxiter
    assert rel(A, x, b) < 1e-8


def test_non_finite_preconditioner_falls_back(system):
    op, b, A = system
    x, info = L.solve_fcg(op, None, b, lambda r: torch.full_like(r, float("nan")), tol=1e-8)
    assert info["fell_back"] and np.isfinite(x.numpy()).all()
    assert rel(A, x, b) < 1e-8
