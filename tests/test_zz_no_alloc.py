"""fem_gemm_tn_f64_noalloc (no workspace, static device scratch) against
fem_gemm_tn_f64 with a caller workspace, through the C ABI: bit-identical C
on two calls in a row, the second smaller than the first (so the second call
reads a scratch that still holds the first call's partials), and FEM_E_LIMIT
for p * q beyond the static scratch.  Also checked against torch's FP64
product (1e-12 relative)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_pair(lib, st, n, p, q, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(n, p, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(n, q, dtype=torch.float64, device="cuda", generator=g)
    C_ref = torch.full((p, q), float("nan"), dtype=torch.float64, device="cuda")
    C_na = torch.full((p, q), float("nan"), dtype=torch.float64, device="cuda")
    wb = int(lib.fem_gemm_tn_work_bytes(p, q))
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    from fem_taichi_amd import _lib
    _lib.call("fem_gemm_tn_f64", A.data_ptr(), p, B.data_ptr(), q, n, p, q, C_ref.data_ptr(), work.data_ptr(), wb,
              st)
    _lib.call("fem_gemm_tn_f64_noalloc", A.data_ptr(), p, B.data_ptr(), q, n, p, q, C_na.data_ptr(), st)
    torch.cuda.synchronize()
    exact = (A.T @ B)
    return C_ref.cpu().numpy(), C_na.cpu().numpy(), exact.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(gpu):
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    lib = _lib.lib()
    st = stream_ptr()
    # large first (many row chunks, p*q at the scratch limit), then smaller
    for n, p, q, seed in ((300_001, 64, 32, 1), (5_003, 24, 8, 2), (17, 3, 5, 3)):
        print(f"\ncase gemm_tn noalloc n={n} p={p} q={q}", flush=True)
        ref, na, exact = _run_pair(lib, st, n, p, q, seed)
        assert np.array_equal(ref, na), np.abs(ref - na).max()
        assert np.linalg.norm(na - exact) <= 1e-12 * np.linalg.norm(exact)


def test_noalloc_rejects_oversize(gpu):
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    import torch
    lib = _lib.lib()
    A = torch.zeros(8, 64, dtype=torch.float64, device="cuda")
    C = torch.zeros(64, 64, dtype=torch.float64, device="cuda")
    rc = lib.fem_gemm_tn_f64_noalloc(A.data_ptr(), 64, A.data_ptr(), 64, 8, 64, 64, C.data_ptr(), stream_ptr())
    assert rc == _lib.FEM_E_LIMIT
    assert "static scratch" in _lib.last_error()
