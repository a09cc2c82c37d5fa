"""Host-side tables of the multigrid hierarchy (CPU only): the per-axis
transfer tables against a brute-force linear interpolation in the index
coordinate, the restriction spans, and the coarse element moduli against
explicit block means."""
import numpy as np
import pytest

from fem_taichi_amd import multigrid as MG


def brute_interp(n, sel):
    P = np.zeros((n, len(sel)))
    for c in range(len(sel) - 1):
        a, b = sel[c], sel[c + 1]
        for f in range(a, b + 1):
            t = (f - a) / (b - a)
            P[f, c] = 1.0 - t if f != b else P[f, c]
            P[f, c + 1] = t if f != a else P[f, c + 1]
    for c, f in enumerate(sel):
        P[f, :] = 0.0
        P[f, c] = 1.0
    return P


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 26, 51, 401])
def test_axis_tables(n):
    sel = MG.axis_select(n)
    assert sel[0] == 0 and sel[-1] == n - 1 and np.all(np.diff(sel) >= 1) and np.all(np.diff(sel) <= 2)
    left, wr, span = MG.axis_tables(n, sel)
    P = np.zeros((n, len(sel)))
    for f in range(n):
        P[f, left[f]] += 1.0 - wr[f]
        if wr[f]:
            assert left[f] + 1 < len(sel)
            P[f, left[f] + 1] += wr[f]
    assert np.allclose(P, brute_interp(n, sel))
    for c in range(len(sel)):
        nz = np.nonzero(P[:, c])[0]
        lo, hi = span[2 * c], span[2 * c + 1]
        assert 0 <= lo <= hi <= n - 1
        assert lo <= nz.min() and nz.max() <= hi


def test_interpolation_matrix_rows_sum_to_one():
    dims = (9, 4, 6)
    sels = [MG.axis_select(n) for n in dims]
    P = MG.interpolation_matrix(dims, sels)
    assert P.shape == (9 * 4 * 6, 5 * 3 * 4)
    assert np.allclose(np.asarray(P.sum(axis=1)).ravel(), 1.0)
    # coinciding nodes map to exactly one coarse node
    cx, cy, cz = np.meshgrid(sels[0], sels[1], sels[2], indexing="ij")
    f = (cx + dims[0] * (cy + dims[1] * cz)).transpose(2, 1, 0).ravel()
    assert np.allclose(P[f].toarray(), np.eye(P.shape[1]))


def test_coarse_modulus_block_means():
    rng = np.random.default_rng(0)
    dims = (8, 5, 6)
    E = rng.random(7 * 4 * 5)
    sels = [MG.axis_select(n) for n in dims]
    Ec = MG.coarse_modulus(E, dims, sels)
    e3 = E.reshape(5, 4, 7)
    ref = [e3[sels[2][k]:sels[2][k + 1], sels[1][j]:sels[1][j + 1], sels[0][i]:sels[0][i + 1]].mean()
           for k in range(len(sels[2]) - 1) for j in range(len(sels[1]) - 1) for i in range(len(sels[0]) - 1)]
    assert np.allclose(Ec, ref)
    assert MG.coarse_modulus(2.0e5, dims, sels) == 2.0e5


def test_bounds_from_fixed_roundtrip():
    fixed = np.zeros((10, 3), dtype=bool)
    fixed[[0, 3], :] = True
    fixed[5, 1] = True
    fixed[7, [0, 2]] = True
    got = np.zeros_like(fixed)
    for ids, flags in MG._bounds_from_fixed(fixed):
        for j in range(3):
            if flags[j] != -1:
                got[ids, j] = True
    assert np.array_equal(got, fixed)
