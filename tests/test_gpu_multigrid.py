"""GPU parity of the multigrid preconditioner pieces and of the multigrid
static solve.

* Grid transfers: fem_mg_restrict / fem_mg_prolong_add against P^T and P of
  the trilinear interpolation matrix built independently in scipy (FP32,
  1e-6 relative).
* Smoother: the three modes of fem_mg_smooth against numpy on the oracle's
  penalised K (FP32 matrix and vectors, 1e-5 relative).
* Solve: the FP64 residual of the result on the oracle's K and F below 1e-8,
  in far fewer iterations than block-Jacobi PCG.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem_oracle as O
from fem_taichi_amd import mesh

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dims", [(7, 6, 5), (26, 6, 4), (5, 3, 9)])
def test_transfers_match_interpolation_matrix(gpu, dims):
    import torch
    from fem_taichi_amd import _lib, multigrid as MG
    from fem_taichi_amd.runtime import stream_ptr
    print(f"\ndims {dims}", flush=True)
    sels = [MG.axis_select(n) for n in dims]
    T = MG.Transfer(dims, sels)
    P = sp.kron(MG.interpolation_matrix(dims, sels), sp.eye(3)).tocsr()
    rng = np.random.default_rng(1)
    xf = rng.standard_normal(P.shape[0]).astype(np.float32)
    xc = rng.standard_normal(P.shape[1]).astype(np.float32)
    bc = torch.empty(P.shape[1], dtype=torch.float32, device="cuda")
    _lib.call("fem_mg_restrict", T.ref, torch.as_tensor(xf).cuda().data_ptr(), bc.data_ptr(), 1, stream_ptr())
    assert rel(bc.cpu().numpy(), P.T @ xf.astype(np.float64)) < 1e-6
    y = torch.as_tensor(xf).cuda()
    _lib.call("fem_mg_prolong_add", T.ref, torch.as_tensor(xc).cuda().data_ptr(), y.data_ptr(), 1, stream_ptr())
    assert rel(y.cpu().numpy(), xf + P @ xc.astype(np.float64)) < 1e-6
    # several column-interleaved vectors at once (5: one column per thread;
    # 8 and 16: four per thread, each column's sums in the same order, so
    # bit-identical to single-vector calls on the columns)
    for nb in (5, 8, 16):
        XF = rng.standard_normal((P.shape[0], nb)).astype(np.float32)
        XC = rng.standard_normal((P.shape[1], nb)).astype(np.float32)
        BC = torch.empty(P.shape[1], nb, dtype=torch.float32, device="cuda")
        _lib.call("fem_mg_restrict", T.ref, torch.as_tensor(XF).cuda().data_ptr(), BC.data_ptr(), nb, stream_ptr())
        assert rel(BC.cpu().numpy(), P.T @ XF.astype(np.float64)) < 1e-6
        Y = torch.as_tensor(XF).cuda()
        _lib.call("fem_mg_prolong_add", T.ref, torch.as_tensor(XC).cuda().data_ptr(), Y.data_ptr(), nb, stream_ptr())
        assert rel(Y.cpu().numpy(), XF + P @ XC.astype(np.float64)) < 1e-6
        for c in range(nb):
            b1 = torch.empty(P.shape[1], dtype=torch.float32, device="cuda")
            _lib.call("fem_mg_restrict", T.ref, torch.as_tensor(np.ascontiguousarray(XF[:, c])).cuda().data_ptr(),
                      b1.data_ptr(), 1, stream_ptr())
            assert torch.equal(b1, BC[:, c]), (nb, c)
            y1 = torch.as_tensor(np.ascontiguousarray(XF[:, c])).cuda()
            _lib.call("fem_mg_prolong_add", T.ref, torch.as_tensor(np.ascontiguousarray(XC[:, c])).cuda().data_ptr(),
                      y1.data_ptr(), 1, stream_ptr())
            assert torch.equal(y1, Y[:, c]), (nb, c)


@pytest.mark.parametrize("layout", ["bsr", "csr"])
def test_smoother_modes(gpu, layout):
    import torch
    from fem_taichi_amd import Model, Physics, _lib, solvers
    from fem_taichi_amd.runtime import stream_ptr
    g = mesh.structured_block(6, 5, 7, 1.0, 0.10, 1, E_sigma=0.3)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    k32 = model.assemble_K("f32", layout)
    lay = _lib.LAYOUT_BSR if layout == "bsr" else _lib.LAYOUT_CSR
    dinv, _ = solvers.BsrOperator(model.dev, model.K_bsr("f32")).block_jacobi()
    n = K.shape[0]
    rng = np.random.default_rng(2)
    x = rng.standard_normal(n).astype(np.float32)
    b = (K @ rng.standard_normal(n)).astype(np.float32)
    D = dinv.cpu().numpy().astype(np.float64).reshape(-1, 3, 3)
    Dinv = sp.block_diag(list(D), format="csr")
    w = 0.7
    xd, bd = torch.as_tensor(x).cuda(), torch.as_tensor(b).cuda()
    out = torch.empty_like(xd)
    st = stream_ptr()
    r_ref = b - K @ x.astype(np.float64)
    _lib.call("fem_mg_smooth", model.dev.ref, k32.data_ptr(), lay, dinv.data_ptr(), w, xd.data_ptr(), bd.data_ptr(),
              out.data_ptr(), 0, st)
    free = np.abs(r_ref) < 1e15        # penalised rows: 1e20 x in FP32 and FP64 differ in the last bits
    assert rel(out.cpu().numpy()[free], r_ref[free]) < 1e-5
    _lib.call("fem_mg_smooth", model.dev.ref, k32.data_ptr(), lay, dinv.data_ptr(), w, xd.data_ptr(), bd.data_ptr(),
              out.data_ptr(), 1, st)
    ref = x + w * (Dinv @ r_ref)
    assert rel(out.cpu().numpy(), ref) < 1e-5
    _lib.call("fem_mg_smooth", model.dev.ref, None, lay, dinv.data_ptr(), w, None, bd.data_ptr(), out.data_ptr(), 2,
              st)
    assert rel(out.cpu().numpy(), w * (Dinv @ b.astype(np.float64))) < 1e-5


@pytest.mark.parametrize("nb", [3, 8, 16, 18])
def test_smoother_multi_matches_single(gpu, nb):
    """fem_mg_smooth_multi on nb columns equals fem_mg_smooth per column."""
    import torch
    from fem_taichi_amd import Model, Physics, _lib, solvers
    from fem_taichi_amd.runtime import stream_ptr
    g = mesh.structured_block(6, 5, 7, 1.0, 0.10, 1, E_sigma=0.3)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    k32 = model.K_bsr("f32")
    dinv, _ = solvers.BsrOperator(model.dev, k32).block_jacobi()
    n = 3 * phys.Num_Node
    rng = np.random.default_rng(3)
    X = torch.as_tensor(rng.standard_normal((n, nb)).astype(np.float32)).cuda()
    B = torch.as_tensor(rng.standard_normal((n, nb)).astype(np.float32)).cuda() * 1e3
    st = stream_ptr()
    for mode in (0, 1, 2):
        OUT = torch.empty_like(X)
        _lib.call("fem_mg_smooth_multi", model.dev.ref, k32.data_ptr(), dinv.data_ptr(), 0.6, X.data_ptr(),
                  B.data_ptr(), OUT.data_ptr(), nb, mode, st)
        for c in range(nb):
            x, b = X[:, c].contiguous(), B[:, c].contiguous()
            o = torch.empty_like(x)
            _lib.call("fem_mg_smooth", model.dev.ref, k32.data_ptr(), _lib.LAYOUT_BSR, dinv.data_ptr(), 0.6,
                      x.data_ptr(), b.data_ptr(), o.data_ptr(), mode, st)
            a, r = OUT[:, c].cpu().numpy().astype(np.float64), o.cpu().numpy().astype(np.float64)
            fin = np.abs(r) < 1e15          # penalised residual rows: 1e20 x, summed in other orders
            assert rel(a[fin], r[fin]) < 1e-5, (mode, c)
        if nb == 16:
            # the 16-column kernel (K staged in LDS) sums each column like the
            # one-column-per-thread kernel: bit-identical to two 8-column calls
            for h in range(2):
                Xh, Bh = X[:, 8 * h:8 * h + 8].contiguous(), B[:, 8 * h:8 * h + 8].contiguous()
                Oh = torch.empty_like(Xh)
                _lib.call("fem_mg_smooth_multi", model.dev.ref, k32.data_ptr(), dinv.data_ptr(), 0.6, Xh.data_ptr(),
                          Bh.data_ptr(), Oh.data_ptr(), 8, mode, st)
                assert torch.equal(Oh, OUT[:, 8 * h:8 * h + 8]), (mode, h)


@pytest.mark.parametrize("nb", [8, 16])
def test_block_fcg_matches_scipy(gpu, nb):
    """The block flexible CG (8 or 16 right-hand sides, one of them zero)
    against scipy's direct solves on the oracle's K: every nonzero column to
    1e-9.  With 16 columns on the blade's node grid the iteration takes p.q
    from the structured-grid product (fem_bsr_spmm16_grid_pq)."""
    import torch
    import scipy.sparse.linalg as sla
    from fem_taichi_amd import Model, Physics, solvers, multigrid as MG
    g = mesh.twisted_blade(40, 8, 8)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    pre = MG.MultigridPreconditioner(model, MG.hierarchy_for(model))
    pp = O.pre_process(g.Node, g.Elem)
    K = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    n = K.shape[0]
    assert model.dev.grid_nodes is not None
    Bh = np.random.default_rng(9).standard_normal((n, nb))
    Bh[:, 3] = 0.0
    X, info = MG.solve_block_fcg_mg(model, K64, torch.as_tensor(Bh).cuda(), pre, tol=1e-9)
    Xh = X.cpu().numpy()
    for c in range(nb):
        if c == 3:
            assert np.all(Xh[:, c] == 0)
            continue
        assert info["converged"][c] and info["rel_residual"][c] < 1e-9
        r = np.linalg.norm(Bh[:, c] - K @ Xh[:, c]) / np.linalg.norm(Bh[:, c])
        assert r < 1e-9
    ref = sla.spsolve(K.tocsc(), Bh[:, 0])
    assert rel(Xh[:, 0], ref) < 1e-5


def mg_cases():
    return {
        "blade_40x8x8": mesh.twisted_blade(40, 8, 8),
        "slab_20x20x18_lognormal": mesh.structured_block(20, 20, 18, 1.0, 0.10, 1, E_sigma=0.3,
                                                         name="slab_20x20x18"),
        "cantilever_12x10x30_jit": mesh.structured_block(12, 10, 30, 1.0, 0.05, 0, name="cant_12x10x30"),
    }


@pytest.mark.parametrize("case", list(mg_cases().keys()))
@pytest.mark.parametrize("min_nodes", [1000, 60])
def test_mg_solve(gpu, case, min_nodes):
    import torch
    from fem_taichi_amd import Model, Physics, solvers, multigrid as MG
    print(f"\ncase {case} min_nodes {min_nodes}", flush=True)
    g = mg_cases()[case]
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    K64 = solvers.BsrOperator(model.dev, model.K_bsr("f64"))
    F = model.F_device()
    x, info = MG.solve_mg(model, K64, F, tol=1e-8, min_nodes=min_nodes, fallback=False)
    det = info.history[0]
    assert det["converged"], det["trace"][-5:]
    assert len(det["levels"]) >= 2
    pp = O.pre_process(g.Node, g.Elem)
    Kref = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    Fref = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    assert O.static_residual(Kref, x.cpu().numpy(), Fref) < 1e-8
    _, it_pcg, _ = solvers.pcg(K64, F, K64.block_jacobi()[0], 1e-8, 100000)
    print(f"levels {det['levels']}: MG-FCG {info.iterations} it, PCG {it_pcg} it", flush=True)
    assert info.iterations * 5 < it_pcg
    torch.cuda.synchronize()


def test_mg_static_solver_method(gpu):
    from fem_taichi_amd import Model, Physics, Static_solver
    g = mesh.twisted_blade(40, 8, 8)
    phys = Physics.from_grid(g)
    model = Model(g.name, phys, g.Materials, Omegas=g.Omegas)
    U = Static_solver(model, 0).Solve_cpu(method="mg", tol=1e-9)
    pp = O.pre_process(g.Node, g.Elem)
    Kref = O.model_K_csr(g.Node, g.Elem, pp, g.EX, g.PO, g.Bounds)
    Fref = O.model_F(g.Node, g.Elem, pp, g.DE, g.Omegas, g.Forces)
    assert O.static_residual(Kref, U, Fref) < 1e-9


@pytest.mark.parametrize("n", [1, 5, 700, 2049])
@pytest.mark.parametrize("nb", [1, 3, 8, 13, 16, 64])
def test_coarsest_dense_products(gpu, n, nb):
    """fem_dense_gemm_nb_f32 (the coarsest level's inverse on nb columns, in
    4-row x 16-column tiles, k split over four slices) and fem_dense_gemv_f32
    against numpy in FP64."""
    import torch
    from fem_taichi_amd import _lib
    from fem_taichi_amd.runtime import stream_ptr
    rng = np.random.default_rng(n * 100 + nb)
    A = rng.standard_normal((n, n)).astype(np.float32)
    X = rng.standard_normal((n, nb)).astype(np.float32)
    Ad, Xd = torch.as_tensor(A).cuda(), torch.as_tensor(X).cuda()
    Y = torch.full((n, nb), np.nan, dtype=torch.float32, device="cuda")
    _lib.call("fem_dense_gemm_nb_f32", Ad.data_ptr(), n, Xd.data_ptr(), Y.data_ptr(), nb, stream_ptr())
    ref = A.astype(np.float64) @ X.astype(np.float64)
    assert rel(Y.cpu().numpy(), ref) < 1e-5
    y = torch.full((n,), np.nan, dtype=torch.float32, device="cuda")
    x = Xd[:, 0].contiguous()
    _lib.call("fem_dense_gemv_f32", Ad.data_ptr(), n, x.data_ptr(), y.data_ptr(), stream_ptr())
    assert rel(y.cpu().numpy(), ref[:, 0]) < 1e-5


@pytest.mark.parametrize("n,nb", [(7, 1), (1000, 1), (77777, 1), (30000, 8), (5001, 10), (20000, 16), (300, 64)])
def test_fcg_steps_match_numpy(gpu, n, nb):
    """fem_fcg_direction / fem_fcg_update (through FcgSteps) against the FCG(1)
    formulas in numpy FP64: two iterations (the first without beta), a masked
    row set, one frozen column; FP32 and FP64 Z; R32 = float(R); results
    reproducible bit for bit."""
    import torch
    from fem_taichi_amd import multigrid as MG
    print(f"\nfcg steps n {n} nb {nb}", flush=True)
    rng = np.random.default_rng(n + nb)
    mask = (rng.random(n) > 0.1).astype(np.float64)
    active = np.ones(nb, dtype=np.int32)
    active[nb // 2] = 0 if nb > 1 else 1
    X0, R0 = rng.standard_normal((n, nb)), rng.standard_normal((n, nb)) * mask[:, None]
    Z1, Z2 = rng.standard_normal((n, nb)).astype(np.float32), rng.standard_normal((n, nb))
    Q1, Q2 = rng.standard_normal((n, nb)), rng.standard_normal((n, nb))

    def ref():
        X, R = X0.copy(), R0.copy()
        P = mask[:, None] * Z1.astype(np.float64)
        pq = (P * Q1).sum(0)
        alpha = np.where(active > 0, (P * R).sum(0) / pq, 0.0)
        X += alpha * P
        R -= alpha * mask[:, None] * Q1
        zm = mask[:, None] * Z2
        beta = (zm * Q1).sum(0) / pq
        P = zm - beta * P
        pq2 = (P * Q2).sum(0)
        alpha2 = np.where(active > 0, (P * R).sum(0) / pq2, 0.0)
        X += alpha2 * P
        R -= alpha2 * mask[:, None] * Q2
        return P, X, R, (R * R).sum(0)

    def run():
        c = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
        st = MG.FcgSteps(n, nb, "cuda")
        X, R, P = c(X0), c(R0), torch.empty(n, nb, dtype=torch.float64, device="cuda")
        R32 = torch.empty(n, nb, dtype=torch.float32, device="cuda")
        m, act, q1, q2 = c(mask), c(active), c(Q1), c(Q2)
        st.direction(c(Z1), m, None, P, True)
        st.update(P, q1, m, act, X, R, R32)
        st.direction(c(Z2), m, q1, P, False)
        rr = st.update(P, q2, m, act, X, R, R32).cpu().numpy()
        return P.cpu().numpy(), X.cpu().numpy(), R.cpu().numpy(), rr, R32.cpu().numpy()

    Pr, Xr, Rr, rrr = ref()
    P, X, R, rr, R32 = run()
    # random P and Q make some p.q small, so alpha (a ratio of sums taken in
    # another order) carries rounding into R: the formulas are checked to
    # 1e-10 and the column sums of squares against the kernel's own R
    for a, b in ((P, Pr), (X, Xr), (R, Rr)):
        assert rel(a, b) < 1e-10
    assert np.allclose(rr, (R * R).sum(0), rtol=1e-12, atol=0)
    assert np.array_equal(R32, R.astype(np.float32))
    if nb > 1:
        assert np.array_equal(X[:, nb // 2], X0[:, nb // 2])
    again = run()
    for a, b in zip(again, (P, X, R, rr, R32)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,nb,nblk", [(7, 1, 1), (77777, 1, 13), (30000, 8, 300), (20000, 16, 2016), (300, 64, 5)])
def test_fcg_split_steps_match_numpy(gpu, n, nb, nblk):
    """fem_fcg_direction_pr / fem_fcg_update_split (p.q from nblk outside
    partials, p.r from the direction pass) against the same two FCG(1)
    iterations in numpy as test_fcg_steps_match_numpy, and against the
    unsplit steps; the tolerance is test_fcg_steps_match_numpy's (random P
    and Q make some p.q small, and the partials here sum in another order)."""
    import torch
    from fem_taichi_amd import multigrid as MG
    print(f"\nfcg split steps n {n} nb {nb} partials {nblk}", flush=True)
    rng = np.random.default_rng(5 * n + nb)
    mask = (rng.random(n) > 0.1).astype(np.float64)
    active = np.ones(nb, dtype=np.int32)
    active[nb // 2] = 0 if nb > 1 else 1
    X0, R0 = rng.standard_normal((n, nb)), rng.standard_normal((n, nb)) * mask[:, None]
    Z1, Z2 = rng.standard_normal((n, nb)).astype(np.float32), rng.standard_normal((n, nb))
    Q1, Q2 = rng.standard_normal((n, nb)), rng.standard_normal((n, nb))
    c = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()

    def partials(P, Q):
        # nblk row ranges of unequal length (some empty when nblk > n)
        cuts = np.sort(rng.integers(0, n + 1, nblk - 1))
        edges = np.concatenate([[0], cuts, [n]])
        return np.stack([(P[a:b] * Q[a:b]).sum(0) for a, b in zip(edges[:-1], edges[1:])])

    def run(split):
        st = MG.FcgSteps(n, nb, "cuda")
        X, R, P = c(X0), c(R0), torch.empty(n, nb, dtype=torch.float64, device="cuda")
        R32 = torch.empty(n, nb, dtype=torch.float32, device="cuda")
        m, act, q1, q2 = c(mask), c(active), c(Q1), c(Q2)
        out = []
        for Z, Q, Qp, first in ((Z1, q1, None, True), (Z2, q2, q1, False)):
            if split:
                st.direction_pr(c(Z), m, Qp, R, P, first)
                pq = c(partials(P.cpu().numpy(), Q.cpu().numpy()))
                rr = st.update_split(P, Q, m, act, pq, nblk, X, R, R32)
            else:
                st.direction(c(Z), m, Qp, P, first)
                rr = st.update(P, Q, m, act, X, R, R32)
        return P.cpu().numpy(), X.cpu().numpy(), R.cpu().numpy(), rr.cpu().numpy(), R32.cpu().numpy()

    P, X, R, rr, R32 = run(True)
    P0, X0u, R0u, rr0, _ = run(False)
    for a, b in ((P, P0), (X, X0u), (R, R0u)):
        assert rel(a, b) < 1e-10
    assert np.allclose(rr, (R * R).sum(0), rtol=1e-12, atol=0)
    assert np.array_equal(R32, R.astype(np.float32))
    if nb > 1:
        assert np.array_equal(X[:, nb // 2], X0[:, nb // 2])
    # the formulas themselves (as in test_fcg_steps_match_numpy)
    Xr, Rr = X0.copy(), R0.copy()
    Pr = mask[:, None] * Z1.astype(np.float64)
    pq = (Pr * Q1).sum(0)
    alpha = np.where(active > 0, (Pr * Rr).sum(0) / pq, 0.0)
    Xr += alpha * Pr
    Rr -= alpha * mask[:, None] * Q1
    zm = mask[:, None] * Z2
    Pr = zm - ((zm * Q1).sum(0) / pq) * Pr
    alpha2 = np.where(active > 0, (Pr * Rr).sum(0) / (Pr * Q2).sum(0), 0.0)
    Xr += alpha2 * Pr
    Rr -= alpha2 * mask[:, None] * Q2
    for a, b in ((P, Pr), (X, Xr), (R, Rr)):
        assert rel(a, b) < 1e-10


def test_fcg_split_rejects_bad_partials(gpu):
    import torch
    from fem_taichi_amd import _lib, multigrid as MG
    st = MG.FcgSteps(10, 2, "cuda")
    V = torch.zeros(10, 2, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        st.update_split(V, V, None, None, torch.zeros(3, dtype=torch.float64, device="cuda"), 2, V, V.clone(), None)
    with pytest.raises(_lib.FemError):
        _lib.call("fem_fcg_update_split", 10, 2, V.data_ptr(), V.data_ptr(), None, None, V.data_ptr(), 0,
                  V.data_ptr(), V.data_ptr(), None, st.rr.data_ptr(), st.work.data_ptr(), st.wb, None)


@pytest.mark.parametrize("n,nb", [(5, 1), (77777, 1), (30000, 8), (20000, 16), (300, 64)])
def test_fcg_residual_matches_numpy(gpu, n, nb):
    """fem_fcg_residual (through FcgSteps.residual): t = (B - Y) mask, its
    column norms and those of B against numpy; R and R32 written on the
    selected columns only (the others keep their contents), bit-exact there
    (one subtraction and one product, as numpy does them); Y, mask, sel and
    R/R32 each absent; the carried FCG scalars in the shared work buffer are
    left alone."""
    import torch
    from fem_taichi_amd import multigrid as MG
    print(f"\nfcg residual n {n} nb {nb}", flush=True)
    rng = np.random.default_rng(3 * n + nb)
    B, Y = rng.standard_normal((n, nb)), rng.standard_normal((n, nb))
    mask = (rng.random(n) > 0.2).astype(np.float64)
    sel = (rng.random(nb) > 0.5).astype(np.int32)
    sel[0] = 1
    c = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    st = MG.FcgSteps(n, nb, "cuda")
    st.work[-4096:] = 7                                  # the carried scalars' region: must survive
    keep = st.work[-4096:].clone()
    R = torch.full((n, nb), -1.0, dtype=torch.float64, device="cuda")
    R32 = torch.full((n, nb), -1.0, dtype=torch.float32, device="cuda")
    tn, bn = st.residual(c(B), c(Y), c(mask), c(sel), R, R32)
    t = (B - Y) * mask[:, None]
    assert np.allclose(tn, np.linalg.norm(t, axis=0), rtol=1e-13, atol=0)
    assert np.allclose(bn, np.linalg.norm(B, axis=0), rtol=1e-13, atol=0)
    Rh, R32h = R.cpu().numpy(), R32.cpu().numpy()
    on = sel > 0
    assert np.array_equal(Rh[:, on], t[:, on]) and np.array_equal(R32h[:, on], t[:, on].astype(np.float32))
    assert np.all(Rh[:, ~on] == -1.0) and np.all(R32h[:, ~on] == -1.0)
    assert torch.equal(st.work[-4096:], keep)
    tn2, _ = st.residual(c(B), None, None, None, R, None)          # R = B
    assert np.allclose(tn2, np.linalg.norm(B, axis=0), rtol=1e-13, atol=0)
    assert np.array_equal(R.cpu().numpy(), B)
    tn3, bn3 = st.residual(c(B), c(Y), None, None, None, None)     # norms only
    assert np.allclose(tn3, np.linalg.norm(B - Y, axis=0), rtol=1e-13, atol=0)
    again = st.residual(c(B), c(Y), c(mask), c(sel), R, R32)
    assert np.array_equal(again[0], tn) and np.array_equal(again[1], bn)
