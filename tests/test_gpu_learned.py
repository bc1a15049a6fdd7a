"""Learned smoother: GPU forward / backward against the float64 oracle
(oracle/learned.py, whose gradient is pinned by finite differences), training
on (residual, error) pairs from plain V-cycles, and the Gauss-Seidel fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(cs):
    import learned as LO

    md = cs.MultigridData(cs.icosphere(0), cs.BinarySubdivision, 3, cs.DiagonalHodge, 2, project=True)
    nn = cs.LearnedSmoother(md, seed=3)
    n, nc = md.sizes[0], md.sizes[1]
    lv = LO.Level(*md.export(0, "A"), R=LO.Csr(*md.export(0, "R"), ncols=n), Pr=LO.Csr(*md.export(0, "P"), ncols=nc))
    return md, nn, lv


def random_params(nn, seed):
    rng = np.random.default_rng(seed)
    p = nn.get_params().astype(np.float64)
    p += 0.1 * rng.standard_normal(p.size)
    return p.astype(np.float32)


def residuals(O, n, nsamp, seed):
    return np.stack([O.uniform_form(seed + i, n) for i in range(nsamp)]).astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b))


def test_parameter_layout(small):
    import learned as LO

    md, nn, lv = small
    assert nn.nparams == LO.n_params()
    p = LO.unpack(nn.get_params().astype(np.float64))
    assert p["alpha"][0] == pytest.approx(2 / 3) and not p["wo"].any() and not p["bo"].any()


@pytest.mark.parametrize("precise", [True, False])
def test_forward_matches_oracle(O, small, precise):
    import torch

    import learned as LO

    md, nn, lv = small
    p = random_params(nn, 1)
    nn.set_params(p)
    R = residuals(O, lv.n, 3, 10)
    c = nn.forward(0, torch.tensor(R, device="cuda"), precise=precise).cpu().numpy()
    for b in range(3):
        ref, _ = LO.forward(p.astype(np.float64), lv, R[b].astype(np.float64))
        err = rel(c[b], ref)
        # FP32 path: rounding of float32 arithmetic; BF16 path: 8-bit mantissa
        # activations and weights in the two 64->32 layers (error stated here)
        assert err <= (1e-5 if precise else 3e-2), err


@pytest.mark.parametrize("kind", ["residual", "error"])
@pytest.mark.parametrize("precise", [True, False])
def test_gradient_matches_oracle(O, small, precise, kind):
    import torch

    import learned as LO

    md, nn, lv = small
    p = random_params(nn, 2)
    nn.set_params(p)
    R = residuals(O, lv.n, 4, 20)
    if kind == "residual":
        loss, g = nn.loss_grad(0, torch.tensor(R, device="cuda"), precise=precise)
        lref, gref = LO.loss_grad(p.astype(np.float64), lv, R.astype(np.float64))
    else:
        E = residuals(O, lv.n, 4, 40)
        loss, g = nn.loss_grad(0, torch.tensor(R, device="cuda"), torch.tensor(E, device="cuda"), precise=precise)
        lref, gref = LO.loss_err_grad(p.astype(np.float64), lv, R.astype(np.float64), E.astype(np.float64))
    # FP32 path: float32 rounding; BF16 path: 8-bit mantissa activations and
    # weights in the two dense layers (errors stated here)
    assert loss == pytest.approx(lref, rel=1e-5 if precise else 3e-2)
    assert rel(g, gref) <= (1e-4 if precise else 0.1), rel(g, gref)


@pytest.mark.parametrize("base,levels", [(0, 3), (2, 5)])
def test_inference_kernels_match_forward(O, cs, base, levels):
    """The V-cycle's inference kernels (layers 1 and 3 in registers) add to u
    exactly the c of the BF16 forward, on every level: the 16-row kernels up
    to 50000 rows, the 64-row kernels above (163842 rows); 642 and 40962 rows
    are not multiples of the 16-row tile, 163842 not of 64."""
    import torch

    md = cs.MultigridData(cs.icosphere(base), cs.BinarySubdivision, levels, cs.DiagonalHodge, 2, project=True)
    nn = cs.LearnedSmoother(md, seed=8)
    nn.set_params(random_params(nn, 4))
    for lv in range(md.nlevels):
        n = md.sizes[lv]
        r = torch.tensor(O.uniform_form(30 + lv, n), dtype=torch.float32, device="cuda")
        u0 = torch.tensor(O.uniform_form(60 + lv, n), dtype=torch.float32, device="cuda")
        c = nn.forward(lv, r.clone())
        u = nn.smooth(lv, r, u0.clone())
        assert torch.equal(u, u0 + c), (lv, n, (u - (u0 + c)).abs().max().item())


def test_inference_with_wide_halos(O, cs):
    """A level above 50,000 rows whose 64-row tiles name more than 191 other
    rows runs layer 1 without the staged halo (the direct gathers).  The
    graph is a ring with five random matchings, so the breadth-first blocks
    cannot keep a row's neighbours together.  The vertex-order call still adds
    exactly the forward's c."""
    import scipy.sparse as sp
    import torch

    rng = np.random.default_rng(3)
    n, agg = 60000, 60
    rows = [np.arange(n), np.arange(n)]
    cols = [(np.arange(n) + 1) % n, (np.arange(n) - 1) % n]
    for _ in range(5):
        p = rng.permutation(n)
        rows += [p[0::2], p[1::2]]
        cols += [p[1::2], p[0::2]]
    W = sp.csr_matrix((np.ones(sum(r.size for r in rows)), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    W.data[:] = 1.0  # duplicate edges count once
    W.setdiag(0)
    W.eliminate_zeros()
    deg = np.asarray(W.sum(axis=1)).ravel()
    A = (sp.diags(deg + 0.5) - W).tocsr()
    A.sort_indices()
    R = sp.csr_matrix((np.full(n, 1.0 / agg), (np.arange(n) // agg, np.arange(n))), shape=(n // agg, n))
    P = sp.csr_matrix((np.ones(n), (np.arange(n), np.arange(n) // agg)), shape=(n, n // agg))
    # the coarse operator only has to be a valid level: a ring of n / agg rows
    m = n // agg
    Ac = (sp.diags(np.full(m, 2.5)) - sp.diags(np.ones(m - 1), 1) - sp.diags(np.ones(m - 1), -1)).tocsr()
    Ac.sort_indices()
    csr = lambda M: (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64))  # noqa: E731
    md = cs.MultigridData.from_matrices([csr(A), csr(Ac)], [csr(R)], [csr(P)], steps=2)
    blk, nb = md.blocks(0)
    # the full 64-row blocks (contiguous in the level's numbering) name far
    # more than 191 rows outside them, so the tiles over them do too
    size = np.bincount(blk, minlength=nb)
    B = sp.csr_matrix((np.ones(n), (blk, np.arange(n))), shape=(nb, n))
    reach = (B @ (A != 0)).tocsr()
    halo = np.diff(reach.indptr) - size
    full = size == 64
    assert full.sum() >= 100 and halo[full].min() > 250, (full.sum(), halo[full].min())
    nn = cs.LearnedSmoother(md, seed=8)
    nn.set_params(random_params(nn, 4))
    r = torch.tensor(O.uniform_form(30, n), dtype=torch.float32, device="cuda")
    u0 = torch.tensor(O.uniform_form(60, n), dtype=torch.float32, device="cuda")
    c = nn.forward(0, r.clone())
    u = nn.smooth(0, r, u0.clone())
    assert torch.equal(u, u0 + c), (u - (u0 + c)).abs().max().item()


@pytest.mark.parametrize("kind", ["residual", "error"])
def test_training_gain_and_solve(O, cs, kind):
    """Training lowers the loss on held-out V-cycle pairs compared with the
    untrained network (damped Jacobi), and the learned float32 V-cycle solves
    to 1e-8 (float64-verified)."""
    import torch

    md = cs.MultigridData(cs.icosphere(1), cs.BinarySubdivision, 3, cs.DiagonalHodge, 2, project=True)  # k=4
    R, E = cs.make_pairs(md, 480, cycles_per_rhs=6, seed=4)
    train, val = R[:400].contiguous(), R[400:].contiguous()
    etrain, evl = (E[:400].contiguous(), E[400:].contiguous()) if kind == "error" else (None, None)
    nn = cs.LearnedSmoother(md, seed=5)
    l0, _ = nn.loss_grad(0, val, evl)
    hist = nn.train(0, train, steps=400, E=etrain, batch=16, lr=2e-3, seed=6)
    l1, _ = nn.loss_grad(0, val, evl)
    assert np.all(np.isfinite(hist)) and l1 < 0.9 * l0, (l0, l1)
    # solve with the network on the finest level
    nn.attach(True, min_rows=0)
    d = O.Dual(O.icosphere(4))
    h0 = d.hodge_diag(0)
    f = O.uniform_form(9, d.nv)
    f -= (h0 @ f) / h0.sum()
    u, cycles, rr = cs.solve(md, torch.tensor(f, device="cuda"), tol=1e-8, variant="mixed")
    L = d.laplacian0()
    true = np.linalg.norm(f - L.matvec(u.cpu().numpy())) / np.linalg.norm(f)
    assert rr <= 1e-8 and true <= 1.01e-8
    st = cs.learned_stats(md)
    assert st["learned_cycles"] >= 1


def test_fallback_to_gauss_seidel(O, cs):
    """A network that does nothing (alpha = 0, zero head) cannot reduce the
    residual by 0.5x: the solve switches to Gauss-Seidel and still converges."""
    import torch

    import learned as LO

    md = cs.MultigridData(cs.icosphere(1), cs.BinarySubdivision, 3, cs.DiagonalHodge, 2, project=True)
    nn = cs.LearnedSmoother(md, seed=7)
    p = LO.unpack(nn.get_params().astype(np.float64))
    p["alpha"][0] = 0.0
    nn.set_params(LO.pack(p).astype(np.float32))
    nn.attach(True, min_rows=0)
    d = O.Dual(O.icosphere(4))
    h0 = d.hodge_diag(0)
    f = O.uniform_form(13, d.nv)
    f -= (h0 @ f) / h0.sum()
    u, cycles, rr = cs.solve(md, torch.tensor(f, device="cuda"), tol=1e-8, variant="mixed")
    assert rr <= 1e-8
    st = cs.learned_stats(md)
    assert st["fallbacks"] == 1
