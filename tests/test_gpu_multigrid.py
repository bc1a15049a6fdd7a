"""GPU multigrid against the oracle and the reference's own residual thresholds
(test/Multigrid.jl:54-90, 92-102, 151-170)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def csr_equal(a, b, exact=True, tol=1e-12):
    (rp, ci, vv), (orp, oci, ovv) = a, b
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    if exact:
        assert np.array_equal(vv, ovv)
    else:
        assert np.max(np.abs(vv - ovv)) <= tol * np.max(np.abs(ovv))


@pytest.fixture(scope="module")
def grid_pair(O, cs):
    base_o = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, compress=False)
    H = O.Hierarchy(base_o, 4, geometric=True)
    base_g = cs.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, False)
    md = cs.MultigridData(base_g, cs.BinarySubdivision, 4, cs.GeometricHodge, 3)
    return H, md


@pytest.fixture(scope="module")
def sphere_pair(O, cs):
    H = O.Hierarchy(O.icosahedron(), 4, project=True, geometric=False)
    md = cs.MultigridData(cs.icosphere(0), cs.BinarySubdivision, 4, cs.DiagonalHodge, 2, project=True)
    return H, md


@pytest.mark.parametrize("which", ["grid", "sphere"])
def test_hierarchy_matches_oracle(which, grid_pair, sphere_pair):
    H, md = grid_pair if which == "grid" else sphere_pair
    assert md.nlevels == len(H.A) and md.sizes == [a.m for a in H.A]
    for lv in range(md.nlevels):
        # operators depend on circumcenters (closed form here, elimination in
        # the oracle): equal to rounding; R and P are exact
        csr_equal(md.export(lv, "A"), H.A[lv].csr(), exact=False)
        if lv + 1 < md.nlevels:
            csr_equal(md.export(lv, "R"), H.R[lv].csr())
            csr_equal(md.export(lv, "P"), H.S[lv].transpose().csr())


def rhs(O, L, seed):
    return L.matvec(0.5 * O.uniform_form(seed, L.n) + 0.5)   # b = L * rand(nv)


def test_reference_thresholds_cg(O, cs, grid_pair):
    import torch

    H, md = grid_pair
    L = H.A[0]
    b = rhs(O, L, 7)
    db = torch.tensor(b, device="cuda")
    res = lambda u: np.linalg.norm(L.matvec(u.cpu().numpy()) - b) / np.linalg.norm(b)  # noqa: E731
    z = lambda: torch.zeros(L.n, dtype=torch.float64, device="cuda")  # noqa: E731
    assert res(cs.dec_Delta_inv(md)(db)) < 1e-6
    assert res(cs.multigrid_vcycles(z(), db, md, 5)) < 1e-7
    assert res(cs.multigrid_wcycles(z(), db, md, 5)) < 1e-7
    assert res(cs.full_multigrid(db, md, 5)) < 1e-6
    assert res(cs.full_multigrid(db, md, 5, cs.cg, 2)) < 1e-7
    r0 = res(cs.multigrid_vcycles(z(), db, md, 5, steps=0))
    r1 = res(cs.multigrid_vcycles(z(), db, md, 5, steps=1))
    assert r1 < r0
    # CG iterates are third-party arithmetic (Krylov.jl): compared with the
    # oracle's restatement only to a tolerance
    u_o = H.mu_cycles(np.zeros(L.n), b, 2, alg=O.CG)
    u_g = cs.multigrid_vcycles(z(), db, md, 2, steps=3).cpu().numpy()
    assert np.linalg.norm(u_g - u_o) <= 1e-8 * np.linalg.norm(u_o)


def test_galerkin_mode(O, cs):
    """GalerkinMode (test/Multigrid.jl:113-148, BinarySubdivision): coarse
    operators R A P equal the oracle's (both accumulate each entry over k
    ascending), and the reference's residual thresholds hold on the GPU."""
    import torch

    base_o = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, compress=False)
    H = O.Hierarchy(base_o, 4, geometric=True, galerkin=True)
    base_g = cs.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, False)
    md = cs.MultigridData(base_g, cs.BinarySubdivision, 4, cs.GeometricHodge, 3, mode=cs.GalerkinMode)
    assert md.sizes == [a.m for a in H.A]
    for lv in range(md.nlevels):
        csr_equal(md.export(lv, "A"), H.A[lv].csr(), exact=False)
    L = H.A[0]
    b = rhs(O, L, 9)
    db = torch.tensor(b, device="cuda")
    res = lambda u: np.linalg.norm(L.matvec(u.cpu().numpy()) - b) / np.linalg.norm(b)  # noqa: E731
    z = lambda: torch.zeros(L.n, dtype=torch.float64, device="cuda")  # noqa: E731
    assert res(cs.multigrid_vcycles(z(), db, md, 5)) < 2e-6
    assert res(cs.multigrid_wcycles(z(), db, md, 5)) < 7e-7
    assert res(cs.full_multigrid(db, md, 5)) < 1e-3
    assert res(cs.full_multigrid(db, md, 5, cs.cg, 2)) < 5e-6


@pytest.mark.parametrize("alg", ["gs", "jacobi"])
def test_smoothers_match_oracle(O, cs, sphere_pair, alg):
    import torch

    H, md = sphere_pair
    L = H.A[0]
    b = rhs(O, L, 9)
    b = b - b.mean()
    oalg, galg = (O.GS, cs.gauss_seidel) if alg == "gs" else (O.JACOBI, cs.jacobi)
    # the oracle sweeps the colours the library chose (checked proper below)
    cols = [md.colors(lv)[0] for lv in range(md.nlevels)]
    u_o = H.mu_cycles(np.zeros(L.n), b, 3, alg=oalg, steps=2, coarse_direct=True, colors=cols)
    u = torch.zeros(L.n, dtype=torch.float64, device="cuda")
    cs.multigrid_vcycles(u, torch.tensor(b, device="cuda"), md, 3, galg, steps=2, coarse_direct=True)
    assert np.linalg.norm(u.cpu().numpy() - u_o) <= 1e-10 * np.linalg.norm(u_o)


@pytest.mark.parametrize("steps", [1, 2, 3])
def test_cg_smoother_forms_match_oracle(O, cs, sphere5_pair, steps):
    """The Krylov CG smoother (Multigrid.jl:860) on levels 10,242 .. 12: the
    finest level runs the multi-workgroup form (two launches per iteration),
    the levels of at most 3,072 rows the one-workgroup form.  The dot products
    add in another order than the oracle's, so u agrees to a tolerance.  A
    zero right-hand side stops CG at once and leaves u zero.

    Steps are kept at 3 or fewer: the 12-row level has three distinct nonzero
    eigenvalues, so CG converges there in 3 steps, and further steps on its
    singular, slightly inconsistent system divide by a p·Ap that rounding
    sets, which moves u along the constants by an amount no tolerance covers.
    This happens in the oracle too: u's mean is 16 at 5 steps, 8e3 at 8 and
    1e8 at 40."""
    import torch

    H, md = sphere5_pair
    L = H.A[0]
    b = rhs(O, L, 11)
    b = b - b.mean()
    u_o = H.mu_cycles(np.zeros(L.n), b, 2, alg=O.CG, steps=steps)
    u = torch.zeros(L.n, dtype=torch.float64, device="cuda")
    cs.multigrid_vcycles(u, torch.tensor(b, device="cuda"), md, 2, cs.cg, steps=steps)
    assert np.linalg.norm(u.cpu().numpy() - u_o) <= 1e-8 * np.linalg.norm(u_o)
    z = torch.zeros(L.n, dtype=torch.float64, device="cuda")
    cs.multigrid_vcycles(z, torch.zeros_like(z), md, 1, cs.cg, steps=steps)
    assert torch.count_nonzero(z).item() == 0


@pytest.mark.parametrize("which", ["grid", "sphere"])
def test_gauss_seidel_colouring(which, grid_pair, sphere_pair):
    """Every level's colouring is proper for the level's operator pattern and
    uses at most four colours (four on refined levels); on each refined level the coarse vertices (the first
    n_{l+1}) form colour 0 and the midpoint of coarse edge (a, b) has colour
    c(a) xor c(b)."""
    H, md = grid_pair if which == "grid" else sphere_pair
    for lv in range(md.nlevels):
        c, k = md.colors(lv)
        rp, ci, _ = md.export(lv, "A")
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        off = rows != ci
        assert k <= 4 and c.min() == 0 and c.max() == k - 1 and (k == 4 or lv == md.nlevels - 1)
        assert not np.any(c[rows[off]] == c[ci[off]])
        if lv + 1 < md.nlevels:
            cc, _ = md.colors(lv + 1)
            nc = md.sizes[lv + 1]
            assert not c[:nc].any()
            m = H.meshes[lv + 1]
            assert np.array_equal(c[nc:], cc[m.v0] ^ cc[m.v1])


def sphere_rhs(O, d, seed):
    """Seeded smooth right-hand side, projected to ⋆0-weighted mean zero."""
    p = d.m.p
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, 4)
    f = c[0] * p[:, 2] + c[1] * p[:, 0] * p[:, 1] + c[2] * (3 * p[:, 2] ** 2 - 1) + c[3] * p[:, 0] ** 3
    h0 = d.hodge_diag(0)
    return f - (h0 @ f) / h0.sum()


@pytest.mark.parametrize("variant", ["plain", "mixed"])
def test_solve_to_tolerance(O, cs, sphere_pair, variant):
    import torch

    H, md = sphere_pair
    L = H.A[0]
    f = sphere_rhs(O, H.duals[0], 3)
    u, cycles, rr = cs.solve(md, torch.tensor(f, device="cuda"), tol=1e-8, variant=variant)
    u = u.cpu().numpy()
    true = np.linalg.norm(f - L.matvec(u)) / np.linalg.norm(f)   # float64 oracle check
    assert rr <= 1e-8 and true <= 1.01e-8
    assert cycles <= 40


@pytest.mark.parametrize("variant", ["plain", "mixed"])
def test_fused_and_graph_paths_are_bit_identical(O, cs, variant, monkeypatch):
    """The single-workgroup coarse tail and the hipGraph replay of each cycle
    give the same iterates as one launch per phase, eagerly issued."""
    import torch

    def run():
        md = cs.MultigridData(cs.icosphere(0), cs.BinarySubdivision, 5, cs.DiagonalHodge, 2, project=True)
        d = O.Dual(O.icosphere(5))
        f = torch.tensor(sphere_rhs(O, d, 11), device="cuda")
        out = []
        for _ in range(3):   # replays of the cached graph, eager work in between
            u, c, r = cs.solve(md, f, tol=1e-8, variant=variant)
            torch.cuda.synchronize()
            assert torch.isfinite(torch.randn(4096, device="cuda").sum())
            out.append((u.clone(), c, r))
        return out

    fast = run()
    monkeypatch.setenv("DEC_MG_NO_FUSE", "1")
    monkeypatch.setenv("DEC_MG_NO_GRAPH", "1")
    slow = run()
    for (u1, c1, r1), (u2, c2, r2) in zip(fast, slow):
        assert c1 == c2 and r1 == r2 and torch.equal(u1, u2)


@pytest.fixture(scope="module")
def sphere5_pair(O, cs):
    """Levels 10,242 .. 12: the finest level (above 3,000 rows) is split into
    blocks and renumbered; the others are one block each."""
    H = O.Hierarchy(O.icosahedron(), 5, project=True, geometric=False)
    md = cs.MultigridData(cs.icosphere(0), cs.BinarySubdivision, 5, cs.DiagonalHodge, 2, project=True)
    return H, md


def test_blocks_partition_levels(sphere5_pair):
    """Blocks cover every vertex, hold at most 512 vertices, and each block of
    the finest level is one connected patch; small levels are one block."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    H, md = sphere5_pair
    for lv in range(md.nlevels):
        blk, nb = md.blocks(lv)
        n = md.sizes[lv]
        assert blk.min() == 0 and blk.max() == nb - 1 and len(np.unique(blk)) == nb
        if n <= 3000:
            assert nb == 1
            continue
        sizes = np.bincount(blk)
        assert sizes.max() <= 512 and nb >= n // 512
        rp, ci, _ = md.export(lv, "A")
        rows = np.repeat(np.arange(n), np.diff(rp))
        same = blk[rows] == blk[ci]
        G = sp.csr_matrix((np.ones(same.sum()), (rows[same], ci[same])), shape=(n, n))
        ncomp, lab = connected_components(G, directed=False)
        assert ncomp <= nb + 8          # BFS-grown: almost every block is one patch
        assert same.mean() > 0.6        # most couplings stay inside a block


def test_block_gs_matches_oracle(O, cs, sphere5_pair):
    """Hybrid block Gauss-Seidel V(2,2) cycles against the oracle's sweep with
    the library's colours and blocks (the direct coarse solve sums in another
    order, so to 1e-10)."""
    import torch

    H, md = sphere5_pair
    L = H.A[0]
    f = sphere_rhs(O, H.duals[0], 5)
    cols = [md.colors(lv)[0] for lv in range(md.nlevels)]
    blks = [md.blocks(lv)[0] for lv in range(md.nlevels)]
    u_o = H.mu_cycles(np.zeros(L.n), f, 3, alg=O.BLOCK_GS, steps=2, coarse_direct=True, colors=cols, blocks=blks)
    u = torch.zeros(L.n, dtype=torch.float64, device="cuda")
    cs.multigrid_vcycles(u, torch.tensor(f, device="cuda"), md, 3, cs.block_gauss_seidel, steps=2,
                         coarse_direct=True)
    assert np.linalg.norm(u.cpu().numpy() - u_o) <= 1e-10 * np.linalg.norm(u_o)
    # odd sweep count (the last sweep copied back) and W-cycles: same check
    u_o = H.mu_cycles(np.zeros(L.n), f, 2, alg=O.BLOCK_GS, mu=2, steps=3, coarse_direct=True, colors=cols,
                      blocks=blks)
    u.zero_()
    cs.multigrid_wcycles(u, torch.tensor(f, device="cuda"), md, 2, cs.block_gauss_seidel, steps=3,
                         coarse_direct=True)
    assert np.linalg.norm(u.cpu().numpy() - u_o) <= 1e-10 * np.linalg.norm(u_o)


@pytest.mark.parametrize("variant", ["plain", "mixed"])
def test_block_gs_solve(O, cs, sphere5_pair, variant):
    """The solve with the hybrid smoother reaches 1e-8 (checked by the oracle's
    float64 operator) in no more cycles than multicolour Gauss-Seidel plus one."""
    import torch

    H, md = sphere5_pair
    L = H.A[0]
    f = sphere_rhs(O, H.duals[0], 3)
    df = torch.tensor(f, device="cuda")
    u, cb, rr = cs.solve(md, df, tol=1e-8, variant=variant, alg=cs.block_gauss_seidel)
    true = np.linalg.norm(f - L.matvec(u.cpu().numpy())) / np.linalg.norm(f)
    assert rr <= 1e-8 and true <= 1.01e-8
    _, cg_, _ = cs.solve(md, df, tol=1e-8, variant=variant, alg=cs.gauss_seidel)
    assert cb <= cg_ + 1


def test_pcg_iterates_match_oracle(O, cs, sphere5_pair):
    """Three iterations of CG preconditioned by one V(2,2) block Gauss-Seidel
    cycle against the oracle's PCG with the oracle's V-cycle (the library's
    colours and blocks) as M: to 1e-9 (the cycles themselves agree to 1e-10)."""
    import torch

    H, md = sphere5_pair
    L = H.A[0]
    f = sphere_rhs(O, H.duals[0], 11)
    cols = [md.colors(lv)[0] for lv in range(md.nlevels)]
    blks = [md.blocks(lv)[0] for lv in range(md.nlevels)]

    def vcycle(r):
        return H.mu_cycles(np.zeros(L.n), r, 1, alg=O.BLOCK_GS, steps=2, coarse_direct=True, colors=cols,
                           blocks=blks)

    x_o, _ = O.pcg(L, f, np.zeros(L.n), 3, vcycle)
    u = torch.zeros(L.n, dtype=torch.float64, device="cuda")
    with pytest.raises(cs.DecError):   # 1e-30 is out of reach: exactly 3 iterations run
        cs.solve(md, torch.tensor(f, device="cuda"), u=u, tol=1e-30, max_cycles=3, variant="pcg")
    assert np.linalg.norm(u.cpu().numpy() - x_o) <= 1e-9 * np.linalg.norm(x_o)


@pytest.mark.parametrize("variant", ["pcg", "pcg_mixed"])
def test_pcg_solve(O, cs, sphere5_pair, variant):
    """CG preconditioned by one V-cycle (float64, or float32 for pcg_mixed)
    reaches 1e-8, checked by the oracle's float64 operator, in fewer
    iterations than the stationary cycles of the same precision."""
    import torch

    H, md = sphere5_pair
    L = H.A[0]
    f = sphere_rhs(O, H.duals[0], 3)
    df = torch.tensor(f, device="cuda")
    u, it, rr = cs.solve(md, df, tol=1e-8, variant=variant)
    true = np.linalg.norm(f - L.matvec(u.cpu().numpy())) / np.linalg.norm(f)
    assert rr <= 1e-8 and true <= 1.01e-8 and abs(true - rr) <= 1e-3 * rr
    _, cyc, _ = cs.solve(md, df, tol=1e-8, variant="plain" if variant == "pcg" else "mixed")
    assert it < cyc, (it, cyc)


@pytest.mark.parametrize("pair", ["sphere_pair", "sphere5_pair"])
@pytest.mark.parametrize("variant", ["plain", "mixed", "pcg", "pcg_mixed"])
def test_solve_zero_rhs_and_converged_start(O, cs, request, pair, variant):
    """b = 0 gives u = 0 after no cycle, and a start that already meets the
    tolerance comes back unchanged after no cycle.  On the renumbered
    hierarchy (sphere5_pair) the library queues its first cycles before it
    reads ||b|| and the starting residual, so these exits discard work
    already queued; on sphere_pair (vertex order) it reads first."""
    import torch

    H, md = request.getfixturevalue(pair)
    n = H.A[0].n
    u = torch.tensor(O.uniform_form(5, n), device="cuda")
    u, c, rr = cs.solve(md, torch.zeros(n, dtype=torch.float64, device="cuda"), u=u, tol=1e-8, variant=variant)
    assert c == 0 and rr == 0.0 and not u.any()
    f = torch.tensor(sphere_rhs(O, H.duals[0], 3), device="cuda")
    x, _, _ = cs.solve(md, f, tol=1e-10, variant=variant)
    x2, c2, rr2 = cs.solve(md, f, u=x.clone(), tol=1e-8, variant=variant)
    assert c2 == 0 and rr2 <= 1e-8 and torch.equal(x2, x)


def test_hierarchy_from_matrices(O, cs, sphere5_pair):
    """MultigridData(ops, rs, ps, steps) (Multigrid.jl:545-546): the oracle's own
    CSR matrices handed to the library give the oracle's cycles (CG smoother
    of the reference, and Gauss-Seidel with the library's greedy colours),
    and the same CG iterates as the hierarchy the library builds itself."""
    import torch

    H, md = sphere5_pair
    csr = lambda M: M.csr()  # noqa: E731
    mdm = cs.MultigridData.from_matrices([csr(a) for a in H.A], [csr(r) for r in H.R],
                                         [csr(s.transpose()) for s in H.S], steps=3)
    assert mdm.sizes == md.sizes
    for lv in range(mdm.nlevels):
        csr_equal(mdm.export(lv, "A"), H.A[lv].csr())
        if lv + 1 < mdm.nlevels:
            csr_equal(mdm.export(lv, "R"), H.R[lv].csr())
            csr_equal(mdm.export(lv, "P"), H.S[lv].transpose().csr())
    L = H.A[0]
    f = sphere_rhs(O, H.duals[0], 8)
    df = torch.tensor(f, device="cuda")
    u_o = H.mu_cycles(np.zeros(L.n), f, 2, alg=O.CG, steps=3)
    u = cs.multigrid_vcycles(torch.zeros(L.n, dtype=torch.float64, device="cuda"), df, mdm, 2, steps=3)
    assert np.linalg.norm(u.cpu().numpy() - u_o) <= 1e-8 * np.linalg.norm(u_o)
    cols = [mdm.colors(lv)[0] for lv in range(mdm.nlevels)]
    u_o = H.mu_cycles(np.zeros(L.n), f, 3, alg=O.GS, steps=2, coarse_direct=True, colors=cols)
    u = cs.multigrid_vcycles(torch.zeros(L.n, dtype=torch.float64, device="cuda"), df, mdm, 3, cs.gauss_seidel,
                             steps=2, coarse_direct=True)
    assert np.linalg.norm(u.cpu().numpy() - u_o) <= 1e-10 * np.linalg.norm(u_o)
    # a solve to 1e-8 on the matrix-built hierarchy, verified by the oracle
    u, c, rr = cs.solve(mdm, df, tol=1e-8)
    assert rr <= 1e-8 and np.linalg.norm(f - L.matvec(u.cpu().numpy())) / np.linalg.norm(f) <= 1.01e-8


def test_from_matrices_rejects_bad_input(cs):
    import scipy.sparse as sp

    from cspaces._lib import DecError

    A = sp.identity(4, format="csr") * 2.0
    R = sp.csr_matrix(np.ones((2, 4)) * 0.25)
    P = sp.csr_matrix(np.ones((4, 3)))          # wrong shape: 4 x 3, coarse level has 2
    with pytest.raises(DecError):
        cs.MultigridData.from_matrices([A, sp.identity(2, format="csr")], [R], [P])
    Z = sp.csr_matrix((4, 4))                   # no diagonal
    with pytest.raises(DecError):
        cs.MultigridData.from_matrices([Z], [], [])
