"""The oracle's preconditioned conjugate gradients (the library's CG solve
variants; no reference counterpart) pinned by its limiting cases: with the
identity as preconditioner it is the oracle's restatement of Krylov.jl's cg,
iterate for iterate; with the exact inverse it solves in one iteration; with
one block Gauss-Seidel V(2,2) cycle it beats the stationary cycles."""
import numpy as np


def _system(O, levels):
    H = O.Hierarchy(O.icosahedron(), levels, project=True)
    A = H.A[0]
    d = H.duals[0]
    rng = np.random.default_rng(5)
    f = rng.uniform(-1, 1, A.m)
    f -= (d.hodge_diag(0) @ f) / d.hodge_diag(0).sum()   # ⋆0-weighted mean zero: consistent
    return H, A, f


def test_identity_preconditioner_is_krylov_cg(O):
    H, A, f = _system(O, 2)
    for k in (1, 2, 5):
        x, _ = O.pcg(A, f, np.zeros(A.m), k, lambda r: r.copy())
        xc = O.cg(A, f, np.zeros(A.m), k)
        assert np.linalg.norm(x - xc) <= 1e-12 * np.linalg.norm(xc), k


def test_exact_preconditioner_solves_in_one_iteration(O):
    H, A, f = _system(O, 1)
    Ad = A.dense()
    Ap = np.linalg.pinv(Ad)
    x, r = O.pcg(A, f, np.zeros(A.m), 1, lambda r: Ap @ r)
    assert np.linalg.norm(f - Ad @ x) <= 1e-10 * np.linalg.norm(f)


def test_vcycle_preconditioner_beats_stationary_cycles(O):
    """Levels 0->4 (2562 rows): one V(2,2) block Gauss-Seidel cycle as M reaches
    1e-8 in fewer iterations than the stationary cycles need."""
    H, A, f = _system(O, 4)
    cols = [O.greedy_colors(a)[1] for a in H.A]
    blks = [np.arange(a.m, dtype=np.int32) // 64 if a.m > 3000 else np.zeros(a.m, np.int32) for a in H.A]

    def vcycle(r):
        return H.mu_cycles(np.zeros(A.m), r, 1, alg=O.BLOCK_GS, steps=2, coarse_direct=True, colors=cols,
                           blocks=blks)

    nf = np.linalg.norm(f)
    u = np.zeros(A.m)
    stat = None
    for c in range(1, 30):
        u = vcycle(f - A.matvec(u)) + u
        if np.linalg.norm(f - A.matvec(u)) <= 1e-8 * nf:
            stat = c
            break
    assert stat is not None
    it = None
    for k in range(1, stat + 1):
        x, r = O.pcg(A, f, np.zeros(A.m), k, vcycle)
        if np.linalg.norm(f - A.matvec(x)) <= 1e-8 * nf:
            it = k
            break
    assert it is not None and it < stat, (it, stat)
