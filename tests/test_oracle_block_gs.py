"""The oracle's hybrid block Gauss-Seidel sweep (the GPU smoother
`block_gauss_seidel`; no counterpart in the reference) pinned by its limiting
cases and by its convergence rate: one block is the multicolour sweep, one
vertex per block is undamped Jacobi, and inside V(2,2) cycles with blocks of
64 vertices it converges about as fast as the multicolour sweep."""
from collections import deque

import numpy as np


def _colors(O, H):
    return [O.greedy_colors(a)[1] for a in H.A]


def bfs_blocks(A, B):
    """Blocks of B vertices grown breadth-first over A's pattern."""
    rp, ci, _ = A.csr()
    n = A.m
    blk = -np.ones(n, np.int32)
    nb = 0
    for s in range(n):
        if blk[s] >= 0:
            continue
        q, cnt = deque([s]), 1
        blk[s] = nb
        while q and cnt < B:
            v = q.popleft()
            for w in ci[rp[v]:rp[v + 1]]:
                if blk[w] < 0 and cnt < B:
                    blk[w] = nb
                    cnt += 1
                    q.append(w)
        nb += 1
    return blk


def test_one_block_is_multicolour_gs(O):
    H = O.Hierarchy(O.icosahedron(), 3, project=True)
    cols = _colors(O, H)
    b = O.uniform_form(3, H.A[0].m)
    b -= b.mean()
    u_gs = H.mu_cycles(np.zeros_like(b), b, 2, alg=O.GS, steps=2, coarse_direct=True, colors=cols)
    u_bgs = H.mu_cycles(np.zeros_like(b), b, 2, alg=O.BLOCK_GS, steps=2, coarse_direct=True, colors=cols,
                        blocks=[np.zeros(a.m, np.int32) for a in H.A])
    assert np.array_equal(u_gs, u_bgs)


def test_singleton_blocks_are_jacobi(O):
    """One level, no coarse solve: the cycle is one forward sweep."""
    H = O.Hierarchy(O.icosphere(2), 0)
    n = H.A[0].m
    cols = _colors(O, H)
    u0, b = O.uniform_form(4, n), O.uniform_form(5, n)
    one = H.mu_cycles(u0, b, 1, alg=O.BLOCK_GS, steps=1, colors=cols, blocks=[np.arange(n, dtype=np.int32)])
    jac = H.mu_cycles(u0, b, 1, alg=O.JACOBI, steps=1, omega=1.0)
    assert np.allclose(one, jac, rtol=1e-12, atol=1e-12 * np.abs(jac).max())
    # and not the multicolour sweep (which uses the fresh values of earlier colours)
    gs = H.mu_cycles(u0, b, 1, alg=O.GS, steps=1, colors=cols)
    assert not np.allclose(one, gs, rtol=1e-6)


def test_convergence_rate_close_to_multicolour_gs(O):
    H = O.Hierarchy(O.icosahedron(), 5, project=True)
    L = H.A[0]
    cols = _colors(O, H)
    blks = [bfs_blocks(a, 64) if a.m > 3000 else np.zeros(a.m, np.int32) for a in H.A]
    assert blks[0].max() + 1 >= L.m // 64
    b = L.matvec(O.uniform_form(6, L.m))
    rates = {}
    for alg in (O.GS, O.BLOCK_GS):
        u = np.zeros(L.m)
        res = [np.linalg.norm(b)]
        for _ in range(4):
            u = H.mu_cycles(u, b, 1, alg=alg, steps=2, coarse_direct=True, colors=cols, blocks=blks)
            res.append(np.linalg.norm(b - L.matvec(u)))
        rates[alg] = (res[-1] / res[1]) ** (1 / 3)
    assert rates[O.GS] < 0.2
    assert rates[O.BLOCK_GS] <= 1.15 * rates[O.GS]
