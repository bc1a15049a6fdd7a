"""The benchmark configurations at their own sizes: the level 3->8 Poisson
solves (655,362 unknowns) checked by the oracle's float64 Laplacian, the
float32 operator set at k=8 against the float64 oracle, the 64-instance batch
of icosphere levels 4-7 against single-operator applies, and the learned
mixed solve on 3->8."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sh_rhs(points, h0, seed, nterms=16, lmax=12):
    """Seeded sum of 16 real orthonormal spherical harmonics (l <= 12),
    projected to ⋆0-weighted mean zero."""
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(seed)
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    theta, phi = np.arccos(np.clip(z, -1, 1)), np.arctan2(y, x)
    f = np.zeros(len(points))
    for _ in range(nterms):
        l = int(rng.integers(1, lmax + 1))
        m = int(rng.integers(-l, l + 1))
        Y = sph_harm_y(l, abs(m), theta, phi)
        Yr = Y.real if m == 0 else math.sqrt(2) * (-1) ** m * (Y.real if m > 0 else Y.imag)
        f += rng.uniform(-1, 1) * Yr
    return f - (h0 @ f) / h0.sum()


@pytest.fixture(scope="module")
def k8(O, cs):
    od = O.Dual(O.icosphere(8))
    L = od.laplacian0()
    f = sh_rhs(od.m.p, od.hodge_diag(0), seed=2024)
    md = cs.MultigridData(cs.icosphere(3), cs.BinarySubdivision, 5, cs.DiagonalHodge, 2, project=True)
    assert md.sizes == [655362, 163842, 40962, 10242, 2562, 642]
    return od, L, f, md


@pytest.mark.parametrize("alg,variant", [(a, v) for a in ("block_gauss_seidel", "gauss_seidel") for v in
                                         ("plain", "mixed")] + [("block_gauss_seidel", "pcg"),
                                                                ("block_gauss_seidel", "pcg_mixed")])
def test_poisson_levels_3_to_8(cs, k8, variant, alg):
    import torch

    od, L, f, md = k8
    u, cycles, rr = cs.solve(md, torch.tensor(f, device="cuda"), tol=1e-8, variant=variant, alg=getattr(cs, alg))
    true = np.linalg.norm(f - L.matvec(u.cpu().numpy())) / np.linalg.norm(f)   # float64, oracle operator
    assert rr <= 1e-8 and true <= 1.01e-8
    assert abs(true - rr) <= 1e-3 * rr
    assert cycles <= (7 if variant.startswith("pcg") else 12)


def test_learned_mixed_solve_levels_3_to_8(cs, k8):
    """A network trained briefly (pairs from plain V-cycles on levels 3->5)
    smooths the float32 cycles on levels 10,242-163,842 of the 3->8 solve;
    the result meets 1e-8 by the oracle's float64 operator."""
    import torch

    od, L, f, md = k8
    mdt = cs.MultigridData(cs.icosphere(3), cs.BinarySubdivision, 2, cs.DiagonalHodge, 2, project=True)
    R, E = cs.make_pairs(mdt, 600, cycles_per_rhs=6, seed=3)
    nn_t = cs.LearnedSmoother(mdt, seed=5)
    nn_t.train(0, R, steps=200, E=E, batch=16, lr=2e-3, seed=6)
    nn = cs.LearnedSmoother(md, seed=5)
    nn.set_params(nn_t.get_params())
    nn.attach(True, min_rows=10242, max_rows=163842)
    try:
        u, cycles, rr = cs.solve(md, torch.tensor(f, device="cuda"), tol=1e-8, variant="mixed")
        st = cs.learned_stats(md)
    finally:
        nn.attach(False)
    true = np.linalg.norm(f - L.matvec(u.cpu().numpy())) / np.linalg.norm(f)
    assert rr <= 1e-8 and true <= 1.01e-8
    assert st["learned_cycles"] >= 1


OPS = ["wedge01", "wedge11", "wedge02", "d0", "d1", "hodge0", "hodge1", "hodge2", "laplacian0"]


def test_float32_operator_set_k8(O, cs, k8):
    """Every operator of the benchmark step in float32 at k=8 against the
    float64 oracle on the same inputs, on a strided sample of rows: the error
    is that of float32 rounding (inputs, tables and sums), stated as the bound
    max|y32 - y64| <= 3e-6 max|y64| and rms(y32 - y64) <= 1e-6 rms-scale."""
    import torch

    od, L, f, md = k8
    gm = cs.dualize(cs.icosphere(8))
    t32 = torch.float32
    V, E, E2, T = (O.uniform_form(s, n) for s, n in ((31, od.nv), (32, od.ne), (33, od.ne), (34, od.nt)))
    dv = lambda x: torch.tensor(x, dtype=t32, device="cuda")  # noqa: E731
    h = {k: od.hodge_diag(k) for k in range(3)}
    d0, d1 = od.differential(0), od.differential(1)
    refs = {
        "wedge01": (cs.dec_wedge_product(0, 1, gm, cast_float=t32)(dv(V), dv(E)), od.wedge01(V, E)),
        "wedge11": (cs.dec_wedge_product(1, 1, gm, cast_float=t32)(dv(E), dv(E2)), od.wedge11(E, E2)),
        "wedge02": (cs.dec_wedge_product(0, 2, gm, cast_float=t32)(dv(V), dv(T)), od.wedge02(V, T)),
        "d0": (cs.dec_differential(0, gm, cast_float=t32) @ dv(V), d0.matvec(V)),
        "d1": (cs.dec_differential(1, gm, cast_float=t32) @ dv(E), d1.matvec(E)),
        "hodge0": (cs.dec_hodge_star(0, gm, cs.DiagonalHodge, cast_float=t32) @ dv(V), h[0] * V),
        "hodge1": (cs.dec_hodge_star(1, gm, cs.DiagonalHodge, cast_float=t32) @ dv(E), h[1] * E),
        "hodge2": (cs.dec_hodge_star(2, gm, cs.DiagonalHodge, cast_float=t32) @ dv(T), h[2] * T),
        "laplacian0": (cs.fast_laplace_beltrami(gm, cs.DiagonalHodge, cast_float=t32) @ dv(V), L.matvec(V)),
    }
    assert sorted(refs) == sorted(OPS)
    for name, (got, ref) in refs.items():
        got = got.cpu().numpy().astype(np.float64)
        rows = np.arange(0, len(ref), 97)
        g, r = got[rows], ref[rows]
        scale = np.abs(ref).max()
        assert np.abs(g - r).max() <= 3e-6 * scale, name
        assert np.sqrt(np.mean((g - r) ** 2)) <= 1e-6 * scale, name


def test_batch_64_instances_levels_4_to_7(cs):
    """The benchmark's batch: 64 instances on a seeded uneven mix of levels
    4-7, one jittered mesh each; after 5 steps every instance equals the
    single-mesh operator's steps bit for bit (same sums, same order), in both
    precisions, and the float32 batch stays within 1e-5 of the float64 one."""
    import torch

    rng = np.random.default_rng(7)
    levels = [int(x) for x in rng.choice([4, 5, 6, 7], size=64, p=[0.4, 0.3, 0.2, 0.1])]
    assert sorted(set(levels)) == [4, 5, 6, 7]
    base = {k: cs.icosphere(k).export() for k in set(levels)}
    inst = []
    for i, k in enumerate(levels):
        e = base[k]
        p = e["points"].copy()
        hm = np.linalg.norm(p[e["v0"]] - p[e["v1"]], axis=1).mean()
        p += np.random.default_rng(1000 + i).uniform(-0.05, 0.05, p.shape) * hm
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        inst.append(cs.dualize(cs.mesh_from_arrays(p, e["v0"], e["v1"], e["e0"], e["e1"], e["e2"],
                                                   e["edge_orientation"], e["tri_orientation"])))
    b64 = cs.HeatBatch(inst, torch.float64)
    b32 = cs.HeatBatch(inst, torch.float32)
    assert b64.total == sum(10 * 4 ** k + 2 for k in levels)
    u0 = torch.empty(b64.total, dtype=torch.float64, device="cuda")
    for i in range(64):
        cs.fill_uniform(b64.instance(u0, i), 500 + i)
    u = u0.clone()
    b64.steps(u, 5, graph=True)
    v = u0.float()
    b32.tune(v.clone())
    b32.steps(v, 5, graph=True)
    for i, m in enumerate(inst):
        for dt, b, st, x in ((torch.float64, b64, u, u0), (torch.float32, b32, v, u0.float())):
            Lop = cs.fast_laplace_beltrami(m, cs.DiagonalHodge, cast_float=dt)
            y = b.instance(x, i).clone()
            for _ in range(5):
                y = y - b.dt[i] * (Lop @ y)
            assert torch.equal(b.instance(st, i), y), (i, dt)
        ref = b64.instance(u, i)
        assert float((b32.instance(v, i).double() - ref).norm() / ref.norm()) <= 1e-5
