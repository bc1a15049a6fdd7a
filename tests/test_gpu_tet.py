"""Tetrahedral complexes through the HIP library against the 3-D oracle
(oracle/tet.py): gluing and orient! bit for bit, the 3-D dual geometry, d, ∂
and d̃ bit for bit, the diagonal ⋆0..⋆3 and inverses, the wedges 0-1, 0-2,
0-3, 1-1, 1-2 and their transposed forms against the reference's pedagogical
wedges, ∇²₀ and Δᵈ0, and the reference's 3-D known answers run on the
library.  FP64 agrees to 1e-12 normwise (relative to the largest entry);
FP32 to 1e-5.

The circumcentric dual of the Blessent cube has dual triangles of zero area
(circumcenters on faces, as the reference's test notes).  The reference's
area formula sqrt(|uu ww - uw²|) turns the rounding of such a triangle into
about 1e-8 of the largest area, differently in any two restatements, so on
that mesh the quantities that sum dual areas (the 2-D dual areas, ⋆1, the
wedges 0-2 and 1-1 and ∇²₀) agree to 1e-7, and the inverse stars of cells
whose dual measure is zero (1/0 in the reference) are left out together with
Δᵈ0, which contains them."""
import math

import numpy as np
import pytest

from test_gpu_ops import F32_TOL, F64_TOL, rel, sparse_close, to_dev

pytestmark = pytest.mark.gpu

CASES = ["single", "blessent_circ", "blessent_bary", "kuhn_bary", "kuhn_circ", "kuhn_jitter_circ", "delaunay_bary"]
DEGENERATE = {"blessent_circ"}
SQRT_TOL = 1e-7


def area_tol(name, dtype="f64"):
    if dtype == "f32":
        return F32_TOL
    return SQRT_TOL if name in DEGENERATE else F64_TOL


def oracle_mesh(T, name):
    if name == "single":
        return T.single_tetrahedron(), False
    if name.startswith("blessent"):
        return T.blessent_cube(), name.endswith("circ")
    if name == "kuhn_bary":
        return T.kuhn_grid(3, 3, 2, jitter=0.2, seed=7), False
    if name == "kuhn_circ":
        return T.kuhn_grid(3, 2, 2, lx=3.0, ly=2.0, lz=1.5, shear=(0.5, 0.25)), True
    if name == "kuhn_jitter_circ":
        return T.kuhn_grid(3, 3, 3, jitter=0.15, seed=2, jitter_boundary=True), True
    # unstructured: scipy's Delaunay tetrahedralization of a jittered 5 x 5 x 5 lattice
    # (682 tetrahedra of varying shape and vertex degree), glued and oriented as
    # EmbeddedDeltaSet3D(mesh) does
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(*[np.arange(5.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    p = g + (rng.random(g.shape) - 0.5) * 0.5
    m = T.glue_tetrahedra(p, Delaunay(p).simplices)
    assert T.orient3(m)
    return m, False


def gpu_mesh3_like(cs, m):
    """The library's mesh of an oracle TetMesh (same topology and orientation state)."""
    return cs.mesh3_from_arrays(m.p, m.v0, m.v1, m.e0, m.e1, m.e2, m.t0, m.t1, m.t2, m.t3,
                                m.eo, m.to, m.teo)


@pytest.fixture(scope="module")
def T():
    import tet

    return tet


@pytest.fixture(scope="module", params=CASES)
def case(request, T, cs):
    m, circ = oracle_mesh(T, request.param)
    gm = gpu_mesh3_like(cs, m)
    od = T.TetDual(m, circumcenter_=circ)
    cs.dualize(gm, cs.Circumcenter if circ else cs.Barycenter)
    return request.param, od, gm


def csr3(A):
    A = A.tocsr()
    A.sort_indices()
    return A.indptr, A.indices, A.data


def tol(dtype):
    return F64_TOL if dtype == "f64" else F32_TOL


def td(torch, dtype):
    return torch.float64 if dtype == "f64" else torch.float32


def test_glue_and_orient_bit_exact(T, cs):
    cases = [T.kuhn_grid_arrays(3, 2, 2), T.kuhn_grid_arrays(2, 2, 2, jitter=0.1, seed=1)]
    for p, tets in cases:
        om = T.glue_tetrahedra(p, tets)
        assert T.orient3(om)
        g = cs.mesh_from_tetrahedra(p, tets).export()
        for a in ("v0", "v1", "e0", "e1", "e2", "t0", "t1", "t2", "t3"):
            assert np.array_equal(g[a], getattr(om, a)), a
        assert np.array_equal(g["tet_orientation"], om.teo)
        assert (g["edge_orientation"] == -1).all() and (g["tri_orientation"] == -1).all()
    # unsorted gluing (glue_tetrahedron!) of the square pyramid, 1-based ids
    pm = T.square_pyramid()
    g = cs.mesh_from_tetrahedra(pm.p, np.array([(0, 1, 2, 4), (1, 2, 3, 4)]) + 1, sorted_=False,
                                index_base=1).export()
    for a in ("v0", "v1", "e0", "e1", "e2", "t0", "t1", "t2", "t3"):
        assert np.array_equal(g[a], getattr(pm, a)), a


def test_dual_geometry(case):
    name, od, gm = case
    g = gm.export_dual()
    scale = np.abs(od.m.p).max()
    assert np.abs(g["dual_point"] - od.dual_point).max() <= F64_TOL * scale
    for key, ref in (("length", od.length), ("area", od.area), ("vol", od.vol),
                     ("dual_length", od.dual_length[:len(g["dual_length"])]), ("dual_vol", od.dual_vol)):
        assert rel(g[key], ref) <= F64_TOL, key
    assert rel(g["dual_area"], od.dual_area[:len(g["dual_area"])]) <= area_tol(name)
    # orientations after the dual complex is built: edges / triangles false, tetrahedra from orient!
    e = gm.export()
    assert np.array_equal(e["edge_orientation"], od.m.eo)
    assert np.array_equal(e["tri_orientation"], od.m.to)
    assert np.array_equal(e["tet_orientation"], od.m.teo)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_derivatives(case, cs, dtype):
    import torch

    name, od, gm = case
    t = td(torch, dtype)
    rng = np.random.default_rng(3)
    mats = [(cs.dec_differential(k, gm, cast_float=t), od.differential(k)) for k in range(3)]
    mats += [(cs.dec_boundary(k, gm, cast_float=t), od.boundary(k)) for k in (1, 2, 3)]
    mats += [(cs.dec_dual_derivative(k, gm, cast_float=t), od.dual_derivative(k)) for k in range(3)]
    for op, ref in mats:
        rp, ci, vv = op.csr()
        R = csr3(ref)
        assert np.array_equal(rp, R[0]) and np.array_equal(ci, R[1]) and np.array_equal(vv, R[2])
        x = rng.uniform(-1, 1, op.cols)
        got = (op @ to_dev(torch, x, t)).cpu().numpy()
        assert rel(got, ref @ x) <= tol(dtype)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_diagonal_hodge_stars(case, cs, dtype):
    import torch

    name, od, gm = case
    t = td(torch, dtype)
    for k in range(4):
        h = od.hodge_diag(k)
        live = np.abs(h) > SQRT_TOL * np.abs(h).max()          # cells of nonzero dual measure
        for inv in (False, True):
            fn = cs.dec_inv_hodge_star if inv else cs.dec_hodge_star
            op = fn(k, gm, hodge=cs.DiagonalHodge, cast_float=t)
            ref = od.inv_hodge_diag(k) if inv else h
            got = op.csr()[2]
            tk = area_tol(name) if k == 1 else F64_TOL
            assert rel(got[live], ref[live]) <= tk, (k, inv)
            x = np.random.default_rng(k).uniform(-1, 1, len(ref))
            y = (op @ to_dev(torch, x, t)).cpu().numpy()
            assert rel(y[live], (ref * x)[live]) <= max(tk, tol(dtype)), (k, inv)
        if name not in DEGENERATE:
            assert live.all()
    # the geometric inverse falls back to the diagonal one for 0 and 3 (FastDEC.jl:750-754)
    for k in (0, 3):
        g = cs.dec_inv_hodge_star(k, gm, hodge=cs.GeometricHodge, cast_float=torch.float64)
        assert np.array_equal(g.csr()[2], cs.dec_inv_hodge_star(k, gm, hodge=cs.DiagonalHodge).csr()[2])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_wedges(case, cs, dtype):
    import torch

    name, od, gm = case
    t = td(torch, dtype)
    rng = np.random.default_rng(9)
    n = {0: od.nv, 1: od.ne, 2: od.nt, 3: od.ntet}
    f = {k: rng.uniform(-1, 1, n[k]) for k in range(4)}
    g = {k: rng.uniform(-1, 1, n[k]) for k in range(4)}
    ref = {(0, 1): od.wedge0k(1, f[0], g[1]), (0, 2): od.wedge0k(2, f[0], g[2]),
           (0, 3): od.wedge0k(3, f[0], g[3]), (1, 1): od.wedge11(f[1], g[1]),
           (1, 2): od.wedge12(f[1], g[2]), (2, 1): od.wedge21(f[2], g[1])}
    for (a, b), r in ref.items():
        op = cs.dec_wedge_product(a, b, gm, cast_float=t)
        tw = area_tol(name, dtype) if (a, b) in ((0, 2), (1, 1)) else tol(dtype)
        got = op(to_dev(torch, f[a], t), to_dev(torch, g[b], t)).cpu().numpy()
        assert rel(got, r) <= tw, (a, b, rel(got, r))
        if a != b:                                         # the (k, 0) / (2, 1) forms swap their arguments
            sw = cs.dec_wedge_product(b, a, gm, cast_float=t)
            got = sw(to_dev(torch, g[b], t), to_dev(torch, f[a], t)).cpu().numpy()
            assert rel(got, r) <= tw, (b, a)
    with pytest.raises(cs.DecError):
        cs.dec_wedge_product(2, 2, gm)
    with pytest.raises(cs.DecError):
        cs.dec_wedge_product(1, 3, gm)


def test_wedge_tables_match_fastdec(case, cs):
    """The kernels' tables against FastDEC's: the 24 0-3 coefficients summed
    per vertex, and the 12 signed 2-1 coefficients with their edges and triangles."""
    import torch

    name, od, gm = case
    op3 = cs.dec_wedge_product(0, 3, gm, cast_float=torch.float64).operator
    idx, c = op3.tables(4, 4)
    verts, coeffs = od.fast_coeffs03()
    assert np.array_equal(idx, verts[::6])
    assert rel(c, coeffs.reshape(4, 6, -1).sum(1)) <= F64_TOL
    op21 = cs.dec_wedge_product(2, 1, gm, cast_float=torch.float64).operator
    idx, c = op21.tables(10, 12)
    ets, coeffs = od.fast_coeffs21()
    assert np.array_equal(idx, ets)
    assert rel(c, coeffs) <= F64_TOL
    assert np.array_equal(np.sign(c), np.sign(coeffs))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_laplacians(case, cs, dtype):
    import torch

    name, od, gm = case
    t = td(torch, dtype)
    tl = area_tol(name, dtype)
    L = cs.fast_laplace_beltrami(gm, hodge=cs.DiagonalHodge, cast_float=t)
    assert sparse_close(L.csr(), csr3(od.laplacian0()), od.nv, tol=area_tol(name))
    rng = np.random.default_rng(4)
    x0, x3 = rng.uniform(-1, 1, od.nv), rng.uniform(-1, 1, od.ntet)
    assert rel((L @ to_dev(torch, x0, t)).cpu().numpy(), od.laplacian0() @ x0) <= tl
    if name in DEGENERATE:
        return                                       # ⋆2⁻¹ of zero-length dual edges: 1/0 in the reference
    D = cs.dual_laplacian(0, gm, cast_float=t)
    assert sparse_close(D.csr(), csr3(od.dual_laplacian0()), od.ntet)
    assert rel((D @ to_dev(torch, x3, t)).cpu().numpy(), od.dual_laplacian0() @ x3) <= tol(dtype)


def test_operators_a_3d_complex_refuses(case, cs):
    name, od, gm = case
    from cspaces import _lib as L

    with pytest.raises(cs.DecError):
        cs.dec_hodge_star(1, gm, hodge=cs.GeometricHodge)
    with pytest.raises(cs.DecError, match="Geometric Hodge star in 3D"):
        cs.dec_inv_hodge_star(2, gm, hodge=cs.GeometricHodge)
    with pytest.raises(cs.DecError):
        cs.fast_laplace_beltrami(gm, hodge=cs.GeometricHodge)
    for kind, a, b in ((L.DEC_OP_WEDGE_DD, 0, 1), (L.DEC_OP_AVG01, 0, 0), (L.DEC_OP_FLAT_SHARP, 0, 0),
                       (L.DEC_OP_WEDGE_PD, 0, 1), (L.DEC_OP_INTERIOR_PRODUCT_DD, 1, 1), (L.DEC_OP_LIE_DD, 1, 1)):
        with pytest.raises(cs.DecError, match="2-D complexes only"):
            cs.Operator(gm, kind, a, b)
    with pytest.raises(cs.DecError):
        cs.dual_laplacian(1, gm)
    with pytest.raises(cs.DecError):
        cs.dec_differential(3, gm)
    with pytest.raises(cs.DecError):
        cs.subdivision(gm)


def test_apply_many_matches_single_applies(cs, T):
    import torch

    gm = cs.kuhn_box(6, 5, 4, jitter=0.2, seed=5)
    cs.dualize(gm, cs.Barycenter)
    ops = [cs.dec_wedge_product(0, 3, gm).operator, cs.dec_wedge_product(1, 2, gm).operator,
           cs.dec_wedge_product(2, 1, gm).operator, cs.dec_differential(2, gm), cs.dec_dual_derivative(0, gm),
           cs.dual_laplacian(0, gm), cs.dec_hodge_star(3, gm, hodge=cs.DiagonalHodge),
           cs.fast_laplace_beltrami(gm, hodge=cs.DiagonalHodge)]
    gen = torch.Generator(device="cuda").manual_seed(1)
    args = []
    for op in ops:
        x = torch.rand(op.cols, dtype=torch.float64, device="cuda", generator=gen)
        y = torch.rand(op.cols2, dtype=torch.float64, device="cuda", generator=gen) if op.cols2 else None
        args.append((op, x, y, torch.empty(op.rows, dtype=torch.float64, device="cuda")))
    cs.apply_many(args)
    for op, x, y, out in args:
        assert torch.equal(out, op.apply(x, y))


def test_reference_known_answers_on_the_library(T, cs):
    import torch

    # wdg01 / wdg02 / wdg03 of ones are ones, wdg11 of ones is zero (test/Operators.jl:261-266)
    om = T.kuhn_grid(3, 2, 2, lx=3.0, ly=2.0, lz=1.5, shear=(0.5, 0.25))
    gm = gpu_mesh3_like(cs, om)
    cs.dualize(gm, cs.Barycenter)
    one = lambda n: torch.ones(n, dtype=torch.float64, device="cuda")  # noqa: E731
    for k, n in ((1, gm.ne), (2, gm.nt), (3, gm.ntet)):
        w = cs.dec_wedge_product(0, k, gm)(one(gm.nv), one(n)).cpu().numpy()
        np.testing.assert_allclose(w, 1.0, rtol=1e-14)
    assert (cs.dec_wedge_product(1, 1, gm)(one(gm.ne), one(gm.ne)) == 0).all()
    # ⋆3((dX ∧ dY) ∧ dZ) = 1 with the right-handed orientation (test/DiscreteExteriorCalculus.jl:1054-1080)
    om = T.kuhn_grid(2, 2, 6, lx=2.0, ly=2.0, lz=12.0, jitter=0.2, seed=3)
    d = T.TetDual(om, circumcenter_=False)
    P, tv = d.m.p, d.tet_vertices
    signed = np.einsum("ik,ik->i", P[tv[:, 1]] - P[tv[:, 0]],
                       np.cross(P[tv[:, 2]] - P[tv[:, 0]], P[tv[:, 3]] - P[tv[:, 0]])) * d.sign(3)
    teo = d.m.teo if signed[0] > 0 else (1 - d.m.teo).astype(np.int8)
    m = om
    gm = cs.mesh3_from_arrays(m.p, m.v0, m.v1, m.e0, m.e1, m.e2, m.t0, m.t1, m.t2, m.t3, None, None, teo)
    cs.dualize(gm, cs.Barycenter)
    d0 = cs.dec_differential(0, gm)
    X, Y, Z = (to_dev(torch, P[:, i], torch.float64) for i in range(3))
    dX, dY, dZ = d0 @ X, d0 @ Y, d0 @ Z
    vol_form = cs.dec_wedge_product(2, 1, gm)(cs.dec_wedge_product(1, 1, gm)(dX, dY), dZ)
    h3 = cs.dec_hodge_star(3, gm, hodge=cs.DiagonalHodge)
    np.testing.assert_allclose((h3 @ vol_form).cpu().numpy(), 1.0, rtol=1e-12)
    # heat equation on the regular tetrahedron diffuses to the mean (:978-992)
    rm = T.regular_tetrahedron()
    gm = gpu_mesh3_like(cs, rm)
    cs.dualize(gm, cs.Circumcenter)
    L = cs.fast_laplace_beltrami(gm, hodge=cs.DiagonalHodge)
    C = to_dev(torch, np.linspace(1, 4, 4), torch.float64)
    mean = C.mean().item()
    for _ in range(5_000):
        C = C - 0.005 * (L @ C)
    np.testing.assert_allclose(C.cpu().numpy(), mean, rtol=1e-8)
    dv = gm.export_dual()["dual_vol"]
    np.testing.assert_allclose(dv, (2 * math.sqrt(2)) ** 3 / (6 * math.sqrt(2)) / 24, rtol=1e-12)


def test_large_box_properties(cs):
    """A 40 x 40 x 40 Kuhn box (384 000 tetrahedra): ⋆0 sums to the volume,
    d1 d0 = 0 and d2 d1 = 0 on random forms, ∇²₀ annihilates constants, and
    wedge 0-3 of a constant 0-form scales the 3-form."""
    import torch

    gm = cs.kuhn_box(40, 40, 40, lx=2.0, ly=1.0, lz=1.0, jitter=0.2, seed=11)
    cs.dualize(gm, cs.Barycenter)
    h0 = cs.dec_hodge_star(0, gm, hodge=cs.DiagonalHodge).csr()[2]
    assert math.isclose(h0.sum(), 2.0, rel_tol=1e-12)
    x = torch.rand(gm.nv, dtype=torch.float64, device="cuda")
    d0, d1, d2 = (cs.dec_differential(k, gm) for k in range(3))
    assert (d1 @ (d0 @ x)).abs().max().item() <= 1e-13
    y = torch.rand(gm.ne, dtype=torch.float64, device="cuda")
    assert (d2 @ (d1 @ y)).abs().max().item() <= 1e-13
    L = cs.fast_laplace_beltrami(gm, hodge=cs.DiagonalHodge)
    one = torch.ones(gm.nv, dtype=torch.float64, device="cuda")
    assert (L @ one).abs().max().item() <= 1e-9
    a = torch.rand(gm.ntet, dtype=torch.float64, device="cuda")
    w = cs.dec_wedge_product(0, 3, gm)(2.5 * one, a)
    assert torch.allclose(w, 2.5 * a, rtol=1e-13, atol=0)
