"""The 3-D oracle (oracle/tet.py) pinned without a GPU by the reference's own
3-D tests and by what the mathematics fixes:

* cell counts of one tetrahedron and of the square pyramid, primal vertices
  of the 24 subsimplices, and the elementary-dual incidences
  (test/DiscreteExteriorCalculus.jl:890-941);
* d̃1 d̃0 = 0 and d̃2 d̃1 = 0 on an oriented tetrahedron and an oriented cube,
  and d̃_{3-k} = (-1)^k d_{k-1}ᵀ (:943-966);
* the regular tetrahedron: 24 circumcentric dual volumes of V/24, the
  circumcenter at the centroid, and heat diffusion to the mean (:968-992);
* the Blessent cube: barycentric dual volumes all 1/144 summing to 1, the
  circumcentric ones not summing to 1 (:994-1030);
* dXdY ∧ dZ = vol on one tetrahedron and ⋆3((dX ∧ dY) ∧ dZ) = 1 on a larger
  mesh (:1032-1080; a Kuhn box replaces TetGen's README mesh);
* the FastDEC 0-k wedges of ones, 1-1 of ones = 0, FastDEC's 0-3 / 2-1 tables
  against the pedagogical wedges, and 1-2 invariance under flipping an edge
  and a triangle with their values (test/Operators.jl:250-300; a Kuhn
  parallelepiped replaces TetGen's);
* Σ_σ |σ| |⋆σ| = C(3, k) · volume for circumcentric duals of meshes whose
  circumcenters lie in their simplices, and L·1 = 0, Δᵈ0 = -⋆3 (d2 ⋆2⁻¹ d2ᵀ)
  symmetric up to ⋆3.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def T():
    import tet

    return tet


def test_single_tetrahedron_counts_and_incidences(T):
    d = T.TetDual(T.single_tetrahedron(), geometry=False)
    m = d.m
    assert d.ndv == m.nv + m.ne + m.nt + m.ntet == 15
    assert d.nde == 6 * 4 + 2 * 6 + 14
    assert d.ndt == 60 and d.ndtet == 24
    subs = d.subsimplices(3)
    assert d.primal_vertex(3, subs[0]).tolist() == np.repeat(np.arange(4), 6).tolist()
    tc = d.tetc[0]
    assert [x.tolist() for x in d.elementary_duals(3)] == [[tc]]
    de = d.elementary_duals(2)[0]
    assert d.D_v0[de].tolist() == [tc] and d.D_v1[de].tolist() == [d.tc[0]]
    dts = d.elementary_duals(1)[0]
    assert d.D_v0[d.D_e0[dts]].tolist() == [tc, tc]
    dtets = d.elementary_duals(0)[0]
    assert d.D_v0[d.D_e0[d.D_t0[dtets]]].tolist() == [tc] * 6


def test_square_pyramid_counts(T):
    d = T.TetDual(T.square_pyramid(), geometry=False)
    assert (d.ndv, d.nde, d.ndt, d.ndtet) == (5 + 1 + 2 + 9 + 6, (6 * 4 + 2 * 6 + 14) * 2 - 12, 60 * 2 - 6, 48)
    subs = d.subsimplices(3)
    assert d.primal_vertex(3, subs[0]).tolist() == np.repeat([0, 1, 2, 4], 6).tolist()
    assert d.primal_vertex(3, subs[1]).tolist() == np.repeat([1, 2, 3, 4], 6).tolist()


def _oriented_tetrahedron(T):
    m = T.glue_tetrahedra(np.eye(4, 3), [(0, 1, 2, 3)], sorted_=False, tet_orientation=1)
    m.to[:] = 1
    m.eo[:] = 1
    return m


@pytest.mark.parametrize("which", ["tetrahedron", "cube"])
def test_dual_derivatives_compose_to_zero(T, which):
    m = _oriented_tetrahedron(T) if which == "tetrahedron" else T.tet_cube(tet_orientation=1)
    d = T.TetDual(m, geometry=False)
    dd = [d.dual_derivative(k) for k in range(3)]
    assert abs(dd[1] @ dd[0]).max() == 0
    assert abs(dd[2] @ dd[1]).max() == 0
    for k in (1, 2, 3):
        assert abs(dd[3 - k] - (-1) ** k * d.differential(k - 1).T).max() == 0
    assert abs(d.differential(1) @ d.differential(0)).max() == 0
    assert abs(d.differential(2) @ d.differential(1)).max() == 0


def test_regular_tetrahedron(T):
    d = T.TetDual(T.regular_tetrahedron(), circumcenter_=True)
    V = (2 * math.sqrt(2)) ** 3 / (6 * math.sqrt(2))
    np.testing.assert_allclose(d.dual_vol, V / 24, rtol=1e-12)
    np.testing.assert_allclose(d.dual_point[d.tetc[0]], d.m.p.sum(0) / 4, atol=1e-15)
    # heat equation with ⋆0⁻¹ d̃2 ⋆1 d0 converges to the mean
    lap = (np.diag(d.inv_hodge_diag(0)) @ d.dual_derivative(2).toarray() @ np.diag(d.hodge_diag(1))
           @ d.differential(0).toarray())
    C = np.linspace(1, 4, d.nv)
    C0 = C.copy()
    for _ in range(100_001):
        C = C + 0.005 * (lap @ C)
    np.testing.assert_allclose(C, C0.mean(), rtol=1e-8)
    np.testing.assert_allclose(-lap, d.laplacian0().toarray(), rtol=1e-14, atol=1e-14)


def test_blessent_cube_dual_volumes(T):
    m = T.blessent_cube()
    circ = T.TetDual(m, circumcenter_=True)
    assert not math.isclose(circ.dual_vol.sum(), 1.0, rel_tol=1.5e-8)
    bary = T.TetDual(m, circumcenter_=False)
    assert math.isclose(bary.dual_vol.sum(), 1.0, rel_tol=1e-13)
    np.testing.assert_allclose(bary.dual_vol, 1 / bary.ndtet, rtol=1e-12)
    np.testing.assert_allclose(bary.vol, 1 / 6, rtol=1e-14)


def _as_vec(d, e):
    m = d.m
    return (m.p[m.v0[e]] - m.p[m.v1[e]]) * d.sign(1)[e][..., None]


def surface_integral_dXdY(d):
    """(cross(e3_vec, e2_vec) sign(2, s, tri))[3] / 2 with (_, e2, e3) the
    triangle's edges (test/DiscreteExteriorCalculus.jl:1035-1046)."""
    m = d.m
    c = np.cross(_as_vec(d, m.e2), _as_vec(d, m.e1)) * d.sign(2)[:, None]
    return c[:, 2] / 2


def test_wedge21_of_area_and_height_is_volume(T):
    d = T.TetDual(T.single_tetrahedron(), circumcenter_=False)
    dZ = d.differential(0) @ d.m.p[:, 2]
    w = d.wedge21(surface_integral_dXdY(d), dZ)
    np.testing.assert_allclose(w, d.vol, rtol=1e-14)
    np.testing.assert_allclose(d.fast_wedge21(surface_integral_dXdY(d), dZ), d.vol, rtol=1e-14)


def _box(T):
    """A 2 x 2 x 12 box (the extent of TetGen's README mesh) of jittered Kuhn
    tetrahedra with barycentric duals; tetrahedron orientation flipped where
    orient! chose the left-handed one, as the reference test does."""
    m = T.kuhn_grid(2, 2, 6, lx=2.0, ly=2.0, lz=12.0, jitter=0.2, seed=3)
    d = T.TetDual(m, circumcenter_=False)
    P, tv = d.m.p, d.tet_vertices
    signed = np.einsum("ik,ik->i", P[tv[:, 1]] - P[tv[:, 0]],
                       np.cross(P[tv[:, 2]] - P[tv[:, 0]], P[tv[:, 3]] - P[tv[:, 0]])) * d.sign(3)
    assert (signed > 0).all() or (signed < 0).all()        # orient! is consistent
    if signed[0] < 0:
        d.m.teo = (1 - d.m.teo).astype(np.int8)
    return d


def test_volume_form_from_basis_one_forms(T):
    d = _box(T)
    P = d.m.p
    d0 = d.differential(0)
    dX, dY, dZ = d0 @ P[:, 0], d0 @ P[:, 1], d0 @ P[:, 2]
    np.testing.assert_allclose(d.wedge21(surface_integral_dXdY(d), dZ), d.vol, rtol=1e-12)
    vol_form = d.wedge21(d.wedge11(dX, dY), dZ)
    np.testing.assert_allclose(d.hodge_diag(3) * vol_form, 1.0, rtol=1e-12)


@pytest.fixture(scope="module")
def dual_meshes_3d(T):
    """single_tetrahedron() and a Kuhn parallelepiped (lx 3, ly 2, lz 1.5,
    top face shifted by (0.5, 0.25)) with barycentric duals."""
    pp = T.kuhn_grid(3, 2, 2, lx=3.0, ly=2.0, lz=1.5, shear=(0.5, 0.25))
    return [T.TetDual(T.single_tetrahedron(), circumcenter_=False), T.TetDual(pp, circumcenter_=False)]


def test_fast_wedges_against_pedagogical(T, dual_meshes_3d):
    rng = np.random.default_rng(11)
    for d in dual_meshes_3d:
        V1, E1, T2, Tet2 = rng.random(d.nv), rng.random(d.ne), rng.random(d.nt), rng.random(d.ntet)
        ones = lambda n: np.ones(n)
        np.testing.assert_allclose(d.wedge0k(1, ones(d.nv), ones(d.ne)), 1.0, rtol=1e-15)
        np.testing.assert_allclose(d.wedge0k(2, ones(d.nv), ones(d.nt)), 1.0, rtol=1e-14)
        np.testing.assert_allclose(d.wedge0k(3, ones(d.nv), ones(d.ntet)), 1.0, rtol=1e-14)
        np.testing.assert_allclose(d.fast_wedge03(ones(d.nv), ones(d.ntet)), 1.0, rtol=1e-14)
        assert np.all(d.wedge11(ones(d.ne), ones(d.ne)) == 0)
        np.testing.assert_allclose(d.fast_wedge03(V1, Tet2), d.wedge0k(3, V1, Tet2), rtol=1e-13)
        np.testing.assert_allclose(d.fast_wedge21(T2, E1), d.wedge12(E1, T2), rtol=1e-13)
        np.testing.assert_allclose(d.fast_wedge21(ones(d.nt), ones(d.ne)), d.wedge12(ones(d.ne), ones(d.nt)),
                                   rtol=1e-13, atol=1e-15)


def test_wedge12_invariant_under_flipped_orientation(T):
    d = T.TetDual(T.single_tetrahedron(), circumcenter_=False)
    rng = np.random.default_rng(5)
    E1, T2 = rng.random(d.ne), rng.random(d.nt)
    for i, j in zip(range(d.ne), range(d.nt)):
        ref = d.fast_wedge21(T2, E1)
        d.m.eo[i] ^= 1
        d.m.to[i] ^= 1
        E1[i] = -E1[i]
        T2[j] = -T2[j]
        np.testing.assert_allclose(d.fast_wedge21(T2, E1), ref, rtol=1e-14)
        np.testing.assert_allclose(d.wedge12(E1, T2), ref, rtol=1e-14)


@pytest.mark.parametrize("mesh", ["regular", "kuhn"])
def test_circumcentric_primal_dual_volume_identity(T, mesh):
    m = T.regular_tetrahedron() if mesh == "regular" else T.kuhn_grid(3, 3, 2, lx=1.5, ly=1.5, lz=1.0)
    d = T.TetDual(m, circumcenter_=True)
    V = d.vol.sum()
    np.testing.assert_allclose(d.hodge_diag(0).sum(), V, rtol=1e-13)
    np.testing.assert_allclose((d.hodge_diag(1) * d.length ** 2).sum(), 3 * V, rtol=1e-13)
    np.testing.assert_allclose((d.hodge_diag(2) * d.area ** 2).sum(), 3 * V, rtol=1e-13)


def test_laplacians(T):
    d = T.TetDual(T.kuhn_grid(3, 2, 2, jitter=0.15, seed=1), circumcenter_=False)
    L = d.laplacian0()
    assert np.abs(L @ np.ones(d.nv)).max() <= 1e-12
    D = d.dual_laplacian0().toarray()
    S = np.diag(1 / d.hodge_diag(3)) @ D
    np.testing.assert_allclose(S, S.T, atol=1e-12 * np.abs(S).max())
    assert np.linalg.eigvalsh((S + S.T) / 2).max() <= 1e-10 * np.abs(S).max()
    interior = np.ones(d.ntet)
    bnd = np.zeros(d.nt, int)
    for k in range(4):
        np.add.at(bnd, d.m.tetrahedron_triangles()[:, k], 1)
    touches = np.isin(d.m.tetrahedron_triangles(), np.flatnonzero(bnd == 1)).any(1)
    assert np.abs((D @ interior)[~touches]).max() <= 1e-10


def test_orient_rejects_a_nonorientable_gluing(T):
    # two tetrahedra sharing triangle (0,1,2) with ∂t3 on both sides: the
    # depth-first pass orients the second opposite to the first, and a third
    # sharing the same face on ∂t3 with the second cannot be oriented
    m = T.glue_tetrahedra(np.zeros((6, 3)), [(0, 1, 2, 3), (0, 1, 2, 4), (0, 1, 2, 5)])
    assert not T.orient3(m)
