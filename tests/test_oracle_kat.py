"""Pins the CPU oracle against CombinatorialSpaces.jl's own known answers.

Every expected value below is copied from a reference test (file:line cited)
or fixed by mathematics (closed forms, invariants).  Indices are 0-based here;
the reference's are 1-based.
"""
import math

import numpy as np
import pytest


def square(O, eo=None, to=None, points=None):
    p = np.zeros((4, 3)) if points is None else points
    m = O.glue_triangles(p, [[0, 1, 2], [0, 2, 3]])
    if eo is not None:
        m.eo[:] = eo
    if to is not None:
        m.to[:] = to
    return m


# --- test/SimplicialSets.jl:93-145 -------------------------------------------
def test_single_triangle_faces(O):
    m = O.glue_triangles(np.zeros((3, 3)), [[0, 1, 2]])
    assert (m.e0[0], m.e1[0], m.e2[0]) == (1, 2, 0)          # (2,3,1) 1-based
    assert list(m.triangle_vertices()[0]) == [0, 1, 2]
    s2 = O.glue_triangles(np.zeros((3, 3)), [[1, 2, 0]], sorted_=True)
    for a in ("v0", "v1", "e0", "e1", "e2"):
        assert np.array_equal(getattr(m, a), getattr(s2, a))


def test_commutative_square(O):
    m = O.glue_triangles(np.zeros((4, 3)), [[0, 1, 2], [0, 3, 2]])
    assert m.nt == 2 and m.ne == 5
    pairs = sorted(zip(m.v1.tolist(), m.v0.tolist()))  # src => tgt
    assert pairs == [(0, 1), (0, 2), (0, 3), (1, 2), (3, 2)]


def test_square_orientation_and_d1(O):
    m = square(O, eo=1)
    assert O.orient(m)
    assert list(m.to) == [1, 1]
    d = O.Dual(m)
    B2 = d.boundary(2).dense()
    assert list(B2[:, 0]) == [1, 1, -1, 0, 0]
    D1 = d.differential(1).dense()
    assert list(D1 @ [45, 3, 34, 0, 0]) == [14, 34]
    assert list(D1 @ [45, 3, 34, 17, 5]) == [14, 46]
    assert np.array_equal(D1, B2.T)


# --- test/DiscreteExteriorCalculus.jl:183-233 --------------------------------
def test_dual_complex_numbering(O):
    m = square(O)
    d = O.Dual(m)
    nv, ne, nt = 4, 5, 2
    assert len(d.dual_point) == nv + ne + nt
    assert len(d.D_v0) == 2 * ne + 6 * nt and len(d.D_e0) == 6 * nt
    # primal_vertex(subsimplices(Tri(1))) == [1,1,2,2,3,3]
    verts, _ = d.wedge02_coeffs()
    assert list(verts[:6]) == [0, 0, 1, 1, 2, 2]
    assert tuple(int(x) for x in d.dual_triangle_vertices(0)) == (0, 6, 9)   # [1,7,10]
    assert tuple(int(x) for x in d.dual_edge_vertices(0)) == (4, 1)          # [5,2]
    counts = [int(np.sum(d.D_v1[d.D_e1] == v)) for v in range(4)]
    assert counts == [4, 2, 4, 2]
    # elementary duals of edge 2 / edge 3 (1-based) end at their edge centers
    assert list(d.D_v1[d.D_v1 == nv + 1]) == [nv + 1]
    assert list(d.D_v1[d.D_v1 == nv + 2]) == [nv + 2, nv + 2]


@pytest.mark.parametrize("explicit", [True, False])
def test_dual_orientation(O, explicit):
    m = square(O, eo=1 if explicit else None, to=1 if explicit else None)
    d = O.Dual(m)
    assert d.D_to.sum() == 6 * m.nt // 2
    per_v = [int(d.D_to[np.nonzero(d.D_v1[d.D_e1] == v)[0]].sum()) for v in range(4)]
    assert per_v == [2, 1, 2, 1]
    assert int(d.D_eo[d.D_v1 == 4 + 2].sum()) == 1
    # dual_derivative(2-k) == (-1)^k d(k-1)'  (Desbrun-Kanso-Tong 4.2)
    assert np.array_equal(d.dual_derivative(1).dense(), -d.differential(0).dense().T)
    assert np.array_equal(d.dual_derivative(0).dense(), d.differential(1).dense().T)


# --- test/DiscreteExteriorCalculus.jl:288-371 (Gillette §2.13) ---------------
def right_triangle(O, circumcenter):
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    m = O.glue_triangles(P, [[0, 1, 2]])
    m.to[:] = 1
    m.eo[:] = 1
    return O.Dual(m, circumcenter=circumcenter)


def test_gillette_barycenter(O):
    d = right_triangle(O, False)
    assert np.allclose(d.dual_point[-1][:2], [1 / 3, 1 / 3], rtol=0, atol=1e-15)
    assert np.allclose(d.hodge_diag(0), [1 / 6] * 3, rtol=1e-14)
    assert np.allclose(d.hodge_diag(1), [math.sqrt(5) / 6, 1 / 6, math.sqrt(5) / 6], rtol=1e-14)
    assert np.allclose(d.geo_hodge1().dense(), [[1 / 3, 0, 1 / 6], [0, 1 / 6, 0], [1 / 6, 0, 1 / 3]], rtol=1e-14)
    assert np.allclose(d.inv_hodge_diag(2), [0.5], rtol=1e-14)
    assert np.allclose(d.inv_hodge_diag(1), [-6 / math.sqrt(5), -6, -6 / math.sqrt(5)], rtol=1e-14)
    delta1 = np.diag(d.inv_hodge_diag(0)) @ d.dual_derivative(1).dense() @ d.geo_hodge1().dense()
    assert np.allclose(delta1, [[3, 0, 3], [-2, 1, -1], [-1, -1, -2]], atol=1e-12)
    # Δ0 = [-6 3 3; 3 -3 0; 3 0 -3] and ∇²0 = -Δ0
    assert np.allclose(d.laplacian0(geometric=True).dense(), -np.array([[-6, 3, 3], [3, -3, 0], [3, 0, -3]]),
                       atol=1e-12)


def test_gillette_circumcenter(O):
    d = right_triangle(O, True)
    assert np.allclose(d.dual_point[-1][:2], [0.5, 0.5], atol=1e-15)
    assert np.allclose(d.hodge_diag(0), [1 / 4, 1 / 8, 1 / 8], rtol=1e-14)
    assert np.allclose(d.hodge_diag(1), [0.5, 0.0, 0.5], atol=1e-15)
    delta1 = np.diag(d.inv_hodge_diag(0)) @ d.dual_derivative(1).dense() @ d.geo_hodge1().dense()
    assert np.allclose(delta1, [[2, 0, 2], [-4, 0, 0], [0, 0, -4]], atol=1e-12)


# --- wedges: test/Operators.jl:225-303, 339-359; DEC tests 538-570 -----------
def grid_345(O):
    """grid_345() (src/CombMeshes.jl:363-383)."""
    P = np.array([[0, 4, 0], [3, 4, 0], [6, 4, 0], [0, 0, 0], [3, 0, 0], [6, 0, 0],
                  [0, -4, 0], [3, -4, 0], [6, -4, 0]], float)
    tris = [[0, 1, 3], [4, 1, 3], [4, 1, 2], [4, 5, 2], [4, 6, 3], [4, 6, 7], [4, 5, 7], [8, 5, 7]]
    m = O.glue_triangles(P, tris, sorted_=True)
    m.eo[:] = 1
    assert O.orient(m)
    m.to = (1 - m.to).astype(np.int8)
    return O.Dual(m, circumcenter=False)


def const_form(d, vec):
    """eval_constant_primal_form: dot(v, p[tgt] - p[src]) * sign(edge)."""
    m = d.m
    s = np.where(d.eo == 1, 1.0, -1.0)
    return (m.p[m.v0] - m.p[m.v1]) @ np.asarray(vec, float) * s


def test_wedge11_dx_dy_is_area(O):
    d = grid_345(O)
    dx, dy = const_form(d, [1, 0, 0]), const_form(d, [0, 1, 0])
    assert np.allclose(d.wedge11(dx, dy), d.area, rtol=1e-12)
    assert np.array_equal(d.wedge11(dx, dy), -d.wedge11(dy, dx))


def test_wedge11_cw_mesh_is_minus_area(O):
    P = np.array([[0, 0, 0], [3, 0, 0], [3, 4, 0]], float)     # tri_345_false
    m = O.glue_triangles(P, [[0, 1, 2]])
    m.to[:] = 0
    m.eo[:] = 1
    d = O.Dual(m, circumcenter=False)
    dx, dy = const_form(d, [1, 0, 0]), const_form(d, [0, 1, 0])
    assert np.allclose(d.wedge11(dx, dy), -d.area, rtol=1e-12)


def test_wedge_identities(O):
    d = O.Dual(O.icosphere(2))
    V = O.uniform_form(1, d.nv)
    E = O.uniform_form(2, d.ne)
    E2 = O.uniform_form(3, d.ne)
    T = O.uniform_form(4, d.nt)
    # wdg01(V_ones, E_ones) .== E_ones
    assert np.array_equal(d.wedge01(np.ones(d.nv), np.ones(d.ne)), np.ones(d.ne))
    # avg₀₁_mat * V .* E == wedge01(V, E), exactly
    A = d.differential(0)
    A.nzval[:] = 0.5
    assert np.array_equal(A.matvec(V) * E, d.wedge01(V, E))
    # α ∧ α = 0 exactly; antisymmetry exactly
    assert np.array_equal(d.wedge11(E, E), np.zeros(d.nt))
    assert np.array_equal(d.wedge11(E, E2), -d.wedge11(E2, E))
    # 1 ∧ β = β for a 2-form (Σ of the six dual areas is the area)
    assert np.allclose(d.wedge02(np.ones(d.nv), T), T, rtol=1e-13)
    # flipping an edge's orientation and the sign of its values preserves 1-1
    e = 17
    w0 = d.wedge11(E, E2)
    d.eo[e] = 1 - d.eo[e]
    E[e], E2[e] = -E[e], -E2[e]
    assert np.allclose(d.wedge11(E, E2), w0, rtol=1e-14, atol=1e-15)


def test_wedge_dd_weights(O):
    d = O.Dual(O.icosphere(2))
    M = d.wedge_dd01_mat()
    D = M.dense()
    assert np.allclose(D.sum(axis=1), 1.0, rtol=1e-15)     # weights sum to one
    assert np.all((D > 0).sum(axis=1) == 2)                  # closed surface: two cofaces


# --- brute force from the points alone ---------------------------------------
def _tri_centers(m, circumcenter):
    """Triangle centers from the three points, independent of the oracle's
    dual complex: the centroid, or the circumcenter in closed form
    a + (|u|² v×(u×v) + |v|² (u×v)×u) / (2|u×v|²), u = b - a, v = c - a."""
    V = m.triangle_vertices()
    a, b, c = m.p[V[:, 0]], m.p[V[:, 1]], m.p[V[:, 2]]
    if not circumcenter:
        return (a + b + c) / 3
    u, v = b - a, c - a
    w = np.cross(u, v)
    num = (u * u).sum(1)[:, None] * np.cross(v, w) + (v * v).sum(1)[:, None] * np.cross(w, u)
    return a + num / (2 * (w * w).sum(1))[:, None]


def _area(a, b, c):
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=-1)


def brute_wedge_dd_weights(m, circumcenter):
    """m[e, t] = |c_t - mid_e| / Σ_{t' ∋ e} |c_t' - mid_e| (wedge_dd_01_mat,
    src/FastDEC.jl:293-305: the elementary duals of e run from e's midpoint to
    the centers of its cofaces), as a dense ne x nt array."""
    ct = _tri_centers(m, circumcenter)
    mid = 0.5 * m.p[m.v0] + 0.5 * m.p[m.v1]
    W = np.zeros((m.ne, m.nt))
    for E in (m.e0, m.e1, m.e2):
        t = np.arange(m.nt)
        W[E, t] = np.linalg.norm(ct - mid[E], axis=1)
    return W / W.sum(axis=1, keepdims=True)


def dd_weights_match(M, m, circumcenter):
    return np.allclose(M.dense(), brute_wedge_dd_weights(m, circumcenter), rtol=0, atol=1e-13)


def _swap_two_weights(M):
    """A planted mistake: the two weights of the two-coface edge whose weights
    differ most are exchanged between its triangles (row sums and positivity
    unchanged)."""
    D = M.dense().copy()
    two = np.flatnonzero((D > 0).sum(axis=1) == 2)
    spread = D[two].max(axis=1) - np.where(D[two] > 0, D[two], np.inf).min(axis=1)
    e = int(two[np.argmax(spread)])
    t1, t2 = np.flatnonzero(D[e])
    D[e, t1], D[e, t2] = D[e, t2], D[e, t1]
    class _Dense:
        def dense(self):
            return D
    return _Dense()


def _jittered_grid(O):
    """triangulated_grid(1, 1, 1/4, √3/8) with its points moved by a seeded
    in-plane jitter of up to 15% of the spacing, so that no two cofaces of an
    edge are mirror images."""
    m = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 8, compress=False)
    rng = np.random.default_rng(5)
    m.p[:, :2] += rng.uniform(-0.15, 0.15, (m.nv, 2)) * (1 / 4)
    return m


@pytest.mark.parametrize("case", ["icosphere2_circumcenter", "grid_barycenter"])
def test_wedge_dd_weights_brute_force(O, case):
    """The oracle's dual-dual 0-1 matrix equals the weights computed from the
    points alone, on a curved closed surface (circumcentric, two cofaces per
    edge) and a jittered flat grid with boundary edges (barycentric, one coface
    there, weight 1); a copy with two weights of one edge swapped fails the check."""
    if case == "icosphere2_circumcenter":
        m, circ = O.icosphere(2), True
    else:
        m, circ = _jittered_grid(O), False
    d = O.Dual(m, circumcenter=circ)
    M = d.wedge_dd01_mat()
    assert dd_weights_match(M, m, circ)
    W = brute_wedge_dd_weights(m, circ)
    assert len(np.unique(np.round(W[W > 0], 12))) > 2       # weights are not all equal
    assert not dd_weights_match(_swap_two_weights(M), m, circ)


def test_wedge_dd_reference_cases(O):
    """test/Operators.jl:419-437 on the flat meshes: f = ⋆(1/√2 dx + 1/√2 dy)
    (geometric ⋆1, the default Hodge), g = 5 on triangles; Λ10(f, g) ≈
    ⋆(5/√2 dx + 5/√2 dy) and Λ10(f, g) == Λ01(g, f) exactly."""
    from test_oracle_musical import flat_primal_meshes

    s2 = 1 / math.sqrt(2)
    for name, m in flat_primal_meshes(O).items():
        d = O.Dual(m, circumcenter=False)
        H1 = d.geo_hodge1()
        f = H1.matvec(_constant_primal_form(d, (s2, s2, 0)))
        g = np.full(d.nt, 5.0)
        l10, l01 = d.wedge_dd10(f, g), d.wedge_dd01(g, f)
        ref = H1.matvec(_constant_primal_form(d, (5 * s2, 5 * s2, 0)))
        assert np.allclose(l10, ref, rtol=1.5e-8, atol=1e-14), name
        assert np.array_equal(l10, l01), name


def brute_wedge02(m, circumcenter, f, a):
    """∧(0, 2) of the reference's pedagogical wedge (wedge_product_zero,
    src/DiscreteExteriorCalculus.jl:2084-2090) from the points alone: a
    triangle's six dual triangles are (p_v, mid_e, c_t) for its vertices v and
    the two edges e of t at v; the weight of f[v] is the sum of its two dual
    areas over the triangle's area."""
    ct = _tri_centers(m, circumcenter)
    V = m.triangle_vertices()
    P = m.p[V]                                          # nt x 3 x 3
    area = _area(P[:, 0], P[:, 1], P[:, 2])
    res = np.zeros(m.nt)
    for i in range(3):
        pv = P[:, i]
        wsum = np.zeros(m.nt)
        for j in range(3):
            if j != i:
                wsum += _area(pv, 0.5 * pv + 0.5 * P[:, j], ct)
        res += wsum / area * f[V[:, i]]
    return res * a


@pytest.mark.parametrize("case", ["icosphere2_circumcenter", "grid_barycenter", "tri_345"])
def test_wedge02_equals_generic_wedge(O, case):
    """dec_wedge_product(0, 2) == ∧(0, 2) (test/Operators.jl:245) with random
    0-forms: the oracle's FastDEC table (six vertex/coefficient slots) agrees
    with the per-simplex wedge built from the points; a copy of the table with
    two vertex slots exchanged does not."""
    if case == "icosphere2_circumcenter":
        m, circ = O.icosphere(2), True
    elif case == "grid_barycenter":
        m, circ = _jittered_grid(O), False
    else:
        from test_oracle_musical import flat_primal_meshes

        m, circ = flat_primal_meshes(O)["tri_345"], True
    d = O.Dual(m, circumcenter=circ)
    f, a = O.uniform_form(11, d.nv), O.uniform_form(12, d.nt)
    got, ref = d.wedge02(f, a), brute_wedge02(m, circ, f, a)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-15)
    if not circ:
        # barycentric: all six dual areas are area/6, so 0-2 is the plain
        # vertex mean and no vertex permutation can change it
        assert np.allclose(got, a * f[m.triangle_vertices()].mean(axis=1), rtol=1e-13, atol=1e-15)
        return
    # planted mistake: in the triangle where it matters most, the first slot
    # points at the vertex of the third slot and vice versa
    verts, c = d.wedge02_coeffs()
    V6, C6 = verts.reshape(d.nt, 6), c.reshape(d.nt, 6)
    t = int(np.argmax(np.abs((C6[:, 0] - C6[:, 2]) * (f[V6[:, 0]] - f[V6[:, 2]]) * a)))
    bad = V6.copy()
    assert bad[t, 0] != bad[t, 2]
    bad[t, [0, 2]] = bad[t, [2, 0]]
    res = np.zeros(d.nt)
    O.lib().om_wedge02(O.c_int(d.nt), O._p(bad.reshape(-1)), O._p(c), O._p(f), O._p(a), O._p(res))
    assert np.array_equal(np.delete(res, t), np.delete(got, t))
    assert not np.isclose(res[t], ref[t], rtol=1e-13, atol=1e-15)


# --- meshes and subdivision: test/CombMeshes.jl:14-84, test/Multigrid.jl:14-41 -
def test_icosphere_counts(O):
    expect = [(12, 30, 20), (42, 120, 80), (162, 480, 320), (642, 1920, 1280),
              (2562, 7680, 5120), (10242, 30720, 20480)]
    for k, (nv, ne, nt) in enumerate(expect):
        m = O.icosphere(k)
        assert (m.nv, m.ne, m.nt) == (nv, ne, nt)
        assert m.nv - m.ne + m.nt == 2
        assert np.allclose(np.linalg.norm(m.p, axis=1), 1.0, rtol=0, atol=4.5e-16)


def test_subdivision_matrix_and_counts(O):
    s = O.triangulated_grid(1, 1, 1, 1, compress=False)
    S = O.subdivision_matrix(s).dense()
    assert np.array_equal(S[:s.nv, :s.nv], np.eye(s.nv))
    for e in range(s.ne):
        col = S[:, s.nv + e]
        rows = np.nonzero(col)[0]
        assert list(rows) == [s.v1[e], s.v0[e]] and list(col[rows]) == [0.5, 0.5]
    m = s
    for _ in range(4):
        t = O.subdivide(m)
        assert (t.nv, t.ne, t.nt) == (m.nv + m.ne, 2 * m.ne + 3 * m.nt, 4 * m.nt)
        assert O.orient(t)
        m = t


def test_restriction_closed_form(O):
    m = O.icosphere(2)
    S = O.subdivision_matrix(m)
    R = O.make_restriction(S).dense()
    deg = np.bincount(np.concatenate([m.v0, m.v1]), minlength=m.nv)
    for i in range(m.nv):
        rs = 1 + deg[i] / 2
        assert R[i, i] == 1 / rs
        mids = np.nonzero(R[i, m.nv:])[0]
        assert len(mids) == deg[i] and np.all(R[i, m.nv + mids] == 0.5 / rs)


def test_refinement_orientation_rule(O):
    """Children 4t, 4t+2 keep the parent's orient! orientation, 4t+1, 4t+3 flip:
    the closed form the HIP library uses instead of a depth-first search."""
    for base in (O.icosahedron(), O.triangulated_grid(1, 1, 0.25, math.sqrt(3) / 8)):
        c = O.Mesh(base.p, base.v0, base.v1, base.e0, base.e1, base.e2)
        assert O.orient(c)
        for _ in range(3):
            f = O.subdivide(c)
            assert O.orient(f)
            rule = np.empty(f.nt, np.int8)
            rule[0::4] = c.to
            rule[1::4] = 1 - c.to
            rule[2::4] = c.to
            rule[3::4] = 1 - c.to
            assert np.array_equal(f.to, rule)
            c = f


# --- geometry and Laplacian invariants ---------------------------------------
def test_geometry_invariants(O):
    m = O.icosphere(3)
    d = O.Dual(m)
    nv, ne, nt = m.nv, m.ne, m.nt
    tc = d.dual_point[nv + ne:]
    tv = m.triangle_vertices()
    r = np.linalg.norm(m.p[tv] - tc[:, None, :], axis=2)
    assert np.allclose(r, r[:, :1], rtol=1e-13)                  # circumcenter equidistant
    assert math.isclose(d.hodge_diag(0).sum(), d.area.sum(), rel_tol=1e-13)
    assert np.allclose(d.dual_area.reshape(6, nt).sum(axis=0), d.area, rtol=1e-13)
    # ⋆0⁻¹ · p2_d2 · area = 1 on the sphere (DEC tests 791-800): Σ of a
    # vertex's dual areas is its ⋆0
    assert np.all(d.hodge_diag(0) > 0) and np.all(d.hodge_diag(1) > 0)


def test_laplacian_properties(O):
    d = O.Dual(O.icosphere(5))
    L = d.laplacian0()
    ones = np.ones(d.nv)
    assert np.max(np.abs(L.matvec(ones))) < 1e-9 * np.max(np.abs(L.nzval))
    K = (np.diag(d.hodge_diag(0)) @ L.dense()) if d.nv < 3000 else None
    h0 = d.hodge_diag(0)
    rp, ci, vv = L.csr()
    # K = ⋆0 L = d0ᵀ ⋆1 d0 is symmetric
    Kd = {}
    for r in range(d.nv):
        for k in range(rp[r], rp[r + 1]):
            Kd[(r, ci[k])] = h0[r] * vv[k]
    asym = max(abs(v - Kd[(c, r)]) for (r, c), v in Kd.items())
    assert asym < 1e-12 * max(abs(v) for v in Kd.values())
    del K
    # z = Y_1 restricted to the sphere: ∇² z ≈ l(l+1) z = 2 z (positive sign convention)
    z = d.m.p[:, 2]
    err = np.linalg.norm(L.matvec(z) - 2 * z) / np.linalg.norm(2 * z)
    assert err < 2e-3


# --- multigrid thresholds: test/Multigrid.jl:54-90, 151-170 -------------------
@pytest.fixture(scope="module")
def grid_hierarchy(O):
    base = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, compress=False)
    return O.Hierarchy(base, 4, geometric=True)


def test_multigrid_thresholds(O, grid_hierarchy):
    H = grid_hierarchy
    L = H.A[0]
    b = L.matvec(O.uniform_form(7, L.n) * 0.5 + 0.5)
    u0 = np.zeros(L.n)
    rel = lambda u: np.linalg.norm(L.matvec(u) - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(H.full_multigrid(b, 5, alg=O.CG, mu=2)) < 1e-6            # dec_Δ⁻¹
    assert rel(H.mu_cycles(u0, b, 5, alg=O.CG, mu=1)) < 1e-7
    assert rel(H.mu_cycles(u0, b, 5, alg=O.CG, mu=2)) < 1e-7
    assert rel(H.full_multigrid(b, 5, alg=O.CG, mu=1)) < 1e-6
    assert rel(H.full_multigrid(b, 5, alg=O.CG, mu=2)) < 1e-7
    r0 = rel(H.mu_cycles(u0, b, 5, alg=O.CG, steps=0))
    r1 = rel(H.mu_cycles(u0, b, 5, alg=O.CG, steps=1))
    assert r1 < r0


def test_galerkin_hierarchy_and_thresholds(O, grid_hierarchy):
    """GalerkinMode (test/Multigrid.jl:113-148 with BinarySubdivision): the
    finest operator is the direct one, coarse operators keep the coarse mesh's
    stencil and annihilate constants (R and P preserve them), and the
    reference's residual thresholds hold."""
    base = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, compress=False)
    G = O.Hierarchy(base, 4, geometric=True, galerkin=True)
    D = grid_hierarchy
    assert np.array_equal(G.A[0].csr()[2], D.A[0].csr()[2])
    for k in range(1, 5):
        rp, ci, vv = G.A[k].csr()
        drp, dci, _ = D.A[k].csr()
        assert np.array_equal(rp, drp) and np.array_equal(ci, dci)
        assert np.abs(G.A[k].matvec(np.ones(G.A[k].n))).max() <= 1e-12 * np.abs(vv).max()
    L = G.A[0]
    b = L.matvec(O.uniform_form(9, L.n) * 0.5 + 0.5)
    u0 = np.zeros(L.n)
    rel = lambda u: np.linalg.norm(L.matvec(u) - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(G.mu_cycles(u0, b, 5, alg=O.CG, mu=1)) < 2e-6
    assert rel(G.mu_cycles(u0, b, 5, alg=O.CG, mu=2)) < 7e-7
    assert rel(G.full_multigrid(b, 5, alg=O.CG, mu=1)) < 1e-3
    assert rel(G.full_multigrid(b, 5, alg=O.CG, mu=2)) < 5e-6


def test_gauss_seidel_direct_vcycle(O, grid_hierarchy):
    H = grid_hierarchy
    L = H.A[0]
    b = L.matvec(O.uniform_form(8, L.n))
    rel = lambda u: np.linalg.norm(L.matvec(u) - b) / np.linalg.norm(b)  # noqa: E731
    u = np.zeros(L.n)
    hist = []
    for _ in range(8):
        u = H.mu_cycles(u, b, 1, alg=O.GS, steps=2, coarse_direct=True)
        hist.append(rel(u))
    assert hist[-1] < 1e-7
    rho = (hist[-1] / hist[2]) ** (1 / 5)
    assert rho < 0.4


def test_coarse_inverse_solves_consistent_systems(O):
    d = O.Dual(O.icosphere(1))
    L = d.laplacian0()
    Z = O.coarse_inverse(L)
    x = O.uniform_form(3, L.n)
    x -= x.mean()
    b = L.matvec(x)
    z = Z @ b
    assert np.linalg.norm(L.matvec(z) - b) < 1e-12 * np.linalg.norm(b)
    assert abs(z.mean()) < 1e-14


def _constant_primal_form(d, X):
    """eval_constant_primal_form (src/DiscreteExteriorCalculus.jl:2179-2184):
    X · (p[∂v0] - p[∂v1]) times the edge's orientation sign."""
    m = d.m
    sign = np.where(d.eo > 0, 1.0, -1.0)
    return (m.p[m.v0] - m.p[m.v1]) @ np.asarray(X, float) * sign


@pytest.mark.parametrize("compress", [False, True])
def test_primal_dual_wedge_reference_cases(O, compress):
    """test/Operators.jl:361-414 on flat triangulated grids: a constant 0-form
    g = 5 scales ⋆f, Λ10(f, g) == Λ01(g, f), and for a = x + 4y,
    Λ10(⋆da, a) matches ⋆(-h ∧ dx + g ∧ dy) on interior edges (atol 1e-10),
    with g = 4x + 16y, h = -x - 4y."""
    dx = 1 / 8
    grid = O.triangulated_grid(1, 1, dx, math.sqrt(3) / 2 * dx, compress=compress)
    d = O.Dual(grid)
    H1 = d.geo_hodge1()
    s2 = 1 / math.sqrt(2)
    f = H1.matvec(_constant_primal_form(d, (s2, s2, 0)))
    g = np.full(d.nv, 5.0)
    l10, l01 = d.wedge_dp10(f, g), d.wedge_pd01(g, f)
    ref = H1.matvec(_constant_primal_form(d, (5 * s2, 5 * s2, 0)))
    assert np.allclose(l10, ref, rtol=1.5e-8, atol=0)
    assert np.array_equal(l10, l01)
    # linear case on interior edges
    m = d.m
    ntri = np.bincount(np.concatenate([m.e0, m.e1, m.e2]), minlength=d.ne)
    interior = ntri == 2
    assert interior.any()
    x, y = m.p[:, 0], m.p[:, 1]
    a = x + 4 * y
    f = H1.matvec(d.differential(0).matvec(a))
    gg, hh = 4 * x + 16 * y, -x - 4 * y
    DX = _constant_primal_form(d, (1, 0, 0))
    DY = _constant_primal_form(d, (0, 1, 0))
    analytic = H1.matvec(-d.wedge01(hh, DX) + d.wedge01(gg, DY))
    got = d.wedge_dp10(f, a)
    assert np.max(np.abs(got[interior] - analytic[interior])) <= 1e-10


def test_primal_dual_wedge_matrix_structure(O):
    """Rows of wedge_pd_01_mat sum to one (the weights of an edge's elementary
    duals sum to one, 5/12 + 5/12 + 2/12 = 1); interior rows hold four
    entries, boundary rows three, all positive."""
    for mesh in (O.icosphere(2), O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 8, compress=False)):
        d = O.Dual(mesh)
        rp, ci, vv = d.wedge_pd01_mat().csr()
        rs = np.add.reduceat(vv, rp[:-1])
        assert np.max(np.abs(rs - 1)) <= 4e-16
        m = d.m
        ntri = np.bincount(np.concatenate([m.e0, m.e1, m.e2]), minlength=d.ne)
        assert np.array_equal(np.diff(rp), np.where(ntri == 2, 4, 3))
        assert (vv > 0).all()
