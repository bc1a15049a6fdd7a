"""The oracle's musical isomorphisms and the operators built on them
(oracle/musical.py), pinned without a GPU by the reference's own tests and by
what the mathematics fixes:

* ♯(LLSDDSharp) and ♭♯ reproduce constant fields and forms on flat meshes
  (test/DiscreteExteriorCalculus.jl:702-725, atol 1e-13);
* the primal-dual / dual-primal 1-1 wedges are antisymmetric, exactly so
  across pd and dp, and f ∧ f vanishes (DiscreteExteriorCalculus.jl:727-749);
* d̃0 ι11(u, u) vanishes on interior edges for a constant flow
  (test/Operators.jl:441-461, atol 1e-13);
* ι11(u, u) of u = ⋆(x dx) converges to x² at second order, and passes the
  reference's signed bound (Operators.jl:563-575);
* the self-advection term -ℒ_dd(u, u) + ½ d̃0 ι11(u, u) of a constant flow on
  the 100x100 grid meets the reference's thresholds (Operators.jl:532-545);
* Δᵈ on the icosphere has the spectrum of the Laplace-Beltrami operator:
  -l(l+1) with multiplicity 2l+1 on dual 0-forms, 2(2l+1) on dual 1-forms,
  and annihilates constant dual 0-forms.
"""
import math

import numpy as np
import pytest


def flat_primal_meshes(O):
    """tri_345, right_scalene_unit_hypot, grid_345 (src/CombMeshes.jl:348-413)
    and the 100x100 grid of test/Operators.jl:65-67, all with barycentric duals."""
    from test_gpu_ops import grid345_arrays

    r = 1 / math.sqrt(2)
    tri = O.glue_triangles(np.array([[0, 0, 0], [3, 0, 0], [3, 4, 0]], float), [[0, 1, 2]])
    tri.eo[:] = 1
    tri.to[:] = 1
    rsu = O.glue_triangles(np.array([[0, 0, 0], [r, 0, 0], [r, r, 0]], float), [[0, 1, 2]], True)
    rsu.eo[:] = 1
    return {"tri_345": tri, "right_scalene_unit_hypot": rsu, "grid_345": grid345_arrays(O),
            "tg": O.triangulated_grid(100, 100, 10, 10, True)}


def _flat_meshes(O):
    return {k: O.Dual(m, circumcenter=False) for k, m in flat_primal_meshes(O).items()}


@pytest.fixture(scope="module")
def flat(O):
    return _flat_meshes(O)


def interior_edges(d):
    cnt = np.zeros(d.ne, int)
    for e in (d.m.e0, d.m.e1, d.m.e2):
        np.add.at(cnt, e, 1)
    return np.flatnonzero(cnt == 2)


def interior_tris(d):
    """setdiff(triangles, boundary_inds(Val(2), s)) (SimplicialSets.jl:1097-1107)."""
    bnd = np.setdiff1d(np.arange(d.ne), interior_edges(d))
    E = np.stack([d.m.e0, d.m.e1, d.m.e2], 1)
    return np.flatnonzero(~np.isin(E, bnd).any(1))


def test_sharp_and_flat_sharp_reproduce_constant_fields(flat):
    import musical as MU

    Xs = np.array([3 / math.sqrt(2), 1 / math.sqrt(2), 0])
    for name, d in flat.items():
        Xb = MU.eval_constant_dual_form(d, Xs)
        E, vec = MU.sharp_lls(d)
        sharp = np.einsum("tik,ti->tk", vec, Xb[E])
        assert min(np.abs(sharp - Xs).max(), np.abs(sharp + Xs).max()) <= 1e-13, name
        pf = MU.eval_constant_primal_form(d, Xs)
        fs = MU.flat_sharp_mat(d) @ Xb
        assert min(np.abs(fs - pf).max(), np.abs(fs + pf).max()) <= 1e-13, name
        # on a flat mesh the pseudo-inverse in the triangle's plane is the literal one
        assert abs(MU.flat_sharp_mat(d, project=True) - MU.flat_sharp_mat(d)).max() <= 1e-15, name


def test_primal_dual_wedge_antisymmetry(flat):
    import musical as MU

    fd, gd = np.array([2.0, 7, 0]), np.array([8.0, 1, 0])
    for name, d in flat.items():
        ops = MU.Ops(d)
        f_, g_ = MU.eval_constant_primal_form(d, fd), MU.eval_constant_primal_form(d, gd)
        ft, gt = MU.eval_constant_dual_form(d, fd), MU.eval_constant_dual_form(d, gd)
        np.testing.assert_allclose(ops.wedge_pd11(f_, gt), -ops.wedge_pd11(g_, ft), rtol=1.5e-8, atol=1e-11)
        np.testing.assert_allclose(ops.wedge_dp11(ft, g_), -ops.wedge_dp11(gt, f_), rtol=1.5e-8, atol=1e-11)
        assert np.array_equal(ops.wedge_pd11(f_, gt), -ops.wedge_dp11(gt, f_)), name
        assert np.array_equal(ops.wedge_dp11(ft, g_), -ops.wedge_pd11(g_, ft)), name
        assert np.abs(ops.wedge_pd11(f_, ft)).max() <= 1e-11, name
        assert np.abs(ops.wedge_pd11(g_, gt)).max() <= 1e-11, name


def test_interior_product_of_constant_flow_is_closed(flat):
    import musical as MU

    for name, d in flat.items():
        inter = interior_edges(d)
        if len(inter) == 0:
            continue
        ops = MU.Ops(d)
        u = ops.H1 @ MU.eval_constant_primal_form(d, np.array([1 / math.sqrt(2), 1 / math.sqrt(2), 0]))
        i = ops.interior11(u, u)
        assert np.abs(np.abs(i) - 1.0).max() <= 1e-12, name         # |u|² = 1
        assert np.abs((ops.dd0 @ i)[inter]).max() <= 1e-13, name


def test_interior_product_of_x_dx_converges_to_x_squared(O):
    import musical as MU

    errs = []
    for h in (1.0, 0.5):
        d = O.Dual(O.triangulated_grid(30, 10, h, h * math.sqrt(3) / 2, True), circumcenter=False)
        ops = MU.Ops(d)
        x = d.m.p[:, 0]
        dx = MU.eval_constant_primal_form(d, np.array([1.0, 0, 0]))
        u = ops.H1 @ d.wedge01(x, dx)
        i = ops.interior11(u, u)[interior_tris(d)]
        xc = d.dual_point[d.nv + d.ne + interior_tris(d), 0]
        assert np.all(i - xc ** 2 < 8e-3)                            # the reference's (signed) bound
        errs.append(np.abs(np.abs(i) - xc ** 2).max())
    assert errs[1] < 0.02 and errs[0] / errs[1] > 3.0, errs          # second order in h


def test_self_advection_of_constant_flow(flat):
    import musical as MU

    d = flat["tg"]
    ops = MU.Ops(d)
    it = interior_tris(d)
    assert len(it) == 162
    for Xs, thr, need in (((1 / math.sqrt(2), 1 / math.sqrt(2), 0), 1e-2, 0.64), ((3, 3, 0), 1e-1, 97 / 162)):
        u = ops.H1 @ MU.eval_constant_primal_form(d, np.array(Xs, float))
        selfadv = ops.lie(u, u) - 0.5 * (ops.dd0 @ ops.interior11(u, u))
        mag = np.sqrt(np.abs(ops.interior11(selfadv, selfadv)))[it]
        assert (mag < thr).mean() >= need, (Xs, (mag < thr).mean())


def _dense(op, n):
    return np.stack([op(np.eye(n)[i]) for i in range(n)], 1)


def test_dual_laplacians_have_the_sphere_spectrum(O):
    import musical as MU

    d = O.Dual(O.icosphere(2))
    ops = MU.Ops(d, project=True)
    assert np.abs(ops.dual_laplacian0(np.ones(d.nt))).max() <= 1e-12
    ev0 = np.sort(np.linalg.eigvals(_dense(ops.dual_laplacian0, d.nt)).real)[::-1]
    assert abs(ev0[0]) <= 1e-10
    np.testing.assert_allclose(ev0[1:4], -2.0, rtol=0.03)
    np.testing.assert_allclose(ev0[4:9], -6.0, rtol=0.03)
    assert ev0[9] < -11.0
    ev1 = np.sort(np.linalg.eigvals(_dense(ops.dual_laplacian1, d.ne)).real)[::-1]
    np.testing.assert_allclose(ev1[:6], -2.0, rtol=0.03)
    np.testing.assert_allclose(ev1[6:16], -6.0, rtol=0.03)
    assert ev1[16] < -11.0


def test_literal_pinv_sharp_is_rounding_noise_on_spheres(O):
    """The reference's ♯ takes LinearAlgebra.pinv (default cut 3 eps σmax) of
    the 3x3 matrix of dual-edge vectors.  On a curved surface those rows lie
    in the triangle's plane up to rounding, so the third singular value is
    rounding noise above the cut and gets inverted: ♭♯ built that way moves by
    O(1) when the dual points move by one unit in the last place, so no
    implementation can reproduce the reference's values there.  The in-plane
    pseudo-inverse (the library's ♯) moves by rounding only, and both forms
    agree on flat meshes (the meshes of the reference's tests)."""
    import musical as M

    d = O.Dual(O.icosphere(3))
    lit, proj = M.flat_sharp_mat(d, project=False), M.flat_sharp_mat(d, project=True)
    d2 = O.Dual(O.icosphere(3))
    rng = np.random.default_rng(0)
    d2.dual_point = d2.dual_point * (1 + rng.choice([-1.0, 1.0], d2.dual_point.shape) * 2.0 ** -52)
    lit2, proj2 = M.flat_sharp_mat(d2, project=False), M.flat_sharp_mat(d2, project=True)
    rel = lambda a, b: abs(a - b).max() / abs(b).max()  # noqa: E731
    assert rel(lit2, lit) > 0.5                 # one ulp in the points: an O(1) change
    assert rel(proj2, proj) < 1e-12             # the in-plane form is well conditioned
    assert rel(lit, proj) > 0.5
