"""The oracle's interpolation matrices (oracle/interp.py, oracle/tet.py) pinned
without a GPU by the reference's own tests (test/DiscreteExteriorCalculus.jl:
780-888) and by what the mathematics fixes:

* constant interpolation: ⋆0⁻¹ p2_d2 1 = nt / Σ area on flat meshes of equal
  areas, ⋆0⁻¹ p2_d2 area = 1 on the icosphere (:780-798);
* d0_p0 on triangulated_grid(30, 10, 0.5, 0.5) with circumcentric duals: its
  pattern is the vertex-triangle incidence |d̃1| |d̃0| (d̃1 d̃0 = 0), it
  reproduces linear functions at interior vertices to rmse 1e-14, a quadratic
  to relative rmse 6e-4 (:800-857);
* p0_d0 on the same grid: linear in x to rmse 3e-13, linear in y to 0.07; x on
  the icosphere k=4 to 0.005 (:859-888);
* p0_d0 = p2_d2ᵀ exactly, and on tetrahedra ⋆0⁻¹ p3_d3 vol = 1 and d0_p0
  reproduces linear functions at the interior vertices of a Kuhn box (its
  vertex stars are centrally symmetric) and constants everywhere.
"""
import math

import numpy as np
import pytest


def rmse(v):
    v = np.asarray(v)
    return math.sqrt(np.sum(v ** 2) / len(v))


@pytest.fixture(scope="module")
def I():
    import interp

    return interp


def test_constant_interpolation_on_flat_meshes(O, I):
    from test_oracle_musical import flat_primal_meshes

    for name, m in flat_primal_meshes(O).items():
        d = O.Dual(m, circumcenter=False)
        res = d.inv_hodge_diag(0) * (I.p2_d2_interpolation(d) @ np.ones(d.nt))
        np.testing.assert_allclose(res, d.nt / d.area.sum(), rtol=1e-12, err_msg=name)


def test_constant_interpolation_on_the_icosphere(O, I):
    d = O.Dual(O.icosphere(1), circumcenter=True)
    res = d.inv_hodge_diag(0) * (I.p2_d2_interpolation(d) @ d.area)
    np.testing.assert_allclose(res, 1.0, rtol=1e-13)


@pytest.fixture(scope="module")
def rect(O):
    m = O.triangulated_grid(30, 10, 0.5, 0.5, True)
    return m, O.Dual(m, circumcenter=True)


def _boundary_vertices(d):
    cnt = np.zeros(d.ne, int)
    for e in (d.m.e0, d.m.e1, d.m.e2):
        np.add.at(cnt, e, 1)
    be = np.flatnonzero(cnt == 1)
    return np.union1d(d.m.v0[be], d.m.v1[be])


def test_d0_p0_on_the_rectangle(O, I, rect):
    m, d = rect
    from musical import _csr

    M = I.d0_p0_interpolation(d).tocsr()
    D0, D1 = _csr(d.dual_derivative(0)), _csr(d.dual_derivative(1))
    assert abs(D1 @ D0).max() == 0
    pattern = (abs(D1) @ abs(D0)).tocsr()
    assert M.shape == pattern.shape == (d.nv, d.nt)
    assert (M != 0).sum() == pattern.nnz and ((M != 0).astype(int) - (pattern != 0).astype(int)).nnz == 0
    P, C = m.p, d.dual_point[d.nv + d.ne:]
    ly, lx, dx = 10, 30, 0.5
    y_in = np.flatnonzero((P[:, 1] != 0) & (P[:, 1] != ly))
    x_in = np.flatnonzero((P[:, 0] > dx) & (P[:, 0] < lx - dx))
    inner = np.setdiff1d(np.arange(d.nv), _boundary_vertices(d))
    for f, idx in ((lambda p: 2 * p[:, 1], y_in), (lambda p: 2 * p[:, 0], x_in),
                   (lambda p: 2 * p[:, 0] + 3 * p[:, 1], inner)):
        assert rmse((f(P) - M @ f(C))[idx]) < 1e-14
    q = lambda p: 0.1 * (p[:, 0] - lx / 2.0) ** 2 + 5
    assert rmse(((q(P) - M @ q(C)) / q(P))[inner]) < 0.0006


def test_p0_d0_on_the_rectangle_and_sphere(O, I, rect):
    m, d = rect
    M = I.p0_d0_interpolation(d)
    assert abs(M - I.p2_d2_interpolation(d).T).max() == 0
    P, C = m.p, d.dual_point[d.nv + d.ne:]
    assert rmse(2 * C[:, 0] - M @ (2 * P[:, 0])) < 3e-13
    assert rmse(2 * C[:, 1] - M @ (2 * P[:, 1])) < 0.07
    s = O.Dual(O.icosphere(4), circumcenter=True)
    Ms = I.p0_d0_interpolation(s)
    assert rmse(s.dual_point[s.nv + s.ne:, 0] - Ms @ s.m.p[:, 0]) < 0.005


def test_tetrahedral_interpolation():
    import tet

    d = tet.TetDual(tet.kuhn_grid(4, 4, 4), circumcenter_=False)
    P3 = d.p3_d3_interpolation()
    np.testing.assert_allclose(d.inv_hodge_diag(0) * (P3 @ d.vol), 1.0, rtol=1e-14)
    D = d.d0_p0_interpolation()
    np.testing.assert_allclose(D @ np.ones(d.ntet), 1.0, rtol=1e-14)
    f = lambda p: 2 * p[:, 0] + 3 * p[:, 1] - p[:, 2]
    inner = np.all((d.m.p > 0) & (d.m.p < 1), axis=1)
    err = D @ f(d.dual_point[d.tetc]) - f(d.m.p)
    assert np.abs(err[inner]).max() < 1e-14
    assert np.abs(err[~inner]).max() > 0.1             # boundary stars are one-sided
