"""The learned-smoother oracle (oracle/learned.py) pinned without a GPU: the
untrained network is damped Jacobi, the correction scales with the residual,
the high-pass filter of the error loss removes what the coarse grid
represents, and both analytic gradients agree with central finite
differences on every parameter group."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def level(O):
    import learned as LO

    H = O.Hierarchy(O.icosahedron(), 2, project=True)       # 162 -> 42 -> 12 vertices
    A = H.A[0].csr()
    R = H.R[0].csr()
    P = H.S[0].transpose().csr()
    n, nc = H.A[0].m, H.A[1].m
    return LO.Level(*A, R=LO.Csr(*R, ncols=n), Pr=LO.Csr(*P, ncols=nc)), H


def random_theta(LO, seed):
    rng = np.random.default_rng(seed)
    return LO.init_params(seed) + 0.1 * rng.standard_normal(LO.n_params())


def test_untrained_network_is_damped_jacobi(O, level):
    import learned as LO

    lv, _ = level
    r = O.uniform_form(3, lv.n)
    c, _ = LO.forward(LO.init_params(7), lv, r)
    np.testing.assert_allclose(c, (2 / 3) * r / lv.D, rtol=1e-14, atol=0)


def test_correction_scales_with_the_residual(O, level):
    import learned as LO

    lv, _ = level
    th = random_theta(LO, 1)
    r = O.uniform_form(4, lv.n)
    c1, _ = LO.forward(th, lv, r)
    c2, _ = LO.forward(th, lv, 8.0 * r)
    np.testing.assert_allclose(c2, 8.0 * c1, rtol=1e-13, atol=1e-15)


def test_high_pass_removes_prolongated_vectors(O, level):
    """H = I - Pr R annihilates nothing in general, but H Pr z = Pr (I - R Pr) z,
    and for the constants (R and Pr preserve them) H 1 = 0."""
    import learned as LO

    lv, H = level
    one = np.ones(lv.n)
    np.testing.assert_allclose(LO.high_pass(lv, one), 0.0, atol=1e-14)
    z = O.uniform_form(5, H.A[1].m)
    np.testing.assert_allclose(LO.high_pass(lv, lv.Pr @ z), lv.Pr @ (z - lv.R @ (lv.Pr @ z)), atol=1e-14)


@pytest.mark.parametrize("kind", ["residual", "error"])
def test_gradient_matches_finite_differences(O, level, kind):
    import learned as LO

    lv, _ = level
    th = random_theta(LO, 2)
    Rs = np.stack([O.uniform_form(10 + i, lv.n) for i in range(3)])
    Es = np.stack([O.uniform_form(20 + i, lv.n) for i in range(3)])
    if kind == "residual":
        f, g = LO.loss_grad(th, lv, Rs)
        assert f == pytest.approx(LO.loss(th, lv, Rs), rel=1e-14)
        E = None
    else:
        f, g = LO.loss_err_grad(th, lv, Rs, Es)
        assert f == pytest.approx(LO.loss_err(th, lv, Rs, Es), rel=1e-14)
        E = Es
    # every parameter group: a spread of indices in each block
    idx, k = [], 0
    for _, shp in LO.SHAPES:
        m = int(np.prod(shp))
        idx += list(range(k, k + m, max(1, m // 12)))
        k += m
    fd = LO.fd_grad(th, lv, Rs, idx, E=E)
    err = np.abs(fd - g[idx]).max() / np.abs(g[idx]).max()
    assert err < 1e-6, err
