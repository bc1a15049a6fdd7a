"""makeSphere (src/CombMeshes.jl:238-342): the oracle's call-by-call
restatement (oracle.make_sphere) pinned by the reference's own tests
(test/CombMeshes.jl:155-170: every point at distance ρ; the 90-degree sphere is
the regular octahedron, 6 / 12 / 8) and by what the construction fixes (Euler
characteristic 2 / 1 / 0 for sphere, cap and band; every edge in two triangles
on the closed sphere, at most two elsewhere; triangle_vertices ascending, as
glue_sorted_triangle! leaves them; an orientable surface).  The library's
vectorised generator (cspaces.make_sphere_arrays) is compared with it index
for index; the GPU operator tests run on the sphere of test/Operators.jl:60
and test/Backends.jl:53, makeSphere(5, 175, 5, 0, 360, 5, 6371+90)."""
import math

import numpy as np
import pytest

PARAMS = [(0, 180, 90, 0, 360, 90, 1), (5, 175, 5, 0, 360, 5, 6371 + 90), (0, 180, 5, 0, 360, 5, 6371 + 90),
          (70, 110, 10, 0, 360, 20, 1), (0, 90, 20, 0, 360, 20, 1), (10, 100, 15, 30, 300, 25, 2.5),
          (90, 90, 0, 0, 360, 30, 1)]


def _edge_use(m):
    cnt = np.zeros(m.ne, int)
    for e in (m.e0, m.e1, m.e2):
        np.add.at(cnt, e, 1)
    return cnt


def test_reference_kats(O):
    rho = 6371 + 90
    s, npi, spi = O.make_sphere(0, 180, 5, 0, 360, 5, rho)
    r = np.linalg.norm(s.p, axis=1)
    np.testing.assert_allclose(r, rho, rtol=math.sqrt(np.finfo(float).eps))
    np.testing.assert_allclose(r, r[0], rtol=math.sqrt(np.finfo(float).eps))
    octa, npi, spi = O.make_sphere(0, 180, 90, 0, 360, 90, 1)
    assert (octa.nv, octa.ne, octa.nt) == (6, 12, 8)
    assert (npi, spi) == (4, 5)
    assert sorted(map(tuple, octa.p.tolist())) == sorted(
        [(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])


@pytest.mark.parametrize("args", PARAMS)
def test_topology(O, args):
    m, npi, spi = O.make_sphere(*args)
    use = _edge_use(m)
    chi = m.nv - m.ne + m.nt
    closed = args[0] == 0 and args[1] == 180 and args[4] == 360
    if m.nt == 0:
        assert m.nv == m.ne                              # one closed parallel
        return
    assert use.max() <= 2 and use.min() >= 1
    if closed:
        assert chi == 2 and (use == 2).all()
    elif args[4] == 360:
        assert chi == (1 if (args[0] == 0) != (args[1] == 180) else 0)
    tv = m.triangle_vertices()
    assert (np.diff(tv, axis=1) > 0).all()
    assert O.orient(m)


def test_closed_sphere_area_converges(O):
    rho = 2.0
    for d, tol in ((10, 0.03), (5, 0.008)):
        m, _, _ = O.make_sphere(0, 180, d, 0, 360, d, rho)
        a = m.p[m.triangle_vertices()]
        area = 0.5 * np.linalg.norm(np.cross(a[:, 1] - a[:, 0], a[:, 2] - a[:, 0]), axis=1).sum()
        assert 0 < 4 * math.pi * rho ** 2 - area < tol * 4 * math.pi * rho ** 2


@pytest.mark.parametrize("args", PARAMS)
def test_library_generator_matches_the_oracle(O, args):
    from cspaces import fastdec

    m, npi, spi = O.make_sphere(*args)
    p, v0, v1, e0, e1, e2, n2, s2 = fastdec.make_sphere_arrays(*args)
    assert np.array_equal(p, m.p)
    for a, b in ((v0, m.v0), (v1, m.v1), (e0, m.e0), (e1, m.e1), (e2, m.e2)):
        assert np.array_equal(a, b)
    assert (n2, s2) == (npi, spi)


def test_argument_errors(O):
    from cspaces import fastdec

    for bad in ((10, 5, 5, 0, 360, 5, 1), (0, 190, 5, 0, 360, 5, 1), (0, 180, 5, 90, 30, 5, 1)):
        with pytest.raises(ValueError):
            O.make_sphere(*bad)
        with pytest.raises(ValueError):
            fastdec.make_sphere_arrays(*bad)
