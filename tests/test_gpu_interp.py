"""The interpolation operators of the HIP library (p2_d2, p0_d0, d0_p0 on
triangle complexes; p3_d3, d0_p0 and the skeleton's p0_d0 on tetrahedra)
against the oracle (oracle/interp.py, oracle/tet.py): matrices entry-wise to
1e-12 of their largest entry, products to 1e-12 in FP64 and 1e-5 in FP32;
and the reference's linear-reproduction test run on the library."""
import math

import numpy as np
import pytest

from test_gpu_ops import F32_TOL, F64_TOL, gpu_mesh_like, rel, sparse_close, to_dev
from test_gpu_tet import csr3, gpu_mesh3_like

pytestmark = pytest.mark.gpu

CASES2 = ["tri_345", "grid_345", "tg", "ico3", "ico3_bary", "rect", "uvsphere_bary"]
CASES3 = ["single", "kuhn_bary", "kuhn_jitter_circ"]


@pytest.fixture(scope="module", params=CASES2)
def case2(request, O, cs):
    from test_oracle_musical import flat_primal_meshes

    name = request.param
    if name.startswith("ico"):
        om, circ = O.icosphere(3), not name.endswith("bary")
    elif name == "rect":
        om, circ = O.triangulated_grid(30, 10, 0.5, 0.5, True), True
    elif name == "uvsphere_bary":
        om, circ = O.make_sphere(5, 175, 5, 0, 360, 5, 6371 + 90)[0], False      # test/Operators.jl:60
    else:
        om, circ = flat_primal_meshes(O)[name], False
    od = O.Dual(om, circumcenter=circ)
    gm = gpu_mesh_like(cs, om)
    cs.dualize(gm, cs.Circumcenter if circ else cs.Barycenter)
    return name, od, gm


@pytest.fixture(scope="module", params=CASES3)
def case3(request, cs):
    import tet

    name = request.param
    if name == "single":
        m, circ = tet.single_tetrahedron(), False
    elif name == "kuhn_bary":
        m, circ = tet.kuhn_grid(3, 3, 2, jitter=0.2, seed=7), False
    else:
        m, circ = tet.kuhn_grid(3, 3, 3, jitter=0.15, seed=2, jitter_boundary=True), True
    gm = gpu_mesh3_like(cs, m)
    od = tet.TetDual(m, circumcenter_=circ)
    cs.dualize(gm, cs.Circumcenter if circ else cs.Barycenter)
    return name, od, gm


def check_op(torch, op, ref, n, dtype):
    assert sparse_close(op.csr(), csr3(ref), n)
    x = np.random.default_rng(n).uniform(-1, 1, op.cols)
    t = torch.float64 if dtype == "f64" else torch.float32
    got = (op @ to_dev(torch, x, t)).cpu().numpy()
    assert rel(got, ref @ x) <= (F64_TOL if dtype == "f64" else F32_TOL)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_interpolation_2d(case2, cs, dtype):
    import interp
    import torch

    name, od, gm = case2
    t = torch.float64 if dtype == "f64" else torch.float32
    check_op(torch, cs.p2_d2_interpolation(gm, cast_float=t), interp.p2_d2_interpolation(od), od.nv, dtype)
    check_op(torch, cs.p0_d0_interpolation(gm, cast_float=t), interp.p0_d0_interpolation(od), od.nt, dtype)
    for h in (cs.GeometricHodge, cs.DiagonalHodge):
        check_op(torch, cs.d0_p0_interpolation(gm, hodge=h, cast_float=t), interp.d0_p0_interpolation(od),
                 od.nv, dtype)
    with pytest.raises(cs.DecError):
        cs.p3_d3_interpolation(gm)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_interpolation_3d(case3, cs, dtype):
    import interp
    import torch

    name, od, gm = case3
    t = torch.float64 if dtype == "f64" else torch.float32
    check_op(torch, cs.p3_d3_interpolation(gm, cast_float=t), od.p3_d3_interpolation(), od.nv, dtype)
    check_op(torch, cs.d0_p0_interpolation(gm, cast_float=t), od.d0_p0_interpolation(), od.nv, dtype)
    check_op(torch, cs.p0_d0_interpolation(gm, cast_float=t), interp.p0_d0_interpolation(od), od.nt, dtype)
    with pytest.raises(cs.DecError):
        cs.p2_d2_interpolation(gm)


def test_d0_p0_reproduces_linear_functions_on_the_library(O, cs):
    """test/DiscreteExteriorCalculus.jl:800-840 on the library's matrix."""
    import torch

    m = O.triangulated_grid(30, 10, 0.5, 0.5, True)
    gm = gpu_mesh_like(cs, m)
    cs.dualize(gm, cs.Circumcenter)
    M = cs.d0_p0_interpolation(gm)
    C = gm.export_dual()["dual_point"][gm.nv + gm.ne:]
    P = m.p
    cnt = np.zeros(gm.ne, int)
    for e in (m.e0, m.e1, m.e2):
        np.add.at(cnt, e, 1)
    be = np.flatnonzero(cnt == 1)
    inner = np.setdiff1d(np.arange(gm.nv), np.union1d(m.v0[be], m.v1[be]))
    f = lambda p: 2 * p[:, 0] + 3 * p[:, 1]  # noqa: E731
    got = (M @ to_dev(torch, f(C), torch.float64)).cpu().numpy()
    err = (f(P) - got)[inner]
    assert math.sqrt(np.sum(err ** 2) / len(err)) < 1e-14
