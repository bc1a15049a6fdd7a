"""♭♯, the primal-dual / dual-primal 1-1 wedges, the dual-dual interior
products, ℒ_dd and the dual Laplacians of the HIP library against the numpy
oracle (oracle/musical.py) on flat meshes (the reference's test meshes) and
icospheres.  The oracle's ♯ is taken in the triangle's plane, which is the
reference's pseudo-inverse on flat meshes (pinned in test_oracle_musical.py)
and the library's definition on curved ones.  FP64 results agree to 1e-12
normwise (relative to the largest entry); FP32 results to 1e-4."""
import math

import numpy as np
import pytest

from test_gpu_ops import F64_TOL, gpu_mesh_like, rel, sparse_close, to_dev

pytestmark = pytest.mark.gpu

F32_CHAIN_TOL = 1e-4
CASES = ["tri_345", "grid_345", "tg", "ico3", "ico3_bary"]


@pytest.fixture(scope="module", params=CASES)
def case(request, O, cs):
    import musical as MU
    from test_oracle_musical import flat_primal_meshes

    name = request.param
    if name.startswith("ico"):
        om = O.icosphere(3)
        circ = not name.endswith("bary")
    else:
        om = flat_primal_meshes(O)[name]
        circ = False
    od = O.Dual(om, circumcenter=circ)
    gm = gpu_mesh_like(cs, om)
    cs.dualize(gm, cs.Circumcenter if circ else cs.Barycenter)
    return name, od, gm, MU.Ops(od, project=True)


def dev(torch, x, td):
    return to_dev(torch, np.asarray(x, float), td)


def check(torch, got, ref, dtype):
    tol = F64_TOL if dtype == "f64" else F32_CHAIN_TOL
    err = rel(got.cpu().numpy(), ref)
    assert err <= tol, err


def test_flat_sharp_matrix(case, O, cs):
    import torch

    name, od, gm, ops = case
    fs = cs.flat_sharp_mat(gm, cast_float=torch.float64)
    R = ops.FS.tocsr()
    R.sort_indices()
    assert sparse_close(fs.csr(), (R.indptr, R.indices, R.data), od.ne)
    x = O.uniform_form(70, od.ne)
    check(torch, fs(dev(torch, x, torch.float64)), ops.FS @ x, "f64")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_wedge_pd_dp_11(case, O, cs, dtype):
    import torch

    name, od, gm, ops = case
    td = torch.float64 if dtype == "f64" else torch.float32
    pd = cs.dec_wedge_product_pd(1, 1, gm, cast_float=td)
    dp = cs.dec_wedge_product_dp(1, 1, gm, cast_float=td)
    f, g = O.uniform_form(71, od.ne), O.uniform_form(72, od.ne)
    a = pd(dev(torch, f, td), dev(torch, g, td))
    b = dp(dev(torch, g, td), dev(torch, f, td))
    check(torch, a, ops.wedge_pd11(f, g), dtype)
    check(torch, b, ops.wedge_dp11(g, f), dtype)
    assert torch.equal(a, -b)                      # Λpd(f, g̃) == -Λdp(g̃, f) exactly


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_interior_products_and_lie_derivative(case, O, cs, dtype):
    import torch

    name, od, gm, ops = case
    td = torch.float64 if dtype == "f64" else torch.float32
    f, g = O.uniform_form(73, od.ne), O.uniform_form(74, od.ne)
    g2 = O.uniform_form(75, od.nv)
    i11 = cs.interior_product_dd(1, 1, gm, cast_float=td)
    i12 = cs.interior_product_dd(1, 2, gm, cast_float=td)
    lie = cs.lie_derivative_dd(1, 1, gm, cast_float=td)
    check(torch, i11(dev(torch, f, td), dev(torch, g, td)), ops.interior11(f, g), dtype)
    check(torch, i12(dev(torch, f, td), dev(torch, g2, td)), ops.interior12(f, g2), dtype)
    check(torch, lie(dev(torch, f, td), dev(torch, g, td)), ops.lie(f, g), dtype)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_dual_laplacians(case, O, cs, dtype):
    import torch

    name, od, gm, ops = case
    td = torch.float64 if dtype == "f64" else torch.float32
    x0, x1 = O.uniform_form(76, od.nt), O.uniform_form(77, od.ne)
    L0 = cs.dual_laplacian(0, gm, cast_float=td)
    L1 = cs.dual_laplacian(1, gm, cast_float=td)
    check(torch, L0 @ dev(torch, x0, td), ops.dual_laplacian0(x0), dtype)
    check(torch, L1 @ dev(torch, x1, td), ops.dual_laplacian1(x1), dtype)


def test_constant_flow_on_the_grid(O, cs):
    """The reference's interior-product property (test/Operators.jl:441-461)
    through the library: d̃0 ι11(u, u) = 0 on interior edges of the grid."""
    import musical as MU
    import torch
    from test_oracle_musical import flat_primal_meshes, interior_edges

    om = flat_primal_meshes(O)["tg"]
    od = O.Dual(om, circumcenter=False)
    gm = gpu_mesh_like(cs, om)
    cs.dualize(gm, cs.Barycenter)
    u = MU.Ops(od).H1 @ MU.eval_constant_primal_form(od, np.array([1 / math.sqrt(2), 1 / math.sqrt(2), 0]))
    ud = dev(torch, u, torch.float64)
    i = cs.interior_product_dd(1, 1, gm, cast_float=torch.float64)(ud, ud)
    di = cs.dec_dual_derivative(0, gm, cast_float=torch.float64) @ i
    assert np.abs(di.cpu().numpy()[interior_edges(od)]).max() <= 1e-13


def test_composites_are_refused_by_apply_many(cs):
    import torch

    gm = cs.icosphere(2)
    cs.dualize(gm)
    lie = cs.lie_derivative_dd(1, 1, gm, cast_float=torch.float64).operator
    x = torch.zeros(lie.cols, dtype=torch.float64, device="cuda")
    with pytest.raises(cs.DecError):
        cs.apply_many([(lie, x, x, torch.empty_like(x))])
    with pytest.raises(cs.DecError):
        cs.interior_product_dd(2, 1, gm)
    with pytest.raises(cs.DecError):
        cs.dec_wedge_product_pd(1, 0, gm)
    from cspaces import _lib as L
    with pytest.raises(cs.DecError, match="primal-dual wedges are 0-1, 1-0, and 1-1"):
        cs.Operator(gm, L.DEC_OP_WEDGE_PD, 1, 0)          # the library refuses it too
    with pytest.raises(cs.DecError, match="dual-primal wedges are 0-1, 1-0, and 1-1"):
        cs.Operator(gm, L.DEC_OP_WEDGE_DP, 0, 1)
