"""GPU parity of meshes, dual geometry and every FastDEC operator against the
CPU oracle, through the C ABI (libdec_hip.so).

Bars: integer / index / orientation work bit-exact; float64 operator outputs
within 1e-12 normwise relative error (‖x - x_ref‖∞ / ‖x_ref‖∞) of the oracle,
and bit-exact where both sides evaluate the same expression in the same order.
The triangle circumcenter is the exception: the oracle inverts the
Cayley-Menger matrix by elimination (as the reference's inv(CM) does), the
HIP library uses the closed-form barycentric coordinates, so every quantity
downstream of a circumcenter (B/C dual lengths, dual areas, ⋆0, ⋆1, the
Laplacian, wedge 0-2 / 1-1 / dual-dual coefficients) agrees to rounding, and
is checked bit-exactly on barycentric duals, where the center formula is shared.
Float32 outputs report their error against the float64 oracle (bound 1e-5).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    d = np.max(np.abs(b)) if b.size else 0.0
    return 0.0 if b.size == 0 else float(np.max(np.abs(a - b)) / (d if d > 0 else 1.0))


def to_dev(torch, x, dtype):
    return torch.tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


# ----------------------------------------------------------------------------- cases
def oracle_icosphere(O, k):
    return O.icosphere(k)


def gpu_mesh_like(cs, m):
    """Product mesh from an oracle mesh's arrays (orientations as given)."""
    return cs.mesh_from_arrays(m.p, m.v0, m.v1, m.e0, m.e1, m.e2, m.eo, m.to)


def grid345_arrays(O):
    P = np.array([[0, 4, 0], [3, 4, 0], [6, 4, 0], [0, 0, 0], [3, 0, 0], [6, 0, 0],
                  [0, -4, 0], [3, -4, 0], [6, -4, 0]], float)
    tris = [[0, 1, 3], [4, 1, 3], [4, 1, 2], [4, 5, 2], [4, 6, 3], [4, 6, 7], [4, 5, 7], [8, 5, 7]]
    m = O.glue_triangles(P, tris, sorted_=True)
    m.eo[:] = 1
    assert O.orient(m)
    m.to = (1 - m.to).astype(np.int8)
    return m


# uvsphere_bary: makeSphere(5, 175, 5, 0, 360, 5, 6371+90) with barycentric
# duals, and tg10_bary / tg10_circ: triangulated_grid(10, 10, 8, 8) with either
# dual, the meshes of test/Operators.jl:60 and test/Backends.jl:49-57
CASES = ["ico3", "ico5", "grid", "grid345_bary", "ico3_bary", "square_unoriented", "uvsphere_bary", "tg10_bary",
         "tg10_circ"]


@pytest.fixture(scope="module", params=CASES)
def case(request, O, cs):
    name = request.param
    circ = True
    if name.startswith("ico"):
        k = int(name[3:4])
        circ = not name.endswith("bary")
        om = O.icosphere(k)
        gm = cs.icosphere(k)
    elif name == "grid":
        om = O.triangulated_grid(1, 1, 1 / 8, math.sqrt(3) / 16, compress=True)
        gm = cs.triangulated_grid(1, 1, 1 / 8, math.sqrt(3) / 16, True)
    elif name == "grid345_bary":
        om = grid345_arrays(O)
        gm = gpu_mesh_like(cs, om)
        circ = False
    elif name.startswith("tg10"):
        om = O.triangulated_grid(10, 10, 8, 8, compress=True)
        gm = cs.triangulated_grid(10, 10, 8, 8, True)
        circ = name.endswith("circ")
    elif name == "uvsphere_bary":
        args = (5, 175, 5, 0, 360, 5, 6371 + 90)
        om = O.make_sphere(*args)[0]
        gm = cs.make_sphere(*args)[0]
        circ = False
    else:
        om = O.glue_triangles(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.2]], float),
                              [[0, 1, 2], [0, 2, 3]])
        gm = gpu_mesh_like(cs, om)
    od = O.Dual(om, circumcenter=circ)
    cs.dualize(gm, cs.Circumcenter if circ else cs.Barycenter)
    return name, od, gm


def exact_case(name):
    """Cases whose dual centers are computed by the same formula on both sides."""
    return name.endswith("bary")


def sparse_close(a, b, n, tol=1e-12):
    """Entry-wise comparison of two CSR matrices whose patterns may differ by
    entries that are rounding noise (absent on one side, |x| ≤ tol·max on the other)."""
    def todict(m):
        rp, ci, vv = m
        return {(r, int(ci[k])): float(vv[k]) for r in range(n) for k in range(rp[r], rp[r + 1])}
    da, db = todict(a), todict(b)
    scale = max(abs(v) for v in db.values())
    keys = set(da) | set(db)
    return max(abs(da.get(k, 0.0) - db.get(k, 0.0)) for k in keys) <= tol * scale


def test_topology_and_orientation_bit_exact(case):
    name, od, gm = case
    g = gm.export()
    m = od.m
    for a in ("v0", "v1", "e0", "e1", "e2"):
        assert np.array_equal(g[a], getattr(m, a)), a
    assert np.array_equal(g["edge_orientation"], od.eo)
    assert np.array_equal(g["tri_orientation"], od.to)        # orient! (DFS) vs the HIP rule
    assert np.array_equal(g["points"], m.p)                   # midpoint + projection, bitwise


def test_dual_geometry(case):
    name, od, gm = case
    g = gm.export_dual()
    assert rel(g["dual_point"], od.dual_point) <= F64_TOL
    for a in ("length", "area", "dual_length", "dual_area"):
        assert rel(g[a], getattr(od, a)) <= F64_TOL, a
    # lengths and edge-center duals involve only correctly rounded operations
    assert np.array_equal(g["length"], od.length)
    assert np.array_equal(g["dual_length"][:2 * od.ne], od.dual_length[:2 * od.ne])


def forms(O, od, seed):
    return (O.uniform_form(seed, od.nv), O.uniform_form(seed + 1, od.ne),
            O.uniform_form(seed + 2, od.ne), O.uniform_form(seed + 3, od.nt))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_wedges(case, O, cs, dtype):
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = F64_TOL if dtype == "f64" else F32_TOL
    V, E, E2, T = forms(O, od, 11)
    dV, dE, dE2, dT = (to_dev(torch, x, td) for x in (V, E, E2, T))
    w01 = cs.dec_wedge_product(0, 1, gm, cast_float=td)
    w10 = cs.dec_wedge_product(1, 0, gm, cast_float=td)
    w02 = cs.dec_wedge_product(0, 2, gm, cast_float=td)
    w20 = cs.dec_wedge_product(2, 0, gm, cast_float=td)
    w11 = cs.dec_wedge_product(1, 1, gm, cast_float=td)
    ref01 = od.wedge01(V, E)
    ref02 = od.wedge02(V, T)
    ref11 = od.wedge11(E, E2)
    r01 = w01(dV, dE).cpu().numpy()
    assert rel(r01, ref01) <= tol
    # wedge 0-0 is the pointwise product (test/Operators.jl:229, 239 compare with ==)
    V2 = O.uniform_form(17, od.nv)
    w00 = cs.dec_wedge_product(0, 0, gm, cast_float=td)(dV, to_dev(torch, V2, td)).cpu().numpy()
    if dtype == "f64":
        assert np.array_equal(w00, V * V2)
    else:
        assert np.array_equal(w00, V.astype(np.float32) * V2.astype(np.float32))
    assert np.array_equal(w10(dE, dV).cpu().numpy(), r01)           # (1,0) swaps its arguments
    assert rel(w02(dV, dT).cpu().numpy(), ref02) <= tol
    assert rel(w20(dT, dV).cpu().numpy(), ref02) <= tol
    assert rel(w11(dE, dE2).cpu().numpy(), ref11) <= tol
    if dtype == "f32":
        # test/Backends.jl:131-135 (Float32 on the device): mean((got - ref)²) < 1e-15
        # against the FP64 wedge of the same, FP32-rounded, inputs
        V3, E3, F3, T3 = (x.astype(np.float32).astype(np.float64) for x in (V, E, E2, T))
        for got, ref in ((r01, od.wedge01(V3, E3)), (w02(dV, dT).cpu().numpy(), od.wedge02(V3, T3)),
                         (w11(dE, dE2).cpu().numpy(), od.wedge11(E3, F3))):
            assert np.mean((got.astype(np.float64) - ref) ** 2) < 1e-15
    if dtype == "f64" and exact_case(name):
        assert np.array_equal(w11(dE, dE2).cpu().numpy(), ref11)
    if dtype == "f64":
        assert np.array_equal(r01, ref01)                             # wedge 0-1 is bit-exact
        assert np.array_equal(w11(dE, dE).cpu().numpy(), np.zeros(od.nt))
        assert np.array_equal(w11(dE, dE2).cpu().numpy(), -w11(dE2, dE).cpu().numpy())
    # the coefficient tables themselves
    idx, co = w11.operator.tables(3, 3)
    e3, c3 = od.wedge11_coeffs()
    assert np.array_equal(idx.T.reshape(-1), e3)
    assert rel(co.T.reshape(-1), c3) <= (F64_TOL if dtype == "f64" else 1e-7)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_wedge_dd(case, O, cs, dtype):
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = F64_TOL if dtype == "f64" else F32_TOL
    F = O.uniform_form(21, od.nt)
    G = O.uniform_form(22, od.ne)
    dd01 = cs.dec_wedge_product_dd(0, 1, gm, cast_float=td)
    dd10 = cs.dec_wedge_product_dd(1, 0, gm, cast_float=td)
    r01 = dd01(to_dev(torch, F, td), to_dev(torch, G, td)).cpu().numpy()
    r10 = dd10(to_dev(torch, G, td), to_dev(torch, F, td)).cpu().numpy()
    ref01 = od.wedge_dd01(F, G)
    ref10 = od.wedge_dd10(G, F)
    assert rel(r01, ref01) <= tol and rel(r10, ref10) <= tol
    if dtype == "f64" and exact_case(name):
        assert np.array_equal(r01, ref01) and np.array_equal(r10, ref10)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_wedge_pd(case, O, cs, dtype):
    """dec_wedge_product_pd(0,1) / dec_wedge_product_dp(1,0) (FastDEC.jl:327-387)."""
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = F64_TOL if dtype == "f64" else F32_TOL
    g = O.uniform_form(61, od.nv)
    f = O.uniform_form(62, od.ne)
    pd = cs.dec_wedge_product_pd(0, 1, gm, cast_float=td)
    dp = cs.dec_wedge_product_dp(1, 0, gm, cast_float=td)
    rpd = pd(to_dev(torch, g, td), to_dev(torch, f, td)).cpu().numpy()
    rdp = dp(to_dev(torch, f, td), to_dev(torch, g, td)).cpu().numpy()
    ref = od.wedge_pd01(g, f)
    assert rel(rpd, ref) <= tol and rel(rdp, ref) <= tol
    assert np.array_equal(rpd, rdp)                   # f .* (M g) == (M g) .* f
    M = od.wedge_pd01_mat()
    rp, ci, vv = pd.operator.csr()
    orp, oci, ovv = M.csr()
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert rel(vv, ovv) <= F64_TOL
    if exact_case(name):
        assert np.array_equal(vv, ovv)
        if dtype == "f64":
            assert np.array_equal(rpd, ref)


def csr_of(O, A):
    rp, ci, vv = A.csr()
    return rp, ci, vv


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_derivatives(case, O, cs, dtype):
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    V, E, _, T = forms(O, od, 31)
    cases = [
        (cs.dec_differential(0, gm, cast_float=td), od.differential(0), V),
        (cs.dec_differential(1, gm, cast_float=td), od.differential(1), E),
        (cs.dec_boundary(1, gm, cast_float=td), od.boundary(1), E),
        (cs.dec_boundary(2, gm, cast_float=td), od.boundary(2), T),
        (cs.dec_dual_derivative(0, gm, cast_float=td), od.dual_derivative(0), T),
        (cs.dec_dual_derivative(1, gm, cast_float=td), od.dual_derivative(1), E),
    ]
    for op, A, x in cases:
        rp, ci, vv = op.csr()
        orp, oci, ovv = csr_of(O, A)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(vv, ovv)
        y = op(to_dev(torch, x, td)).cpu().numpy()
        ref = A.matvec(x)
        if dtype == "f64":
            assert np.array_equal(y, ref)        # ±1 sums in the reference's column order
        else:
            assert rel(y, ref) <= F32_TOL


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_hodge(case, O, cs, dtype):
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = F64_TOL if dtype == "f64" else F32_TOL
    for k, n in ((0, od.nv), (1, od.ne), (2, od.nt)):
        x = O.uniform_form(40 + k, n)
        h = od.hodge_diag(k)
        ih = od.inv_hodge_diag(k)
        hs = cs.dec_hodge_star(k, gm, cs.DiagonalHodge, cast_float=td)
        assert rel(hs.csr()[2], h) <= F64_TOL
        if exact_case(name):
            assert np.array_equal(hs.csr()[2], h)
        assert rel(hs(to_dev(torch, x, td)).cpu().numpy(), h * x) <= tol
        if np.all(np.isfinite(ih)):
            ihs = cs.dec_inv_hodge_star(k, gm, cs.DiagonalHodge, cast_float=td)
            assert rel(ihs(to_dev(torch, x, td)).cpu().numpy(), ih * x) <= tol
    # geometric ⋆1: sparse matrix of the same pattern and values as the oracle's
    # geometric ⋆1 (Operators.jl:194-199 compares it with atol = 1e-12): on a
    # circumcentric dual its off-diagonal entries are rounding noise, so the
    # pattern is compared only where the centers are computed identically
    H = od.geo_hodge1()
    gh = cs.dec_hodge_star(1, gm, cs.GeometricHodge, cast_float=td)
    assert sparse_close(gh.csr(), H.csr(), od.ne)
    if exact_case(name):
        rp, ci, vv = gh.csr()
        orp, oci, ovv = H.csr()
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(vv, ovv)
    x = O.uniform_form(45, od.ne)
    assert rel(gh(to_dev(torch, x, td)).cpu().numpy(), H.matvec(x)) <= tol


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_inv_geometric_hodge1(case, O, cs, dtype):
    """Geometric ⋆1⁻¹ is x -> (-⋆1) \\ x (FastDEC.jl:744-748).  The expected
    value is a sparse LU solve of the oracle's ⋆1 (itself pinned by the golden
    fixtures); test/Operators.jl:215-220 compares with rtol = 1e-12."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    rp, ci, vv = od.geo_hodge1().csr()
    H = sp.csr_matrix((vv, ci, rp), shape=(od.ne, od.ne)).tocsc()
    lu = sla.splu(-H)
    inv = cs.dec_inv_hodge_star(1, gm, cs.GeometricHodge, cast_float=td)
    gh = cs.dec_hodge_star(1, gm, cs.GeometricHodge, cast_float=td)
    assert inv.shape == (od.ne, od.ne)
    for seed in (46, 47):              # a second solve reuses the work vectors
        x = O.uniform_form(seed, od.ne)
        ref = lu.solve(x)
        assert np.linalg.norm(H @ ref + x) <= 1e-13 * np.linalg.norm(x)
        z = inv(to_dev(torch, x, td))
        it, relres = inv.solve_info()
        if dtype == "f64":
            assert relres <= 1e-14 and 0 < it < 200, (it, relres)
            assert rel(z.cpu().numpy(), ref) <= F64_TOL
            assert rel(gh(z).cpu().numpy(), -x) <= 1e-13      # ⋆1 (⋆1⁻¹ x) = -(-x)
        else:
            assert relres <= 1e-6 and 0 < it < 100, (it, relres)
            assert rel(z.cpu().numpy(), ref) <= F32_TOL
    zero = torch.zeros(od.ne, dtype=td, device="cuda")
    assert torch.count_nonzero(inv(zero)).item() == 0
    with pytest.raises(cs.DecError):
        inv.csr()                      # a solve has no matrix to export
    with pytest.raises(cs.DecError):
        cs.apply_many([(inv, zero, None, torch.empty_like(zero))])
    with pytest.raises(cs.DecError):
        gh.solve_info()


def test_inv_geometric_hodge1_graph_replay(O, cs):
    """Captured into a graph, the solve records its whole iteration budget;
    iterations past convergence return at once, so every replay equals the
    eager solve bit for bit (barycentric dual: BiCGSTAB)."""
    import torch

    gm = cs.icosphere(4)
    cs.dualize(gm, cs.Barycenter)
    inv = cs.dec_inv_hodge_star(1, gm, cs.GeometricHodge, cast_float=torch.float64)
    x = torch.empty(inv.rows, dtype=torch.float64, device="cuda")
    out = torch.empty_like(x)
    x.copy_(torch.from_numpy(O.uniform_form(60, inv.rows)))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        inv.apply(x, out=out)                       # warm up outside capture
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        inv.apply(x, out=out)
    for seed in (61, 62, 63):
        x.copy_(torch.from_numpy(O.uniform_form(seed, inv.rows)))
        g.replay()
        torch.cuda.synchronize()
        replayed = out.clone()
        eager = inv(x)
        torch.cuda.synchronize()
        assert torch.equal(replayed, eager)
        assert inv.solve_info()[1] <= 1e-14


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("hodge", ["diagonal", "geometric"])
def test_laplacian(case, O, cs, dtype, hodge):
    import torch

    name, od, gm = case
    td = torch.float64 if dtype == "f64" else torch.float32
    tol = F64_TOL if dtype == "f64" else F32_TOL
    geo = hodge == "geometric"
    L = od.laplacian0(geometric=geo)
    op = cs.fast_laplace_beltrami(gm, cs.GeometricHodge if geo else cs.DiagonalHodge, cast_float=td)
    if geo and not exact_case(name):
        assert sparse_close(op.csr(), L.csr(), od.nv)
    else:
        rp, ci, vv = op.csr()
        orp, oci, ovv = L.csr()
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert rel(vv, ovv) <= F64_TOL
        if exact_case(name):
            assert np.array_equal(vv, ovv)         # reference SpGEMM order reproduced
    x = O.uniform_form(51, od.nv)
    y = op(to_dev(torch, x, td)).cpu().numpy()
    ref = L.matvec(x)
    assert rel(y, ref) <= tol
    if dtype == "f64" and exact_case(name):
        assert np.array_equal(y, ref)


def test_avg01(case, O, cs):
    import torch

    name, od, gm = case
    V, E, _, _ = forms(O, od, 61)
    A = cs.avg01_mat(gm)
    dV = to_dev(torch, V, torch.float64)
    dE = to_dev(torch, E, torch.float64)
    w = cs.dec_wedge_product(0, 1, gm)(dV, dE)
    assert torch.equal(A @ dV * dE, w)         # avg₀₁_mat*V .* E == wedge01 (Operators.jl:339-359)


def test_argument_errors(cs):
    import torch

    gm = cs.icosphere(1)
    with pytest.raises(cs.DecError, match="Unsupported combination of degrees 1 and 2"):
        cs.dec_wedge_product(1, 2, gm)
    with pytest.raises(cs.DecError, match="dual-dual wedges are 0-1 and 1-0"):
        cs.dec_wedge_product_dd(1, 1, gm)
    op = cs.dec_differential(0, gm)
    with pytest.raises(ValueError):
        op(torch.zeros(5, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        op(torch.zeros(gm.nv, dtype=torch.float32, device="cuda"))


def test_large_mesh_properties(cs):
    """icosphere k=8 (655,362 V): size-independent identities of the kernels."""
    import torch

    gm = cs.icosphere(8)
    cs.dualize(gm)
    assert (gm.nv, gm.ne, gm.nt) == (655362, 1966080, 1310720)
    f64 = torch.float64
    onesV = torch.ones(gm.nv, dtype=f64, device="cuda")
    onesE = torch.ones(gm.ne, dtype=f64, device="cuda")
    w01 = cs.dec_wedge_product(0, 1, gm)
    assert torch.equal(w01(onesV, onesE), onesE)
    E = cs.fill_uniform(torch.empty(gm.ne, dtype=f64, device="cuda"), 5)
    w11 = cs.dec_wedge_product(1, 1, gm)
    assert torch.count_nonzero(w11(E, E)).item() == 0
    # d1 d0 = 0 exactly on integer-valued 0-forms
    x = torch.randint(-1000, 1000, (gm.nv,), device="cuda").to(f64)
    d0 = cs.dec_differential(0, gm)
    d1 = cs.dec_differential(1, gm)
    assert torch.count_nonzero(d1 @ (d0 @ x)).item() == 0
    # L 1 = 0 (up to rounding) and ⋆0 sums to the surface area
    L = cs.fast_laplace_beltrami(gm, cs.DiagonalHodge)
    assert (L @ onesV).abs().max().item() < 1e-8
    h0 = cs.dec_hodge_star(0, gm, cs.DiagonalHodge) @ onesV
    area = (cs.dec_hodge_star(2, gm, cs.DiagonalHodge) @ torch.ones(gm.nt, dtype=f64, device="cuda"))
    assert abs(h0.sum().item() - (1 / area).sum().item()) < 1e-10 * 4 * math.pi
    assert abs(h0.sum().item() - 4 * math.pi) < 1e-4


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_apply_many_matches_single_applies(cs, dtype):
    """One grouped launch of the whole operator set (plus the dual-dual wedge,
    the boundaries and avg01) gives bit-identical outputs to one launch per
    operator, including a misaligned form that takes the scalar path."""
    import torch

    td = torch.float64 if dtype == "f64" else torch.float32
    m = cs.dualize(cs.icosphere(4))
    ops = [cs.dec_wedge_product(0, 1, m, cast_float=td).operator,
           cs.dec_wedge_product(1, 1, m, cast_float=td).operator,
           cs.dec_wedge_product(0, 2, m, cast_float=td).operator,
           cs.dec_wedge_product_dd(0, 1, m, cast_float=td).operator,
           cs.dec_wedge_product_dd(1, 0, m, cast_float=td).operator,
           cs.dec_differential(0, m, cast_float=td), cs.dec_differential(1, m, cast_float=td),
           cs.dec_boundary(1, m, cast_float=td), cs.dec_boundary(2, m, cast_float=td),
           cs.dec_hodge_star(0, m, cs.DiagonalHodge, cast_float=td),
           cs.dec_hodge_star(1, m, cs.GeometricHodge, cast_float=td),
           cs.dec_inv_hodge_star(2, m, cs.DiagonalHodge, cast_float=td),
           cs.fast_laplace_beltrami(m, cs.DiagonalHodge, cast_float=td),
           cs.avg01_mat(m, cast_float=td),
           cs.dec_wedge_product_pd(0, 1, m, cast_float=td).operator,
           cs.dec_wedge_product_dp(1, 0, m, cast_float=td).operator]
    items, seed = [], 40
    for k, op in enumerate(ops):
        x = cs.fill_uniform(torch.empty(op.cols + (1 if k == 1 else 0), dtype=td, device="cuda"), seed)
        x = x[1:] if k == 1 else x               # odd offset: scalar path
        y = cs.fill_uniform(torch.empty(op.cols2, dtype=td, device="cuda"), seed + 1) if op.cols2 else None
        items.append((op, x, y, torch.full((op.rows,), float("nan"), dtype=td, device="cuda")))
        seed += 2
    cs.apply_many(items)
    for op, x, y, out in items:
        ref = op.apply(x, y)
        assert torch.equal(out, ref)


def test_apply_many_splits_long_lists(cs):
    """A list longer than one grouped launch holds (37 applications of six
    operators, repeats included) is split over several launches and still
    matches one launch per application bit for bit."""
    import torch

    f64 = torch.float64
    m = cs.dualize(cs.icosphere(3))
    ops = [cs.dec_wedge_product(0, 1, m).operator, cs.dec_wedge_product(1, 1, m).operator,
           cs.dec_differential(0, m), cs.dec_differential(1, m),
           cs.dec_hodge_star(1, m, cs.DiagonalHodge), cs.fast_laplace_beltrami(m, cs.DiagonalHodge)]
    items = []
    for k in range(37):
        op = ops[k % len(ops)]
        x = cs.fill_uniform(torch.empty(op.cols, dtype=f64, device="cuda"), 100 + 2 * k)
        y = cs.fill_uniform(torch.empty(op.cols2, dtype=f64, device="cuda"), 101 + 2 * k) if op.cols2 else None
        items.append((op, x, y, torch.full((op.rows,), float("nan"), dtype=f64, device="cuda")))
    cs.apply_many(items)
    for op, x, y, out in items:
        assert torch.equal(out, op.apply(x, y))


def test_spatial_schedules_cover_every_row(O, cs):
    """On icosphere k=6 the gather operators run from spatial tile schedules
    (tiles dealt to the XCDs by position): single and grouped applications
    equal each other bit for bit and match the oracle, so every row is
    computed exactly once."""
    import torch

    k = 6
    od = O.Dual(O.icosphere(k))
    gm = cs.dualize(cs.icosphere(k))
    f64 = torch.float64
    V, E, E2, T = forms(O, od, 41)
    dV, dE, dE2, dT = (to_dev(torch, x, f64) for x in (V, E, E2, T))
    w01 = cs.dec_wedge_product(0, 1, gm)
    w02 = cs.dec_wedge_product(0, 2, gm)
    w11 = cs.dec_wedge_product(1, 1, gm)
    d0, d1 = cs.dec_differential(0, gm), cs.dec_differential(1, gm)
    L = cs.fast_laplace_beltrami(gm, cs.DiagonalHodge)
    singles = [w01(dV, dE), w02(dV, dT), w11(dE, dE2), d0 @ dV, d1 @ dE, L @ dV]
    refs = [od.wedge01(V, E), od.wedge02(V, T), od.wedge11(E, E2), od.differential(0).matvec(V),
            od.differential(1).matvec(E), od.laplacian0().matvec(V)]
    for got, ref in zip(singles, refs):
        assert rel(got.cpu().numpy(), ref) <= F64_TOL
    assert np.array_equal(singles[0].cpu().numpy(), refs[0])          # wedge 0-1 bit-exact
    assert np.array_equal(singles[3].cpu().numpy(), refs[3])          # d0 bit-exact
    ops = [w01.operator, w02.operator, w11.operator, d0, d1, L]
    items = [(op, x, y, torch.full((op.rows,), float("nan"), dtype=f64, device="cuda"))
             for op, x, y in zip(ops, [dV, dV, dE, dV, dE, dV], [dE, dT, dE2, None, None, None])]
    cs.apply_many(items)
    for (op, x, y, out), ref in zip(items, singles):
        assert torch.equal(out, ref)
