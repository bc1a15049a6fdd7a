"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py
from the oracle): the oracle must still reproduce them bit for bit (CPU), and
the HIP library must match them to the parity bars (GPU): indices and
orientations bit-exact, float64 outputs within 1e-12 normwise relative,
bit-exact where both sides evaluate the same expression (wedge 0-1, d, ∂,
and on barycentric duals every operator)."""
import glob
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "*.npz"))
                  if not os.path.basename(p).startswith("tet_"))
TET_FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "tet_*.npz")))
OPS3 = ["wedge01", "wedge02", "wedge03", "wedge11", "wedge21", "d0", "d1", "d2", "boundary3", "dual_d0", "dual_d1",
        "dual_d2", "hodge0", "hodge1", "hodge2", "hodge3", "laplacian0", "dual_laplacian0", "p3_d3", "d0_p0",
        "p0_d0"]
# ±1 matrices summed in the same (ascending column) order agree bit for bit
EXACT3 = {"d0", "d1", "d2", "boundary3", "dual_d0", "dual_d1", "dual_d2"}
OPS = ["wedge01", "wedge11", "wedge02", "wedge_dd01", "wedge_dd10", "wedge_pd01", "wedge_dp10", "d0", "d1",
       "boundary1", "boundary2",
       "dual_d0", "dual_d1", "hodge0", "hodge1", "hodge2", "inv_hodge0", "inv_hodge1", "inv_hodge2",
       "geo_hodge1", "laplacian0", "laplacian0_geometric",
       "flat_sharp", "wedge_pd11", "wedge_dp11", "inv_hodge1_geometric", "interior11", "interior12", "lie",
       "dual_laplacian0", "dual_laplacian1"]
EXACT_ALWAYS = {"wedge01", "d0", "d1", "boundary1", "boundary2", "dual_d0", "dual_d1"}
# on barycentric duals the center formula is shared: these agree bit for bit too
EXACT_BARY = {"wedge11", "wedge_dd01", "wedge_dd10", "wedge_pd01", "wedge_dp10", "hodge0", "hodge1", "hodge2", "geo_hodge1", "laplacian0",
              "laplacian0_geometric"}


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) == 2 and len(TET_FIXTURES) == 1


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_reproduces_fixture(O, path):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_fixtures

    f = load(path)
    mesh = O.Mesh(f["points"], f["v0"], f["v1"], f["e0"], f["e1"], f["e2"], f["edge_orient"], f["tri_orient"])
    now = make_fixtures.outputs(mesh, circumcenter="circumcenter" in path)
    assert set(now) == set(f)
    for k, v in f.items():
        if k in make_fixtures.LAPACK_KEYS:
            assert np.max(np.abs(now[k] - v)) <= 1e-13 * np.max(np.abs(v)), k
        else:
            assert np.array_equal(now[k], v), k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_hip_matches_fixture(cs, path):
    import torch

    f = load(path)
    circ = "circumcenter" in path
    gm = cs.mesh_from_arrays(f["points"], f["v0"], f["v1"], f["e0"], f["e1"], f["e2"], f["edge_orient"],
                             f["tri_orient"])
    cs.dualize(gm, cs.Circumcenter if circ else cs.Barycenter)
    g = gm.export()
    for a in ("v0", "v1", "e0", "e1", "e2"):
        assert np.array_equal(g[a], f[a]), a
    assert np.array_equal(g["edge_orientation"], f["edge_orient"])
    assert np.array_equal(g["tri_orientation"], f["tri_orient"])
    gd = gm.export_dual()
    for a in ("length", "area", "dual_length", "dual_area"):
        rel = np.max(np.abs(gd[a] - f[a])) / np.max(np.abs(f[a]))
        assert rel <= 1e-12, a
        if not circ:
            assert np.array_equal(gd[a], f[a]), a
    dv = lambda k: torch.tensor(f[k], dtype=torch.float64, device="cuda")  # noqa: E731
    got = {
        "wedge01": cs.dec_wedge_product(0, 1, gm)(dv("V"), dv("E")),
        "wedge11": cs.dec_wedge_product(1, 1, gm)(dv("E"), dv("E2")),
        "wedge02": cs.dec_wedge_product(0, 2, gm)(dv("V"), dv("T")),
        "wedge_dd01": cs.dec_wedge_product_dd(0, 1, gm)(dv("DT"), dv("DE")),
        "wedge_dd10": cs.dec_wedge_product_dd(1, 0, gm)(dv("DE"), dv("DT")),
        "wedge_pd01": cs.dec_wedge_product_pd(0, 1, gm)(dv("V"), dv("DE")),
        "wedge_dp10": cs.dec_wedge_product_dp(1, 0, gm)(dv("DE"), dv("V")),
        "d0": cs.dec_differential(0, gm) @ dv("V"), "d1": cs.dec_differential(1, gm) @ dv("E"),
        "boundary1": cs.dec_boundary(1, gm) @ dv("E"), "boundary2": cs.dec_boundary(2, gm) @ dv("T"),
        "dual_d0": cs.dec_dual_derivative(0, gm) @ dv("DT"), "dual_d1": cs.dec_dual_derivative(1, gm) @ dv("DE"),
        "hodge0": cs.dec_hodge_star(0, gm, cs.DiagonalHodge) @ dv("V"),
        "hodge1": cs.dec_hodge_star(1, gm, cs.DiagonalHodge) @ dv("E"),
        "hodge2": cs.dec_hodge_star(2, gm, cs.DiagonalHodge) @ dv("T"),
        "inv_hodge0": cs.dec_inv_hodge_star(0, gm, cs.DiagonalHodge) @ dv("V"),
        "inv_hodge1": cs.dec_inv_hodge_star(1, gm, cs.DiagonalHodge) @ dv("E"),
        "inv_hodge2": cs.dec_inv_hodge_star(2, gm, cs.DiagonalHodge) @ dv("T"),
        "geo_hodge1": cs.dec_hodge_star(1, gm, cs.GeometricHodge) @ dv("E"),
        "laplacian0": cs.fast_laplace_beltrami(gm, cs.DiagonalHodge) @ dv("V"),
        "laplacian0_geometric": cs.fast_laplace_beltrami(gm, cs.GeometricHodge) @ dv("V"),
        "flat_sharp": cs.flat_sharp_mat(gm) @ dv("DE"),
        "wedge_pd11": cs.dec_wedge_product_pd(1, 1, gm)(dv("E"), dv("DE")),
        "wedge_dp11": cs.dec_wedge_product_dp(1, 1, gm)(dv("DE"), dv("E")),
        "inv_hodge1_geometric": cs.dec_inv_hodge_star(1, gm, cs.GeometricHodge) @ dv("E"),
        "interior11": cs.interior_product_dd(1, 1, gm)(dv("DE"), dv("E2")),
        "interior12": cs.interior_product_dd(1, 2, gm)(dv("DE"), dv("V")),
        "lie": cs.lie_derivative_dd(1, 1, gm)(dv("DE"), dv("E2")),
        "dual_laplacian0": cs.dual_laplacian(0, gm) @ dv("DT"),
        "dual_laplacian1": cs.dual_laplacian(1, gm) @ dv("DE"),
    }
    assert set(got) == set(OPS)
    for k in OPS:
        x, ref = got[k].cpu().numpy(), f[k]
        rel = np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300)
        assert rel <= 1e-12, (k, rel)
        if k in EXACT_ALWAYS or (not circ and k in EXACT_BARY):
            assert np.array_equal(x, ref), k


@pytest.mark.parametrize("path", TET_FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_reproduces_tet_fixture(path):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_fixtures
    import tet

    f = load(path)
    mesh = tet.TetMesh(f["points"], f["v0"], f["v1"], f["e0"], f["e1"], f["e2"], f["t0"], f["t1"], f["t2"],
                       f["t3"], f["edge_orient"], f["tri_orient"], f["tet_orient"])
    now = make_fixtures.outputs3(mesh, circumcenter="circumcenter" in path)
    assert set(now) == set(f)
    for k, v in f.items():
        assert np.array_equal(now[k], v), k


@pytest.mark.gpu
@pytest.mark.parametrize("path", TET_FIXTURES, ids=lambda p: os.path.basename(p))
def test_hip_matches_tet_fixture(cs, path):
    import torch

    f = load(path)
    gm = cs.mesh3_from_arrays(f["points"], f["v0"], f["v1"], f["e0"], f["e1"], f["e2"], f["t0"], f["t1"], f["t2"],
                              f["t3"], f["edge_orient"], f["tri_orient"], f["tet_orient"])
    cs.dualize(gm, cs.Circumcenter if "circumcenter" in path else cs.Barycenter)
    g = gm.export()
    for a in ("v0", "v1", "e0", "e1", "e2", "t0", "t1", "t2", "t3"):
        assert np.array_equal(g[a], f[a]), a
    assert np.array_equal(g["tet_orientation"], f["tet_orient"])
    gd = gm.export_dual()
    for a in ("vol", "dual_vol"):
        assert np.max(np.abs(gd[a] - f[a])) <= 1e-12 * np.max(np.abs(f[a])), a
    dv = lambda k: torch.tensor(f[k], dtype=torch.float64, device="cuda")  # noqa: E731
    D = cs.DiagonalHodge
    got = {
        "wedge01": cs.dec_wedge_product(0, 1, gm)(dv("V"), dv("E")),
        "wedge02": cs.dec_wedge_product(0, 2, gm)(dv("V"), dv("T")),
        "wedge03": cs.dec_wedge_product(0, 3, gm)(dv("V"), dv("X")),
        "wedge11": cs.dec_wedge_product(1, 1, gm)(dv("E"), dv("E")),
        "wedge21": cs.dec_wedge_product(2, 1, gm)(dv("T"), dv("E")),
        "d0": cs.dec_differential(0, gm) @ dv("V"), "d1": cs.dec_differential(1, gm) @ dv("E"),
        "d2": cs.dec_differential(2, gm) @ dv("T"), "boundary3": cs.dec_boundary(3, gm) @ dv("X"),
        "dual_d0": cs.dec_dual_derivative(0, gm) @ dv("X"), "dual_d1": cs.dec_dual_derivative(1, gm) @ dv("T"),
        "dual_d2": cs.dec_dual_derivative(2, gm) @ dv("E"),
        "hodge0": cs.dec_hodge_star(0, gm, D) @ dv("V"), "hodge1": cs.dec_hodge_star(1, gm, D) @ dv("E"),
        "hodge2": cs.dec_hodge_star(2, gm, D) @ dv("T"), "hodge3": cs.dec_hodge_star(3, gm, D) @ dv("X"),
        "laplacian0": cs.fast_laplace_beltrami(gm, D) @ dv("V"),
        "dual_laplacian0": cs.dual_laplacian(0, gm) @ dv("X"),
        "p3_d3": cs.p3_d3_interpolation(gm) @ dv("X"), "d0_p0": cs.d0_p0_interpolation(gm) @ dv("X"),
        "p0_d0": cs.p0_d0_interpolation(gm) @ dv("V"),
    }
    assert set(got) == set(OPS3)
    for k in OPS3:
        x, ref = got[k].cpu().numpy(), f[k]
        rel = np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300)
        assert rel <= 1e-12, (k, rel)
        if k in EXACT3:
            assert np.array_equal(x, ref), k
