"""Mesh files (cspaces/meshio.py) parsed on the host: OBJ faces (fans, the
v/vt/vn forms, negative indices), binary and ASCII STL with points merged by
position in order of first occurrence (src/MeshInterop.jl:85-104, 154-165),
Float32 coordinates as MeshIO reads them, and OBJ output.  The GPU test builds
library meshes from files and compares them with the oracle's glue + orient!.
tests/golden/assets holds the reference's own mesh files (test/assets/square.obj
and square.stl, data only), read as test/MeshInterop.jl:16-25 reads them."""
import os
import struct

import numpy as np
import pytest


def test_obj_faces_are_fans_and_indices_resolve(tmp_path):
    from cspaces import meshio

    f = tmp_path / "quad.obj"
    f.write_text("# a quad and a triangle\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0.1\nv 2 0 0\n"
                 "vt 0 0\nvn 0 0 1\n"
                 "f 1/1/1 2/1/1 3/1/1 4/1/1\n"
                 "f -4//1 5//1 -3//1\n")
    P, T = meshio.read_obj(str(f))
    assert P.shape == (5, 3) and P.dtype == np.float64
    assert P[3, 2] == np.float64(np.float32(0.1))               # read as Float32
    assert T.tolist() == [[0, 1, 2], [0, 2, 3], [1, 4, 2]]


def _square_corners():
    # two triangles of the unit square sharing the diagonal corners 0 and 2
    return [[(0, 0, 0), (1, 0, 0), (1, 1, 0)], [(0, 0, 0), (1, 1, 0), (0, 1, 0)]]


def test_binary_and_ascii_stl_merge_points_in_first_occurrence_order(tmp_path):
    from cspaces import meshio

    tris = _square_corners()
    b = tmp_path / "sq.stl"
    with open(b, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", len(tris)))
        for t in tris:
            fh.write(struct.pack("<3f", 0, 0, 1))
            for c in t:
                fh.write(struct.pack("<3f", *c))
            fh.write(struct.pack("<H", 0))
    a = tmp_path / "sq_ascii.stl"
    a.write_text("solid sq\n" + "".join(
        "facet normal 0 0 1\nouter loop\n" + "".join(f"vertex {x} {y} {z}\n" for x, y, z in t) +
        "endloop\nendfacet\n" for t in tris) + "endsolid sq\n")
    for f in (b, a):
        P, T = meshio.read_mesh(str(f))
        assert P.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
        assert T.tolist() == [[0, 1, 2], [0, 2, 3]]
        P2, T2 = meshio.read_stl(str(f), force_unique=False)       # six corners, unmerged
        assert len(P2) == 6 and T2.tolist() == [[0, 1, 2], [3, 4, 5]]


def test_stl_keeps_signed_zeros_apart(tmp_path):
    """Corners merge when Julia's isequal holds (MeshInterop.jl:84-103):
    -0.0 and 0.0 are different keys there, so a corner written as -0.0 is a
    vertex of its own."""
    from cspaces import meshio

    a = tmp_path / "signed_zero.stl"
    a.write_text("solid z\nfacet normal 0 0 1\nouter loop\n"
                 "vertex 0 0 0\nvertex 1 0 0\nvertex 1 1 0\nendloop\nendfacet\n"
                 "facet normal 0 0 1\nouter loop\n"
                 "vertex -0.0 0 0\nvertex 1 1 0\nvertex 0 1 0\nendloop\nendfacet\nendsolid z\n")
    P, T = meshio.read_stl(str(a))
    assert len(P) == 5
    assert T.tolist() == [[0, 1, 2], [3, 2, 4]]
    assert np.signbit(P[3, 0]) and not np.signbit(P[0, 0])


def test_obj_round_trip_and_glue(O, tmp_path):
    from cspaces import meshio

    m = O.icosphere(2)
    P = m.p.astype(np.float32).astype(np.float64)                   # Float32-representable
    tv = m.triangle_vertices()
    f = tmp_path / "ico2.obj"
    meshio.write_obj(str(f), P, tv)
    P2, T2 = meshio.read_obj(str(f))
    assert np.array_equal(P2, P) and np.array_equal(T2, tv)
    g = O.glue_triangles(P2, T2, sorted_=True)                      # what load_mesh glues
    assert (g.nv, g.ne, g.nt) == (162, 480, 320) and g.nv - g.ne + g.nt == 2
    assert O.orient(g)


ASSETS = os.path.join(os.path.dirname(__file__), "golden", "assets")


def _triangle_vertices(v0, v1, e0, e1, e2):
    """triangle_vertices (src/SimplicialSets.jl:305-309) from ∂ arrays."""
    return np.stack([v1[e1], v0[e2], v0[e1]], axis=1)


@pytest.mark.parametrize("name", ["square.obj", "square.stl"])
def test_reference_assets_glue_as_the_reference_tests_expect(O, name):
    """test/MeshInterop.jl:16-22: both files load, and triangle_vertices(s, 1)
    == [1, 2, 3] (0-based [0, 1, 2]); the STL's six corners merge into four
    points in order of first occurrence."""
    from cspaces import meshio

    P, T = meshio.read_mesh(os.path.join(ASSETS, name))
    assert P.tolist() == [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0]]
    assert T.tolist() == [[0, 1, 2], [0, 3, 2]]
    g = O.glue_triangles(P, T, sorted_=True)
    assert O.orient(g)
    assert (g.nv, g.ne, g.nt) == (4, 5, 2)
    assert _triangle_vertices(g.v0, g.v1, g.e0, g.e1, g.e2)[0].tolist() == [0, 1, 2]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["square.obj", "square.stl"])
def test_reference_assets_load_on_the_library(O, cs, name):
    """The same files through cs.load_mesh: the oracle's glue + orient!
    index for index, triangle_vertices(s, 1) == [1, 2, 3], and the barycentric
    dual of test/MeshInterop.jl:24-25 (areas 1/2 each, ⋆0 summing to 1)."""
    import torch
    from cspaces import meshio

    f = os.path.join(ASSETS, name)
    gm = cs.load_mesh(f)
    P, T = meshio.read_mesh(f)
    om = O.glue_triangles(P, T, sorted_=True)
    assert O.orient(om)
    a = gm.export()
    for k in ("v0", "v1", "e0", "e1", "e2"):
        assert np.array_equal(a[k], getattr(om, k)), k
    assert np.array_equal(a["tri_orientation"], om.to)
    assert _triangle_vertices(a["v0"], a["v1"], a["e0"], a["e1"], a["e2"])[0].tolist() == [0, 1, 2]
    cs.dualize(gm, cs.Barycenter)
    f64 = torch.float64
    area = 1.0 / (cs.dec_hodge_star(2, gm, cs.DiagonalHodge, cast_float=f64)
                  @ torch.ones(gm.nt, dtype=f64, device="cuda")).cpu().numpy()
    assert np.allclose(area, 0.5, rtol=0, atol=1e-15)
    h0 = (cs.dec_hodge_star(0, gm, cs.DiagonalHodge, cast_float=f64)
          @ torch.ones(gm.nv, dtype=f64, device="cuda")).sum().item()
    assert abs(h0 - 1.0) < 1e-15
    od = O.Dual(om, circumcenter=False)
    assert np.allclose((cs.dec_hodge_star(0, gm, cs.DiagonalHodge, cast_float=f64)
                        @ torch.ones(gm.nv, dtype=f64, device="cuda")).cpu().numpy(),
                       od.hodge_diag(0), rtol=1e-14, atol=0)


def test_unknown_extension_is_refused(tmp_path):
    from cspaces import meshio

    f = tmp_path / "m.ply"
    f.write_text("ply\n")
    with pytest.raises(ValueError):
        meshio.read_mesh(str(f))


@pytest.mark.gpu
def test_load_mesh_matches_the_oracle_glue(O, cs, tmp_path):
    from cspaces import meshio

    m = O.icosphere(2)
    P = m.p.astype(np.float32).astype(np.float64)
    f = tmp_path / "ico2.obj"
    meshio.write_obj(str(f), P, m.triangle_vertices())
    gm = cs.load_mesh(str(f))
    P2, T2 = meshio.read_obj(str(f))
    om = O.glue_triangles(P2, T2, sorted_=True)
    assert O.orient(om)
    a = gm.export()
    for k in ("v0", "v1", "e0", "e1", "e2"):
        assert np.array_equal(a[k], getattr(om, k)), k
    assert np.array_equal(a["tri_orientation"], om.to)
    assert np.array_equal(a["points"], P2)
    # the loaded mesh dualizes and its Laplacian annihilates constants
    cs.dualize(gm)
    import torch
    L = cs.fast_laplace_beltrami(gm, cs.DiagonalHodge, cast_float=torch.float64)
    one = torch.ones(gm.nv, dtype=torch.float64, device="cuda")
    assert (L @ one).abs().max().item() <= 1e-10
    # written back, the library mesh gives the same triangles
    g = tmp_path / "back.obj"
    meshio.save_obj(str(g), gm)
    P3, T3 = meshio.read_obj(str(g))
    assert np.array_equal(T3, om.triangle_vertices()) and np.array_equal(P3, P2)


def test_json_acset_round_trip_2d_and_3d(O, tmp_path):
    """JSON ACSets in ACSets.jl's layout: parts as arrays of objects, 1-based
    face maps, null for an unset orientation; `_id` and dual parts ignored."""
    import json

    import tet
    from cspaces import meshio

    m = O.icosphere(1)
    a = {"points": m.p, "v0": m.v0, "v1": m.v1, "e0": m.e0, "e1": m.e1, "e2": m.e2,
         "edge_orientation": np.where(np.arange(m.ne) % 3 == 0, -1, 1).astype(np.int8),
         "tri_orientation": np.full(m.nt, -1, np.int8)}
    f = tmp_path / "ico1.json"
    meshio.write_json_acset(str(f), a)
    b = meshio.read_json_acset(str(f))
    for k, v in a.items():
        assert np.array_equal(np.asarray(b[k]), np.asarray(v)), k
    raw = json.loads(f.read_text())
    assert raw["E"][0]["edge_orientation"] is None and raw["E"][1]["edge_orientation"] is True
    assert raw["Tri"][0]["∂e0"] == int(m.e0[0]) + 1
    # a saved dual complex carries dual parts as well; the reader keeps the primal ones
    raw["DualV"] = [{"_id": 1, "dual_point": [0, 0, 0]}]
    c = meshio.read_json_acset(json.dumps(raw))
    assert np.array_equal(c["e2"], m.e2)
    tm = tet.kuhn_grid(2, 1, 1)
    a3 = {"points": tm.p, "v0": tm.v0, "v1": tm.v1, "e0": tm.e0, "e1": tm.e1, "e2": tm.e2,
          "t0": tm.t0, "t1": tm.t1, "t2": tm.t2, "t3": tm.t3, "tet_orientation": tm.teo}
    g = tmp_path / "kuhn.json"
    meshio.write_json_acset(str(g), a3)
    b3 = meshio.read_json_acset(str(g))
    for k, v in a3.items():
        assert np.array_equal(np.asarray(b3[k]), np.asarray(v)), k
    assert (b3["edge_orientation"] == -1).all()
    with pytest.raises(ValueError):
        raw["E"][0]["∂v0"] = 10_000
        meshio.read_json_acset(json.dumps(raw))


@pytest.mark.gpu
def test_load_json_acset_builds_the_same_mesh(O, cs, tmp_path):
    from cspaces import meshio

    m = O.icosphere(2)
    f = tmp_path / "ico2.json"
    meshio.write_json_acset(str(f), {"points": m.p, "v0": m.v0, "v1": m.v1, "e0": m.e0, "e1": m.e1, "e2": m.e2})
    gm = cs.load_json_acset(str(f))
    cs.dualize(gm)
    a = gm.export()
    for k in ("v0", "v1", "e0", "e1", "e2"):
        assert np.array_equal(a[k], getattr(m, k)), k
    assert np.array_equal(a["points"], m.p)
