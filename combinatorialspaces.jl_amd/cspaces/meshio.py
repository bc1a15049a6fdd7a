"""Mesh files on either side of the DEC path: OBJ, STL and JSON ACSets in,
OBJ and JSON ACSets out.

The reference reads meshes with FileIO/MeshIO into a GeometryBasics.Mesh and
glues its triangles (src/MeshInterop.jl:85-165); it writes them back through
GeometryBasics.Mesh(s) (src/MeshInterop.jl:17-25).  This module restates
those steps on numpy arrays:

* coordinates are read as MeshIO reads them by default, as Float32 points
  (Point3f), and widened exactly to float64;
* OBJ faces with more than three corners are split into a fan
  (v1, v_k, v_{k+1}), as GeometryBasics triangulates a polygon face;
* STL lists each corner by position, so points equal in all three Float32
  coordinates are merged, numbered in order of first occurrence
  (unique_index_map with force_unique, src/MeshInterop.jl:85-104);
* the mesh is then glue_sorted_triangle! of each face in file order followed
  by orient! (dec_mesh_from_triangles).

JSON ACSets (CombMeshes.loadmesh(Rectangle_30x10()) reads one with
parse_json_acset, src/CombMeshes.jl:42-55) hold one array per part type (V, E,
Tri, and Tet for a 3-D set), one object per part with its face maps (1-based)
and attributes: `point` as a coordinate array, orientations as booleans or
null (unset).  `_id` keys and dual parts (DualV, DualE, ... of a saved dual
complex) are ignored: the library recomputes the dual.  No ACSet file of the
reference is available offline, so the reader follows ACSets.jl's documented
layout and is checked by round trips only.

Parsing is host work and needs no GPU; `load_mesh` then builds the library's
mesh handle.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np


def _f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def read_obj(path):
    """(points nv x 3 float64, triangles nt x 3 int32, 0-based) of an OBJ file."""
    pts, tris = [], []
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                pts.append([float(t) for t in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    v = int(tok.split("/")[0])
                    idx.append(v - 1 if v > 0 else len(pts) + v)      # negative: relative to the end
                for k in range(1, len(idx) - 1):
                    tris.append([idx[0], idx[k], idx[k + 1]])
    P = _f32(np.array(pts, np.float64).reshape(-1, 3))
    T = np.array(tris, np.int32).reshape(-1, 3)
    if T.size and (T.min() < 0 or T.max() >= len(P)):
        raise ValueError(f"{path}: face refers to a vertex that does not exist")
    return P, T


def _stl_corners(path):
    with open(path, "rb") as fh:
        data = fh.read()
    if len(data) >= 84:
        (n,) = struct.unpack_from("<I", data, 80)
        if 84 + 50 * n == len(data):                                   # binary STL
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            return rec["v"].reshape(-1, 3).astype(np.float64)
    corners = []
    for line in data.decode("utf-8", errors="replace").splitlines():
        parts = line.split()
        if parts and parts[0] == "vertex":
            corners.append([float(t) for t in parts[1:4]])
    return _f32(np.array(corners, np.float64).reshape(-1, 3))


def unique_index_map(coords, force_unique):
    """(ind_map, uniques): with force_unique, equal points share the index of
    their first occurrence and `uniques` lists those first occurrences.
    Points are equal as Julia's isequal compares them (the Dict of
    MeshInterop.jl:84-103): bit for bit, so -0.0 and 0.0 differ, while every
    NaN equals every other NaN."""
    n = len(coords)
    if not force_unique:
        return np.arange(n), np.arange(n)
    c64 = np.array(coords, np.float64).reshape(n, -1)
    c64[np.isnan(c64)] = np.nan                      # one canonical NaN payload
    keys = c64.view(np.uint64)
    seen = {}
    ind_map = np.empty(n, np.int64)
    uniques = []
    for j, c in enumerate(map(tuple, keys)):
        k = seen.get(c)
        if k is None:
            k = seen[c] = len(uniques)
            uniques.append(j)
        ind_map[j] = k
    return ind_map, np.array(uniques, np.int64)


def read_stl(path, force_unique=True):
    """(points, triangles) of an ASCII or binary STL file, corners merged by position."""
    corners = _stl_corners(path)
    ind_map, uniques = unique_index_map(corners, force_unique)
    return corners[uniques], ind_map.reshape(-1, 3).astype(np.int32)


def _orient(rows, key):
    out = np.full(len(rows), -1, np.int8)
    for i, r in enumerate(rows):
        o = r.get(key)
        if o is not None:
            out[i] = 1 if bool(o) else 0
    return out


def read_json_acset(path_or_text):
    """The primal parts of a JSON ACSet (EmbeddedDeltaSet2D / 3D or a dual
    complex of one): dict with points (nv x 3), 0-based ∂v0, ∂v1, ∂e0..∂e2,
    ∂t0..∂t3 (3-D) and the orientations (1, 0, -1 unset)."""
    text = path_or_text
    if not text.lstrip().startswith("{"):
        with open(path_or_text, "r", encoding="utf-8") as fh:
            text = fh.read()
    data = json.loads(text)
    V, E, T = data.get("V", []), data.get("E", []), data.get("Tri", [])
    pts = np.zeros((len(V), 3))
    for i, r in enumerate(V):
        p = r["point"]
        if isinstance(p, dict):                      # {"x": .., "y": .., "z": ..} or {"data": [...]}
            p = p.get("data", [p.get("x", 0.0), p.get("y", 0.0), p.get("z", 0.0)])
        pts[i, :len(p)] = p[:3]
    hom = lambda rows, key: np.array([int(r[key]) - 1 for r in rows], np.int32)  # noqa: E731
    out = {"points": pts, "v0": hom(E, "∂v0"), "v1": hom(E, "∂v1"),
           "e0": hom(T, "∂e0"), "e1": hom(T, "∂e1"), "e2": hom(T, "∂e2"),
           "edge_orientation": _orient(E, "edge_orientation"), "tri_orientation": _orient(T, "tri_orientation")}
    if "Tet" in data:
        X = data["Tet"]
        for k in range(4):
            out[f"t{k}"] = hom(X, f"∂t{k}")
        out["tet_orientation"] = _orient(X, "tet_orientation")
    for key in ("v0", "v1"):
        if out[key].size and (out[key].min() < 0 or out[key].max() >= len(pts)):
            raise ValueError("JSON ACSet: an edge refers to a vertex that does not exist")
    return out


def write_json_acset(path, arrays):
    """A JSON ACSet of the primal parts in `arrays` (the layout read_json_acset
    returns, or Mesh.export()); 1-based face maps, `_id` on every part."""
    def orient(a, i):
        o = int(a[i])
        return None if o < 0 else bool(o)

    P = np.asarray(arrays["points"], np.float64).reshape(-1, 3)
    eo = arrays.get("edge_orientation", np.full(len(arrays["v0"]), -1))
    to = arrays.get("tri_orientation", np.full(len(arrays["e0"]), -1))
    data = {
        "V": [{"_id": i + 1, "point": p} for i, p in enumerate(P.tolist())],
        "E": [{"_id": i + 1, "∂v0": int(a) + 1, "∂v1": int(b) + 1, "edge_orientation": orient(eo, i)}
              for i, (a, b) in enumerate(zip(arrays["v0"], arrays["v1"]))],
        "Tri": [{"_id": i + 1, "∂e0": int(a) + 1, "∂e1": int(b) + 1, "∂e2": int(c) + 1,
                 "tri_orientation": orient(to, i)}
                for i, (a, b, c) in enumerate(zip(arrays["e0"], arrays["e1"], arrays["e2"]))],
    }
    if "t0" in arrays:
        teo = arrays.get("tet_orientation", np.full(len(arrays["t0"]), -1))
        data["Tet"] = [{"_id": i + 1, **{f"∂t{k}": int(arrays[f"t{k}"][i]) + 1 for k in range(4)},
                        "tet_orientation": orient(teo, i)} for i in range(len(arrays["t0"]))]
    with open(path, "w", encoding="utf-8") as fh:
        json.dump(data, fh, ensure_ascii=False)


def load_json_acset(path_or_text):
    """parse_json_acset(EmbeddedDeltaSet2D/3D, text): the library mesh of a JSON
    ACSet, with its boundary maps and orientations as stored."""
    from .fastdec import mesh3_from_arrays, mesh_from_arrays

    a = read_json_acset(path_or_text)
    if "t0" in a:
        return mesh3_from_arrays(a["points"], a["v0"], a["v1"], a["e0"], a["e1"], a["e2"], a["t0"], a["t1"],
                                 a["t2"], a["t3"], a["edge_orientation"], a["tri_orientation"], a["tet_orientation"])
    return mesh_from_arrays(a["points"], a["v0"], a["v1"], a["e0"], a["e1"], a["e2"], a["edge_orientation"],
                            a["tri_orientation"])


def read_mesh(path, force_unique=None):
    """Points and triangles of an OBJ or STL file (EmbeddedDeltaSet2D(fn),
    src/MeshInterop.jl:154-165: STL merges points by position)."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".stl":
        return read_stl(path, True if force_unique is None else force_unique)
    if ext == ".obj":
        P, T = read_obj(path)
        if force_unique:
            ind_map, uniques = unique_index_map(P, True)
            return P[uniques], ind_map[T].astype(np.int32)
        return P, T
    raise ValueError(f"{path}: only .obj and .stl are read")


def load_mesh(path, force_unique=None):
    """EmbeddedDeltaSet2D(fn): the library's mesh of an OBJ or STL file."""
    from .fastdec import mesh_from_triangles

    P, T = read_mesh(path, force_unique)
    return mesh_from_triangles(P, T, sorted_=True)


def write_obj(path, points, triangles):
    """An OBJ file of points and triangles (GeometryBasics.Mesh(s) written by
    MeshIO: `v` lines, then 1-based `f` lines)."""
    P = np.asarray(points, np.float64).reshape(-1, 3)
    T = np.asarray(triangles, np.int64).reshape(-1, 3)
    with open(path, "w", encoding="utf-8") as fh:
        for p in P.tolist():                     # shortest round-trip decimal of each float64
            fh.write(f"v {p[0]!r} {p[1]!r} {p[2]!r}\n")
        for t in T:
            fh.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def save_obj(path, mesh):
    """GeometryBasics.Mesh(s) of a library mesh (points and triangle_vertices,
    src/MeshInterop.jl:21-25) written as OBJ."""
    a = mesh.export()
    tv = np.stack([a["v1"][a["e1"]], a["v0"][a["e2"]], a["v0"][a["e1"]]], axis=1)
    write_obj(path, a["points"], tv)
