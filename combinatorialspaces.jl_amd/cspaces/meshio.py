"""Mesh files on either side of the DEC path: OBJ and STL in, OBJ out.

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

Parsing is host work and needs no GPU; `load_mesh` then builds the library's
mesh handle.
"""
from __future__ import annotations

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
    their first occurrence and `uniques` lists those first occurrences."""
    n = len(coords)
    if not force_unique:
        return np.arange(n), np.arange(n)
    seen = {}
    ind_map = np.empty(n, np.int64)
    uniques = []
    for j, c in enumerate(map(tuple, np.asarray(coords))):
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
