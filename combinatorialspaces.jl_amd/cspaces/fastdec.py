"""FastDEC operators on the GPU, with CombinatorialSpaces.jl's call signatures.

Mirrors src/FastDEC.jl (and its CUDA placement in
ext/CombinatorialSpacesCUDAExt.jl) over the C ABI of libdec_hip.so.  Forms are
torch tensors on the GPU; `backend` plays the role of `Val(:HIP)`, and
`cast_float` (torch.float32 / torch.float64) the reference's cast_float.
Operators that the reference returns as sparse matrices are returned as
`Operator` objects: `op @ x` or `op(x)` applies them on the device and
`op.csr()` exports the matrix for inspection.
"""
from __future__ import annotations

import ctypes
from ctypes import c_int32, c_int64, c_void_p

import numpy as np

from . import _lib as L
from ._lib import DecError, check, lib, ptr

HIP = "HIP"


class DiagonalHodge:
    code = L.DEC_DIAGONAL_HODGE


class GeometricHodge:
    code = L.DEC_GEOMETRIC_HODGE


class Circumcenter:
    code = L.DEC_CIRCUMCENTER


class Barycenter:
    code = L.DEC_BARYCENTER


def _torch():
    import torch

    return torch


def _dtype_code(cast_float):
    torch = _torch()
    if cast_float is None or cast_float in (torch.float64, np.float64, "float64", "f64"):
        return L.DEC_F64
    if cast_float in (torch.float32, np.float32, "float32", "f32"):
        return L.DEC_F32
    raise ValueError(f"unsupported cast_float {cast_float!r}")


def _torch_dtype(code):
    torch = _torch()
    return torch.float64 if code == L.DEC_F64 else torch.float32


def _stream(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def _hodge(h):
    if h is None:
        return GeometricHodge.code
    if isinstance(h, int):
        return h
    return h.code


# ---------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------
class Mesh:
    """A primal complex on the GPU -- EmbeddedDeltaSet2D (dim 2) or
    EmbeddedDeltaSet3D (dim 3) -- that becomes its own dual complex
    (EmbeddedDeltaDualComplex2D / 3D) once `dualize` has run."""

    def __init__(self, handle, owned=True):
        self._h = c_void_p(handle) if not isinstance(handle, c_void_p) else handle
        self._owned = owned
        nv, ne, nt, ntet, dim = c_int32(), c_int32(), c_int32(), c_int32(), c_int32()
        check(lib().dec_mesh_dim(self._h, ctypes.byref(dim)))
        self.dim = dim.value
        if self.dim == 3:
            check(lib().dec_mesh3_counts(self._h, ctypes.byref(nv), ctypes.byref(ne), ctypes.byref(nt),
                                         ctypes.byref(ntet)))
        else:
            check(lib().dec_mesh_counts(self._h, ctypes.byref(nv), ctypes.byref(ne), ctypes.byref(nt)))
        self.nv, self.ne, self.nt = nv.value, ne.value, nt.value
        self.ntet = ntet.value

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            try:
                lib().dec_mesh_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def export(self):
        """Host copies: dict of ∂v0, ∂v1, ∂e0, ∂e1, ∂e2 (0-based), orientations, points."""
        out = {
            "v0": np.zeros(self.ne, np.int32), "v1": np.zeros(self.ne, np.int32),
            "e0": np.zeros(self.nt, np.int32), "e1": np.zeros(self.nt, np.int32),
            "e2": np.zeros(self.nt, np.int32),
            "edge_orientation": np.zeros(self.ne, np.int8),
            "tri_orientation": np.zeros(self.nt, np.int8),
            "points": np.zeros((self.nv, 3)),
        }
        check(lib().dec_mesh_export(self._h, ptr(out["v0"]), ptr(out["v1"]), ptr(out["e0"]),
                                    ptr(out["e1"]), ptr(out["e2"]), ptr(out["edge_orientation"]),
                                    ptr(out["tri_orientation"]), ptr(out["points"])))
        if self.dim == 3:
            for k in range(4):
                out[f"t{k}"] = np.zeros(self.ntet, np.int32)
            out["tet_orientation"] = np.zeros(self.ntet, np.int8)
            check(lib().dec_mesh3_export(self._h, ptr(out["t0"]), ptr(out["t1"]), ptr(out["t2"]),
                                         ptr(out["t3"]), ptr(out["tet_orientation"])))
        return out

    def export_dual(self):
        nv, ne, nt = self.nv, self.ne, self.nt
        if self.dim == 3:
            n = self.ntet
            out = {"dual_point": np.zeros((nv + ne + nt + n, 3)), "length": np.zeros(ne),
                   "area": np.zeros(nt), "vol": np.zeros(n), "dual_length": np.zeros(2 * ne + 6 * nt),
                   "dual_area": np.zeros(6 * nt), "dual_vol": np.zeros(24 * n)}
            check(lib().dec_mesh3_export_dual(self._h, ptr(out["dual_point"]), ptr(out["length"]),
                                              ptr(out["area"]), ptr(out["vol"]), ptr(out["dual_length"]),
                                              ptr(out["dual_area"]), ptr(out["dual_vol"])))
            return out
        out = {"dual_point": np.zeros((nv + ne + nt, 3)), "length": np.zeros(ne),
               "area": np.zeros(nt), "dual_length": np.zeros(2 * ne + 6 * nt),
               "dual_area": np.zeros(6 * nt)}
        check(lib().dec_mesh_export_dual(self._h, ptr(out["dual_point"]), ptr(out["length"]),
                                         ptr(out["area"]), ptr(out["dual_length"]),
                                         ptr(out["dual_area"])))
        return out


def mesh_from_arrays(points, v0, v1, e0, e1, e2, edge_orientation=None, tri_orientation=None,
                     index_base=0) -> Mesh:
    """EmbeddedDeltaSet2D from boundary maps (MeshTopology / topo_to_mesh)."""
    a = [np.ascontiguousarray(x, np.int32) for x in (v0, v1, e0, e1, e2)]
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    eo = None if edge_orientation is None else np.ascontiguousarray(edge_orientation, np.int8)
    to = None if tri_orientation is None else np.ascontiguousarray(tri_orientation, np.int8)
    h = c_void_p()
    check(lib().dec_mesh_create(len(p), len(a[0]), len(a[2]), ptr(a[0]), ptr(a[1]), ptr(a[2]),
                                ptr(a[3]), ptr(a[4]), ptr(eo), ptr(to), ptr(p), index_base,
                                ctypes.byref(h)))
    return Mesh(h)


def mesh_from_triangles(points, triangles, sorted_=True, index_base=0) -> Mesh:
    """EmbeddedDeltaSet2D from a triangle list: glue_sorted_triangle! (or
    glue_triangle! with sorted_=False) per triangle, then orient!
    (src/MeshInterop.jl:116-131)."""
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    h = c_void_p()
    check(lib().dec_mesh_from_triangles(len(p), len(t), ptr(p), ptr(t), int(bool(sorted_)), index_base,
                                        ctypes.byref(h)))
    return Mesh(h)


def mesh3_from_arrays(points, v0, v1, e0, e1, e2, t0, t1, t2, t3, edge_orientation=None,
                      tri_orientation=None, tet_orientation=None, index_base=0) -> Mesh:
    """EmbeddedDeltaSet3D from boundary maps (SimplicialSets.jl:418-606)."""
    a = [np.ascontiguousarray(x, np.int32) for x in (v0, v1, e0, e1, e2, t0, t1, t2, t3)]
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    o = [None if x is None else np.ascontiguousarray(x, np.int8)
         for x in (edge_orientation, tri_orientation, tet_orientation)]
    h = c_void_p()
    check(lib().dec_mesh3_create(len(p), len(a[0]), len(a[2]), len(a[5]), *[ptr(x) for x in a],
                                 *[ptr(x) for x in o], ptr(p), index_base, ctypes.byref(h)))
    return Mesh(h)


def mesh_from_tetrahedra(points, tets, sorted_=True, index_base=0) -> Mesh:
    """EmbeddedDeltaSet3D(mesh) from a tetrahedron list: glue_sorted_tetrahedron!
    (or glue_tetrahedron! with sorted_=False) per tetrahedron, then orient!
    (src/MeshInterop.jl:141-153)."""
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    h = c_void_p()
    check(lib().dec_mesh3_from_tets(len(p), len(t), ptr(p), ptr(t), int(bool(sorted_)), index_base,
                                    ctypes.byref(h)))
    return Mesh(h)


def kuhn_box_arrays(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0, shear=(0.0, 0.0),
                    jitter_boundary=False):
    """Points and tetrahedra of an nx x ny x nz box of cubes, each cut into
    the six Kuhn tetrahedra 000 → +e_a → +e_a+e_b → 111 over the permutations
    (a, b, c) of the axes in lexicographic order; vertex (i, j, k) has id
    i + (nx+1)(j + (ny+1)k), cubes run with i fastest.  Interior points move by
    up to `jitter` cell widths (numpy default_rng(seed)), boundary points too
    with `jitter_boundary`; `shear` = (dx, dy) slides the top face, as
    CombMeshes.parallelepiped does.  A structured
    stand-in for the reference's TetGen meshes, which need TetGen.jl."""
    import itertools

    h = np.array([lx / nx, ly / ny, lz / nz])
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ijk = np.stack([i.ravel(order="F"), j.ravel(order="F"), k.ravel(order="F")], 1)
    p = ijk * h
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((ijk > 0) & (ijk < [nx, ny, nz]), axis=1) | jitter_boundary
        p[inner] += (rng.random((int(inner.sum()), 3)) * 2 - 1) * jitter * h
    p[:, 0] += shear[0] * p[:, 2] / lz
    p[:, 1] += shear[1] * p[:, 2] / lz
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    o = np.stack([ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")], 1)   # cubes, i fastest
    vid = lambda q: q[:, 0] + (nx + 1) * (q[:, 1] + (ny + 1) * q[:, 2])  # noqa: E731
    unit = np.eye(3, dtype=np.int64)
    tets = np.empty((len(o), 6, 4), np.int64)
    for s, (a, b, c) in enumerate(itertools.permutations(range(3))):
        q1 = o + unit[a]
        q2 = q1 + unit[b]
        tets[:, s] = np.stack([vid(o), vid(q1), vid(q2), vid(q2 + unit[c])], 1)
    return p, tets.reshape(-1, 4)


def kuhn_box(nx, ny, nz, **kw) -> Mesh:
    """The Kuhn box of kuhn_box_arrays glued as EmbeddedDeltaSet3D(mesh)."""
    p, t = kuhn_box_arrays(nx, ny, nz, **kw)
    return mesh_from_tetrahedra(p, t)


def _range_len(a, d, b):
    """length(a:d:b) for the integer-valued steps makeSphere is called with."""
    return int(np.floor((b - a) / d + 1e-9)) + 1 if b >= a else 0


def _sind(x):
    """sin of x degrees, exact at multiples of 90."""
    r = np.fmod(np.asarray(x, np.float64), 360.0)
    out = np.sin(np.radians(r))
    q = np.fmod(r, 90.0) == 0
    out[q] = np.array([0.0, 1.0, 0.0, -1.0])[(np.round(r[q] / 90).astype(int)) % 4]
    return out


def _cosd(x):
    return _sind(np.asarray(x, np.float64) + 90.0)


def make_sphere_arrays(min_lat, max_lat, d_lat, min_long, max_long, d_long, radius):
    """Points and boundary maps of makeSphere (src/CombMeshes.jl:238-342): a
    latitude-longitude sphere or band of radius `radius`, latitudes measured
    from the north pole in degrees.  Vertices run parallel by parallel, then
    the north and south poles when min_lat = 0 / max_lat = 180.  Edges are
    numbered as the reference creates them: each parallel's closing edge (when
    max_long = 360) and the first parallel's own edges are added explicitly,
    every other edge on first use by a triangle (glue_sorted_triangle!), which
    fixes edge and triangle ids index for index.  Returns (points, v0, v1, e0,
    e1, e2, north_pole, south_pole) with 0-based ids, -1 for a missing pole.
    Coordinates use numpy's sin/cos of radians with exact values at multiples
    of 90 degrees (the reference uses Julia's sind/cosd, which may round
    differently in the last bit)."""
    if (not (0 <= min_lat <= 180) or not (0 <= max_lat <= 180) or not (0 <= min_long <= 360)
            or max_lat < min_lat or max_long < min_long):
        raise ValueError("Mins must be less than Maxs, lats must be in [0,180], and longs must be in [0,360].")
    if min_lat == max_lat and d_lat == 0:
        d_lat = 1
    north = min_lat == 0
    south = max_lat == 180
    if north:
        min_lat += d_lat
    if south:
        max_lat -= d_lat
    close = max_long == 360
    if close:
        max_long -= d_long
    npar = _range_len(min_lat, d_lat, max_lat)
    nm = _range_len(min_long, d_long, max_long)
    if npar < 1 or nm < 2 or (close and nm < 3):
        raise ValueError("makeSphere: too few parallels or meridians for a triangulated surface")
    theta = min_lat + d_lat * np.arange(npar, dtype=np.float64)
    phi = min_long + d_long * np.arange(nm, dtype=np.float64)
    st, ct = _sind(theta)[:, None], _cosd(theta)[:, None]
    pts = np.stack([(radius * st) * _cosd(phi)[None, :], (radius * st) * _sind(phi)[None, :],
                    np.broadcast_to(radius * ct, (npar, nm))], axis=2).reshape(-1, 3)
    extra = []
    if north:
        extra.append([0.0, 0.0, radius * 1.0])
    if south:
        extra.append([radius * 0.0, 0.0, radius * -1.0])
    if extra:
        pts = np.vstack([pts, np.array(extra)])
    npole = npar * nm if north else -1
    spole = npar * nm + (1 if north else 0) if south else -1
    # The creation stream: explicit edges (kind 0, always new) and sorted
    # triangles (kind 1, three get_edge! lookups each), in reference order.
    # events: rows (kind, a, b, c), with a < b (< c for triangles)
    ev = []
    i = np.arange(nm - 1)
    for p in range(npar):
        o = p * nm
        if close:
            ev.append(np.array([[0, o, o + nm - 1, -1]]))
        if p == 0:
            ev.append(np.stack([np.zeros(nm - 1, int), o + i, o + i + 1, -np.ones(nm - 1, int)], 1))
            continue
        ta = np.stack([o + i, o + i + 1, o - nm + i], 1)
        tb = np.stack([o + 1 + i, o - nm + i, o - nm + 1 + i], 1)
        tl = np.array([[o + nm - 1, o, o - 1], [o - nm, o, o - 1]]) if close else np.zeros((0, 3), int)
        t = np.sort(np.vstack([ta, tb, tl]), axis=1)
        ev.append(np.hstack([np.ones((len(t), 1), int), t]))
    for pole, start in ((npole, 0), (spole, nm * (npar - 1))):
        if pole < 0:
            continue
        t = np.stack([np.full(nm - 1, pole), start + i, start + i + 1], 1)
        if close:
            t = np.vstack([t, [[pole, start, start + nm - 1]]])
        t = np.sort(t, axis=1)
        ev.append(np.hstack([np.ones((len(t), 1), int), t]))
    ev = np.vstack(ev)
    tri = ev[:, 0] == 1
    # every edge reference in stream order: explicit edges once, triangles as
    # (v0 v1), (v1 v2), (v0 v2) -- the order of glue_triangle!'s get_edge! calls
    ref = np.where(tri[:, None], ev[:, [1, 2, 1]], ev[:, [1, 1, 1]])
    ref = np.stack([ref, np.where(tri[:, None], ev[:, [2, 3, 3]], ev[:, [2, 2, 2]])], 2)
    used = np.where(tri[:, None], True, np.array([True, False, False])[None, :])
    pairs = ref[used]                                   # (src, tgt), stream order
    key = pairs[:, 0].astype(np.int64) * (len(pts) + 1) + pairs[:, 1]
    if len(np.unique(key[~np.repeat(tri, used.sum(1))])) != int((~tri).sum()):
        raise ValueError("makeSphere: these parameters repeat an explicit edge")
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")             # edge ids by first appearance
    rank = np.empty(len(uniq), np.int64)
    rank[order] = np.arange(len(uniq))
    eid = rank[inv.ravel()]
    v1 = pairs[first[order], 0]
    v0 = pairs[first[order], 1]
    slot = np.cumsum(used.ravel()).reshape(used.shape) - 1
    te = eid[slot[tri]]                                  # (ea, eb, ec) per triangle
    # add_triangle!(ea, eb, ec) stores ∂e0 = eb, ∂e1 = ec, ∂e2 = ea
    return pts, v0, v1, te[:, 1], te[:, 2], te[:, 0], npole, spole


def make_sphere(min_lat, max_lat, d_lat, min_long, max_long, d_long, radius):
    """makeSphere(minLat, maxLat, dLat, minLong, maxLong, dLong, radius)
    (src/CombMeshes.jl:238-342): (mesh, north_pole, south_pole), 0-based pole
    ids or -1.  Orientations are left unset, as the reference leaves them:
    dualize sets edge orientations to false and orients the triangles."""
    p, v0, v1, e0, e1, e2, npole, spole = make_sphere_arrays(min_lat, max_lat, d_lat, min_long, max_long,
                                                             d_long, radius)
    return mesh_from_arrays(p, v0, v1, e0, e1, e2), npole, spole


def icosphere(k: int) -> Mesh:
    """Pole-aligned icosahedron refined k times, new points on the unit sphere."""
    h = c_void_p()
    check(lib().dec_mesh_icosphere(k, ctypes.byref(h)))
    return Mesh(h)


def triangulated_grid(lx, ly, dx, dy, compress=True) -> Mesh:
    """triangulated_grid(lx, ly, dx, dy, Point3d, compress) (CombMeshes.jl:79-130)."""
    h = c_void_p()
    check(lib().dec_mesh_triangulated_grid(lx, ly, dx, dy, int(bool(compress)), ctypes.byref(h)))
    return Mesh(h)


def subdivision(s: Mesh, project=False) -> Mesh:
    """subdivision(s, BinarySubdivision()) (Multigrid.jl:447-452)."""
    h = c_void_p()
    check(lib().dec_mesh_subdivide(s.handle, int(bool(project)), ctypes.byref(h)))
    return Mesh(h)


def dualize(s: Mesh, center=Circumcenter) -> Mesh:
    """dualize(s, alg): dual complex + geometry, computed on the GPU in float64."""
    check(lib().dec_mesh_dualize(s.handle, center.code if hasattr(center, "code") else int(center)))
    return s


# ---------------------------------------------------------------------------
# Operators
# ---------------------------------------------------------------------------
class Operator:
    """A device operator; unary operators behave like the reference's sparse
    matrices (`op @ x`), binary ones (wedges) are called as `op(α, β)`."""

    def __init__(self, sd: Mesh, kind, m=0, n=0, hodge=GeometricHodge, cast_float=None):
        self.mesh = sd
        self.kind, self.m, self.n = kind, m, n
        self.dtype_code = _dtype_code(cast_float)
        self.dtype = _torch_dtype(self.dtype_code)
        h = c_void_p()
        check(lib().dec_op_create(sd.handle, kind, m, n, _hodge(hodge), self.dtype_code, ctypes.byref(h)))
        self._h = h
        r, c, c2 = c_int64(), c_int64(), c_int64()
        check(lib().dec_op_shape(h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(c2)))
        self.rows, self.cols, self.cols2 = r.value, c.value, c2.value

    @property
    def handle(self):
        return self._h

    @property
    def shape(self):
        return (self.rows, self.cols)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().dec_op_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def _check_form(self, x, n, name):
        torch = _torch()
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise TypeError(f"{name} must be a GPU torch tensor")
        if x.dtype != self.dtype:
            raise TypeError(f"{name} has dtype {x.dtype}, operator computes in {self.dtype}")
        if x.numel() != n or not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous with {n} entries, got {tuple(x.shape)}")

    def apply(self, x, y=None, out=None, stream=None):
        torch = _torch()
        self._check_form(x, self.cols, "x")
        if self.cols2:
            if y is None:
                raise TypeError("this operator takes two forms")
            self._check_form(y, self.cols2, "y")
        if out is None:
            out = torch.empty(self.rows, dtype=self.dtype, device=x.device)
        else:
            self._check_form(out, self.rows, "out")
        check(lib().dec_op_apply(self._h, c_void_p(x.data_ptr()),
                                 c_void_p(y.data_ptr()) if y is not None else None,
                                 c_void_p(out.data_ptr()), _stream(stream)))
        return out

    __call__ = apply

    def _args(self, x, y, out):
        self._check_form(x, self.cols, "x")
        if self.cols2:
            if y is None:
                raise TypeError("this operator takes two forms")
            self._check_form(y, self.cols2, "y")
        self._check_form(out, self.rows, "out")
        return self._h.value, x.data_ptr(), (y.data_ptr() if y is not None else None), out.data_ptr()

    def __matmul__(self, x):
        return self.apply(x)

    def solve_info(self, stream=None):
        """(iterations, ‖x - ⋆1 z‖/‖x‖) of the last apply of the geometric
        ⋆1⁻¹ solve (synchronises the stream)."""
        it, rr = c_int32(), ctypes.c_double()
        check(lib().dec_op_solve_info(self._h, ctypes.byref(it), ctypes.byref(rr), _stream(stream)))
        return it.value, rr.value

    def csr(self):
        """(rowptr, col, val) of the operator's matrix (float64 values)."""
        nnz = c_int64()
        check(lib().dec_op_export_csr(self._h, ctypes.byref(nnz), None, None, None))
        rp = np.zeros(self.rows + 1, np.int32)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        check(lib().dec_op_export_csr(self._h, None, ptr(rp), ptr(col), ptr(val)))
        return rp, col, val

    def dense(self):
        rp, col, val = self.csr()
        D = np.zeros((self.rows, self.cols))
        for r in range(self.rows):
            for k in range(rp[r], rp[r + 1]):
                D[r, col[k]] += val[k]
        return D

    def tables(self, nidx, ncoeff):
        idx = np.zeros(nidx * self.rows, np.int32)
        coeff = np.zeros(max(ncoeff, 1) * self.rows)
        check(lib().dec_op_export_tables(self._h, ptr(idx), ptr(coeff)))
        return idx.reshape(nidx, self.rows), coeff.reshape(max(ncoeff, 1), self.rows)[:ncoeff]


def _check_backend(backend):
    if backend not in (HIP, "hip", None):
        raise ValueError("this library only implements the HIP backend")


def dec_wedge_product(m, n, sd: Mesh, backend=HIP, arr_cons=None, cast_float=None):
    """dec_wedge_product(m, n, sd, Val(:HIP), arr_cons, cast_float) (FastDEC.jl:255-289):
    returns a function (α, β) -> α ∧ β computed on the GPU."""
    _check_backend(backend)
    if (m, n) == (0, 0):
        return lambda f, g: f * g
    op = Operator(sd, L.DEC_OP_WEDGE, m, n, cast_float=cast_float)
    fn = lambda a, b, out=None, stream=None: op.apply(a, b, out=out, stream=stream)  # noqa: E731
    fn.operator = op
    return fn


def dec_c_wedge_product_(m, n, res, a, b, op: Operator, stream=None):
    """In-place form of dec_c_wedge_product! (FastDEC.jl:217-232): res .= α ∧ β."""
    return op.apply(a, b, out=res, stream=stream)


def dec_wedge_product_dd(m, n, sd: Mesh, backend=HIP, cast_float=None):
    """dec_wedge_product_dd (FastDEC.jl:313-325): dual-dual wedge 0-1 / 1-0."""
    _check_backend(backend)
    op = Operator(sd, L.DEC_OP_WEDGE_DD, m, n, cast_float=cast_float)
    fn = lambda f, g, out=None, stream=None: op.apply(f, g, out=out, stream=stream)  # noqa: E731
    fn.operator = op
    return fn


def _binary(op):
    fn = lambda a, b, out=None, stream=None: op.apply(a, b, out=out, stream=stream)  # noqa: E731
    fn.operator = op
    return fn


def dec_wedge_product_pd(m, n, sd: Mesh, backend=HIP, cast_float=None):
    """dec_wedge_product_pd (FastDEC.jl:379-393): primal m-form ∧ dual n-form.
    (0,1): (g, f) -> (M g) .* f, g primal 0-form, f dual 1-form, M = wedge_pd_01_mat.
    (1,1): (f, g) -> ∧11(f, ♭♯ g), f primal 1-form, g dual 1-form."""
    _check_backend(backend)
    if (m, n) not in ((0, 1), (1, 1)):
        raise DecError(L.DEC_ERR_UNSUPPORTED, f"Unsupported combination of degrees {m} and {n}. Ensure that their "
                       "sum is not greater than the degree of the complex. The currently supported primal-dual "
                       "wedges are 0-1, 1-0, and 1-1.")
    return _binary(Operator(sd, L.DEC_OP_WEDGE_PD, m, n, cast_float=cast_float))


def dec_wedge_product_dp(m, n, sd: Mesh, backend=HIP, cast_float=None):
    """dec_wedge_product_dp (FastDEC.jl:374-377, 396-400): dual m-form ∧ primal n-form.
    (1,0): (f, g) -> f .* (M g), f dual 1-form, g primal 0-form.
    (1,1): (f, g) -> ∧11(♭♯ f, g), f dual 1-form, g primal 1-form."""
    _check_backend(backend)
    if (m, n) not in ((1, 0), (1, 1)):
        raise DecError(L.DEC_ERR_UNSUPPORTED, f"Unsupported combination of degrees {m} and {n}. Ensure that their "
                       "sum is not greater than the degree of the complex. The currently supported dual-primal "
                       "wedges are 0-1, 1-0, and 1-1.")
    return _binary(Operator(sd, L.DEC_OP_WEDGE_DP, m, n, cast_float=cast_float))


def flat_sharp_mat(sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """♭♯_mat(sd) (DiscreteExteriorCalculus.jl:1985-1992): a dual 1-form made
    primal by ♭(DPPFlat) ∘ ♯(LLSDDSharp); `op @ x`, `op.csr()`."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_FLAT_SHARP, cast_float=cast_float)


def interior_product_dd(m, n, sd: Mesh, backend=HIP, cast_float=None):
    """interior_product_dd (FastDEC.jl:758-786).
    (1,1): (f, g) -> ⋆2 ∧pd11(⋆1⁻¹ g, f), dual 1-forms -> dual 0-form.
    (1,2): (f, g) -> ⋆1 ♭♯ ((Λ01 ⋆0⁻¹ g) .* f), dual 1-form f, dual 2-form g."""
    _check_backend(backend)
    return _binary(Operator(sd, L.DEC_OP_INTERIOR_PRODUCT_DD, m, n, cast_float=cast_float))


def lie_derivative_dd(m, n, sd: Mesh, backend=HIP, cast_float=None):
    """ℒ_dd(1, 1, sd) (FastDEC.jl:788-803): (f, g) -> -(d̃0 ι11(f, g)) - ι12(f, d̃1 g)."""
    _check_backend(backend)
    return _binary(Operator(sd, L.DEC_OP_LIE_DD, m, n, cast_float=cast_float))


ℒ_dd = lie_derivative_dd


def dual_laplacian(n, sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """Δᵈ(Val(n), sd) (FastDEC.jl:808-860) on a 2-D complex: the dual 0-form
    (n = 0) or dual 1-form (n = 1) Laplacian; `op @ x`.  Both contain the
    geometric ⋆1⁻¹ solve."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_DUAL_LAPLACIAN, n, cast_float=cast_float)


def dec_differential(n, sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """dec_differential(n, sd) (FastDEC.jl:468-469)."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_DIFFERENTIAL, n, cast_float=cast_float)


def dec_boundary(n, sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """dec_boundary(n, sd) (FastDEC.jl:433-437)."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_BOUNDARY, n, cast_float=cast_float)


def dec_dual_derivative(n, sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """dec_dual_derivative(n, sd) (FastDEC.jl:443-453)."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_DUAL_DERIVATIVE, n, cast_float=cast_float)


def dec_hodge_star(n, sd: Mesh, hodge=GeometricHodge, backend=HIP, cast_float=None) -> Operator:
    """dec_hodge_star(n, sd, hodge) (FastDEC.jl:641-716)."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_HODGE, n, hodge=hodge, cast_float=cast_float)


def dec_inv_hodge_star(n, sd: Mesh, hodge=GeometricHodge, backend=HIP, cast_float=None) -> Operator:
    """dec_inv_hodge_star(n, sd, hodge) (FastDEC.jl:721-748).  With
    GeometricHodge and n = 1 the operator is the solve x -> (-⋆1) \\ x of
    FastDEC.jl:744-748, run as a Krylov iteration on the device."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_INV_HODGE, n, hodge=hodge, cast_float=cast_float)


def p2_d2_interpolation(sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """p2_d2_interpolation(sd) (DiscreteExteriorCalculus.jl:2016-2031): primal
    2-forms to dual 2-forms, nv x nt."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_P2_D2_INTERPOLATION, cast_float=cast_float)


def p3_d3_interpolation(sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """p3_d3_interpolation(sd) (DiscreteExteriorCalculus.jl:2038-2053): primal
    3-forms to dual 3-forms on a tetrahedral complex, nv x ntet."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_P3_D3_INTERPOLATION, cast_float=cast_float)


def d0_p0_interpolation(sd: Mesh, hodge=GeometricHodge, backend=HIP, cast_float=None) -> Operator:
    """d0_p0_interpolation(sd; hodge) (FastDEC.jl:915-926): dual 0-forms to
    primal 0-forms, ⋆0⁻¹ p2_d2 ⋆2⁻¹ (⋆0⁻¹ p3_d3 ⋆3⁻¹ on tetrahedra)."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_D0_P0_INTERPOLATION, hodge=hodge, cast_float=cast_float)


def p0_d0_interpolation(sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """p0_d0_interpolation(sd) (FastDEC.jl:928-941): primal 0-forms to the
    dual 0-forms at triangle centers, nt x nv."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_P0_D0_INTERPOLATION, cast_float=cast_float)


def avg01_mat(sd: Mesh, backend=HIP, cast_float=None) -> Operator:
    """avg₀₁_mat(sd) (FastDEC.jl:874-892)."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_AVG01, cast_float=cast_float)


def fast_laplace_beltrami(sd: Mesh, hodge=GeometricHodge, backend=HIP, cast_float=None) -> Operator:
    """-1 * ⋆0⁻¹ * d̃1 * ⋆1 * d0 (test/Multigrid.jl:46-48), i.e. ∇²(0, sd); on a
    3-D complex -⋆0⁻¹ d̃2 ⋆1 d0, which needs hodge=DiagonalHodge."""
    _check_backend(backend)
    return Operator(sd, L.DEC_OP_LAPLACIAN0, hodge=hodge, cast_float=cast_float)


def fill_uniform(x, seed, stream=None):
    """x[i] = 2u - 1 with u from splitmix64(seed * 2^32 + i) (device-side)."""
    code = L.DEC_F64 if x.dtype == _torch().float64 else L.DEC_F32
    check(lib().dec_fill_uniform(c_void_p(x.data_ptr()), x.numel(), seed, code, _stream(stream)))
    return x


def apply_many(items, stream=None):
    """One launch for several applications: items = [(op, x, y, out), ...]
    (y None for one-form operators), all operators of one dtype.  The items
    run concurrently: no out may overlap another item's x, y or out (the
    library raises DecError).  Same results as op.apply(x, y, out=out) for
    each item."""
    items = list(items)
    if not items:
        return
    args = [op._args(x, y, out) for op, x, y, out in items]
    n = len(args)
    arr = lambda vals: (c_void_p * n)(*vals)  # noqa: E731
    check(lib().dec_op_apply_many(n, arr([a[0] for a in args]), arr([a[1] for a in args]),
                                  arr([a[2] for a in args]), arr([a[3] for a in args]), _stream(stream)))
