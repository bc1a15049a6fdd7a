"""Python face of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: what the HIP library is tested against.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

Every routine mirrors a reference routine of CombinatorialSpaces.jl v0.10.1;
the citations live beside the C restatement in dec_oracle.c.  Indices are
0-based here (the reference is 1-based).
"""
from __future__ import annotations

import ctypes
import math
import os
from ctypes import POINTER, c_double, c_int, c_ulonglong, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", _HERE, "all"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.om_glue_triangles.restype = c_int
        _lib.om_orient2.restype = c_int
        _lib.om_triangulated_grid.restype = c_int
        _lib.om_grid_counts.restype = c_int
        _lib.om_greedy_colors.restype = c_int
        _lib.om_splitmix_u01.restype = c_double
        _lib.om_splitmix_u01.argtypes = [c_ulonglong, c_ulonglong]
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


class _OCSC(ctypes.Structure):
    _fields_ = [("m", c_int), ("n", c_int), ("nnz", c_int),
                ("colptr", POINTER(c_int)), ("rowval", POINTER(c_int)),
                ("nzval", POINTER(c_double))]


class CSC:
    """A CSC matrix held in numpy arrays (0-based colptr/rowval)."""

    def __init__(self, m, n, colptr, rowval, nzval):
        self.m, self.n = int(m), int(n)
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int32)
        self.rowval = np.ascontiguousarray(rowval, dtype=np.int32)
        self.nzval = np.ascontiguousarray(nzval, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.colptr[-1])

    def struct(self) -> _OCSC:
        s = _OCSC()
        s.m, s.n, s.nnz = self.m, self.n, self.nnz
        s.colptr = self.colptr.ctypes.data_as(POINTER(c_int))
        s.rowval = self.rowval.ctypes.data_as(POINTER(c_int))
        s.nzval = self.nzval.ctypes.data_as(POINTER(c_double))
        return s

    @staticmethod
    def _take(s: _OCSC) -> "CSC":
        n, nnz = s.n, s.nnz
        cp = np.ctypeslib.as_array(s.colptr, shape=(n + 1,)).copy()
        rv = np.ctypeslib.as_array(s.rowval, shape=(max(nnz, 1),))[:nnz].copy()
        nz = np.ctypeslib.as_array(s.nzval, shape=(max(nnz, 1),))[:nnz].copy()
        out = CSC(s.m, n, cp, rv, nz)
        lib().om_csc_free(ctypes.byref(s))
        return out

    def dense(self):
        D = np.zeros((self.m, self.n))
        for j in range(self.n):
            for k in range(self.colptr[j], self.colptr[j + 1]):
                D[self.rowval[k], j] += self.nzval[k]
        return D

    def matvec(self, x):
        y = np.zeros(self.m)
        lib().om_spmv(ctypes.byref(self.struct()), _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def rmatvec(self, x):
        """transpose(A) * x."""
        y = np.zeros(self.n)
        lib().om_spmv_t(ctypes.byref(self.struct()), _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def transpose(self) -> "CSC":
        s = _OCSC()
        lib().om_transpose(ctypes.byref(self.struct()), ctypes.byref(s))
        return CSC._take(s)

    def csr(self):
        """(rowptr, colind, val) with columns ascending within each row."""
        t = self.transpose()
        return t.colptr, t.rowval, t.nzval

    def __matmul__(self, other: "CSC") -> "CSC":
        s = _OCSC()
        lib().om_spgemm(ctypes.byref(self.struct()), ctypes.byref(other.struct()), ctypes.byref(s))
        return CSC._take(s)


def sparse(m, n, I, J, V) -> CSC:
    """SparseArrays.sparse(I, J, V, m, n) with duplicates summed in input order."""
    I = np.ascontiguousarray(I, np.int32)
    J = np.ascontiguousarray(J, np.int32)
    V = np.ascontiguousarray(V, np.float64)
    s = _OCSC()
    lib().om_sparse(c_int(m), c_int(n), c_int(len(I)), _p(I), _p(J), _p(V), ctypes.byref(s))
    return CSC._take(s)


# --------------------------------------------------------------------------
# Seeded inputs
# --------------------------------------------------------------------------
def uniform_form(seed: int, n: int) -> np.ndarray:
    x = np.zeros(n)
    lib().om_uniform_form(c_ulonglong(seed), c_int(n), _p(x))
    return x


# --------------------------------------------------------------------------
# Primal meshes
# --------------------------------------------------------------------------
class Mesh:
    """Primal 2-D delta set: ∂v0/∂v1 per edge, ∂e0/∂e1/∂e2 per triangle."""

    def __init__(self, p, v0, v1, e0, e1, e2, eo=None, to=None):
        self.p = np.ascontiguousarray(p, np.float64).reshape(-1, 3)
        self.v0 = np.ascontiguousarray(v0, np.int32)
        self.v1 = np.ascontiguousarray(v1, np.int32)
        self.e0 = np.ascontiguousarray(e0, np.int32)
        self.e1 = np.ascontiguousarray(e1, np.int32)
        self.e2 = np.ascontiguousarray(e2, np.int32)
        self.nv, self.ne, self.nt = len(self.p), len(self.v0), len(self.e0)
        self.eo = (np.full(self.ne, -1, np.int8) if eo is None
                   else np.ascontiguousarray(eo, np.int8))
        self.to = (np.full(self.nt, -1, np.int8) if to is None
                   else np.ascontiguousarray(to, np.int8))

    def triangle_vertices(self):
        """triangle_vertices (SimplicialSets.jl:305-309)."""
        return np.stack([self.v1[self.e1], self.v0[self.e2], self.v0[self.e1]], axis=1)


def glue_triangles(points, tris, sorted_=False) -> Mesh:
    tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    nv, nt = len(points), len(tris)
    v0 = np.zeros(3 * nt, np.int32); v1 = np.zeros(3 * nt, np.int32)
    e0 = np.zeros(nt, np.int32); e1 = np.zeros(nt, np.int32); e2 = np.zeros(nt, np.int32)
    ne = lib().om_glue_triangles(c_int(nv), c_int(nt), _p(tris), c_int(int(sorted_)),
                                 _p(v0), _p(v1), _p(e0), _p(e1), _p(e2))
    return Mesh(points, v0[:ne], v1[:ne], e0, e1, e2)


def orient(m: Mesh) -> bool:
    to = np.zeros(m.nt, np.int8)
    ok = lib().om_orient2(c_int(m.ne), c_int(m.nt), _p(m.e0), _p(m.e1), _p(m.e2), _p(to))
    if ok:
        m.to = to
    return bool(ok)


def icosahedron() -> Mesh:
    p = np.zeros((12, 3)); tv = np.zeros((20, 3), np.int32)
    lib().om_icosahedron_points(_p(p)); lib().om_icosahedron_faces(_p(tv))
    return glue_triangles(p, tv, sorted_=True)


def triangulated_grid(lx, ly, dx, dy, compress=True) -> Mesh:
    nx = c_int(); ny = c_int()
    lib().om_grid_counts(c_double(lx), c_double(ly), c_double(dx), c_double(dy),
                         ctypes.byref(nx), ctypes.byref(ny))
    nv, nt = nx.value * ny.value, 2 * (nx.value - 1) * (ny.value - 1)
    p = np.zeros((nv, 3)); v0 = np.zeros(3 * nt, np.int32); v1 = np.zeros(3 * nt, np.int32)
    e0 = np.zeros(nt, np.int32); e1 = np.zeros(nt, np.int32); e2 = np.zeros(nt, np.int32)
    eo = np.zeros(3 * nt, np.int8); to = np.zeros(nt, np.int8)
    ne = lib().om_triangulated_grid(c_double(lx), c_double(ly), c_double(dx), c_double(dy),
                                    c_int(int(compress)), _p(p), _p(v0), _p(v1), _p(e0),
                                    _p(e1), _p(e2), _p(eo), _p(to))
    return Mesh(p, v0[:ne], v1[:ne], e0, e1, e2, eo[:ne], to)


def refine(m: Mesh) -> Mesh:
    nef, ntf = 2 * m.ne + 3 * m.nt, 4 * m.nt
    fv0 = np.zeros(nef, np.int32); fv1 = np.zeros(nef, np.int32)
    fe0 = np.zeros(ntf, np.int32); fe1 = np.zeros(ntf, np.int32); fe2 = np.zeros(ntf, np.int32)
    lib().om_refine(c_int(m.nv), c_int(m.ne), c_int(m.nt), _p(m.v0), _p(m.v1), _p(m.e0),
                    _p(m.e1), _p(m.e2), _p(fv0), _p(fv1), _p(fe0), _p(fe1), _p(fe2))
    return fv0, fv1, fe0, fe1, fe2


def propagate_points(m: Mesh, project=False) -> np.ndarray:
    pf = np.zeros((m.nv + m.ne, 3))
    lib().om_propagate_points(c_int(m.nv), c_int(m.ne), _p(m.v0), _p(m.v1), _p(m.p),
                              _p(pf), c_int(int(project)))
    return pf


def subdivide(m: Mesh, project=False) -> Mesh:
    """subdivision(s, BinarySubdivision()): new mesh, orientations unset."""
    pf = propagate_points(m, project)
    return Mesh(pf, *refine(m))


def subdivision_matrix(m: Mesh) -> CSC:
    s = _OCSC()
    lib().om_subdivision_matrix(c_int(m.nv), c_int(m.ne), _p(m.v0), _p(m.v1), ctypes.byref(s))
    return CSC._take(s)


def make_restriction(S: CSC) -> CSC:
    s = _OCSC()
    lib().om_make_restriction(ctypes.byref(S.struct()), ctypes.byref(s))
    return CSC._take(s)


def make_sphere(min_lat, max_lat, d_lat, min_long, max_long, d_long, radius):
    """makeSphere (src/CombMeshes.jl:238-342), restated call by call on a
    growing delta set: add_vertices!, add_sorted_edge(s)!, and
    glue_sorted_triangle! = add_triangle!(get_edge!(a,b), get_edge!(b,c),
    get_edge!(a,c)) on the sorted ids, get_edge! reusing the first edge with
    that (∂v1, ∂v0) (src/SimplicialSets.jl:323-364).  The argument check
    tests min_long twice, as the reference does.  Coordinates: sin/cos of
    radians, exact at multiples of 90 degrees.  Returns (Mesh, north_pole,
    south_pole) with 0-based pole ids, -1 when absent."""
    if (not (0 <= min_lat <= 180) or not (0 <= max_lat <= 180) or not (0 <= min_long <= 360)
            or not (0 <= min_long <= 360) or max_lat < min_lat or max_long < min_long):
        raise ValueError("Mins must be less than Maxs, lats must be in [0,180], and longs must be in [0,360].")

    def sind(x):
        r = math.fmod(x, 360.0)
        if math.fmod(r, 90.0) == 0:
            return (0.0, 1.0, 0.0, -1.0)[int(round(r / 90)) % 4]
        return math.sin(math.radians(r))

    def cosd(x):
        return sind(x + 90.0)

    def steps(a, d, b):                       # a:d:b
        out, k = [], 0
        while a + k * d <= b + 1e-9 * abs(d):
            out.append(a + k * d)
            k += 1
        return out

    pts, V0, V1, E0, E1, E2, first = [], [], [], [], [], [], {}

    def add_vertex(p):                        # 1-based ids, as in Julia
        pts.append(p)
        return len(pts)

    def add_edge(src, tgt):
        V1.append(src)
        V0.append(tgt)
        first.setdefault((src, tgt), len(V0))
        return len(V0)

    def add_sorted_edge(a, b):
        return add_edge(min(a, b), max(a, b))

    def get_edge(src, tgt):
        return first[(src, tgt)] if (src, tgt) in first else add_edge(src, tgt)

    def glue_sorted_triangle(i, j, k):
        a, b, c = sorted((i, j, k))
        x, y, z = get_edge(a, b), get_edge(b, c), get_edge(a, c)
        E0.append(y)
        E1.append(z)
        E2.append(x)

    def sph2car(rho, phi, theta):
        return (rho * sind(theta) * cosd(phi), rho * sind(theta) * sind(phi), rho * cosd(theta))

    if min_lat == max_lat and d_lat == 0:
        d_lat = 1
    north = south = False
    if min_lat == 0:
        min_lat += d_lat
        north = True
    if max_lat == 180:
        max_lat -= d_lat
        south = True
    close = False
    if max_long == 360:
        max_long -= d_long
        close = True
    lats, longs = steps(min_lat, d_lat, max_lat), steps(min_long, d_long, max_long)
    npar, nm = len(lats), len(longs)
    for theta in lats:
        off = len(pts) + 1
        for phi in longs:
            add_vertex(sph2car(radius, phi, theta))
        if close:
            add_sorted_edge(off + nm - 1, off)
        if off == 1:
            for a in range(off, off + nm - 1):
                add_sorted_edge(a, a + 1)
            continue
        for i, j, k in zip(range(off, off + nm - 1), range(off + 1, off + nm), range(off - nm, off - 1)):
            glue_sorted_triangle(i, j, k)
        for i, j, k in zip(range(off + 1, off + nm), range(off - nm, off - 1), range(off - nm + 1, off)):
            glue_sorted_triangle(i, j, k)
        if close:
            glue_sorted_triangle(off + nm - 1, off, off - 1)
            glue_sorted_triangle(off - nm, off, off - 1)
    npole = spole = 0
    if north:
        npole = add_vertex(sph2car(radius, 0, 0))
        for i in range(1, nm):
            glue_sorted_triangle(npole, i, i + 1)
        if close:
            glue_sorted_triangle(npole, 1, nm)
    if south:
        start = nm * (npar - 1) + 1
        spole = add_vertex(sph2car(radius, 0, 180))
        for i in range(start, start + nm - 1):
            glue_sorted_triangle(spole, i, i + 1)
        if close:
            glue_sorted_triangle(spole, start, start + nm - 1)
    z = lambda v: np.asarray(v, np.int64) - 1  # noqa: E731
    m = Mesh(np.array(pts, np.float64), z(V0), z(V1), z(E0), z(E1), z(E2))
    return m, npole - 1, spole - 1


def icosphere(k: int) -> Mesh:
    """Pole-aligned icosahedron refined k times, midpoints projected to |p| = 1."""
    m = icosahedron()
    for _ in range(k):
        m = subdivide(m, project=True)
    return m


# --------------------------------------------------------------------------
# Dual complex + FastDEC operators
# --------------------------------------------------------------------------
class Dual:
    """EmbeddedDeltaDualComplex2D with subdivide_duals!(sd, center)."""

    def __init__(self, m: Mesh, circumcenter=True):
        self.m = m
        nv, ne, nt = m.nv, m.ne, m.nt
        self.nv, self.ne, self.nt = nv, ne, nt
        nde, ndt = 2 * ne + 6 * nt, 6 * nt
        self.eo = m.eo.copy(); self.to = m.to.copy()
        self.D_v0 = np.zeros(nde, np.int32); self.D_v1 = np.zeros(nde, np.int32)
        self.D_e0 = np.zeros(ndt, np.int32); self.D_e1 = np.zeros(ndt, np.int32)
        self.D_e2 = np.zeros(ndt, np.int32)
        self.D_eo = np.zeros(nde, np.int8); self.D_to = np.zeros(ndt, np.int8)
        lib().om_dual_complex(c_int(nv), c_int(ne), c_int(nt), _p(m.v0), _p(m.v1), _p(m.e0),
                              _p(m.e1), _p(m.e2), _p(self.eo), _p(self.to), _p(self.D_v0),
                              _p(self.D_v1), _p(self.D_e0), _p(self.D_e1), _p(self.D_e2),
                              _p(self.D_eo), _p(self.D_to))
        self.dual_point = np.zeros((nv + ne + nt, 3))
        self.length = np.zeros(ne); self.area = np.zeros(nt)
        self.dual_length = np.zeros(nde); self.dual_area = np.zeros(ndt)
        lib().om_dual_geometry(c_int(nv), c_int(ne), c_int(nt), _p(m.v0), _p(m.v1), _p(m.e0),
                               _p(m.e1), _p(m.e2), _p(self.D_v0), _p(self.D_v1),
                               _p(self.D_e0), _p(self.D_e1), _p(self.D_e2), _p(m.p),
                               c_int(int(circumcenter)), _p(self.dual_point), _p(self.length),
                               _p(self.area), _p(self.dual_length), _p(self.dual_area))

    def dual_triangle_vertices(self, dt):
        return (self.D_v1[self.D_e1[dt]], self.D_v1[self.D_e0[dt]], self.D_v0[self.D_e0[dt]])

    def dual_edge_vertices(self, de):
        return (self.D_v0[de], self.D_v1[de])

    # --- Hodge stars ---
    def hodge_diag(self, k):
        n = (self.nv, self.ne, self.nt)[k]
        h = np.zeros(n)
        m = self.m
        lib().om_hodge_diag(c_int(k), c_int(self.nv), c_int(self.ne), c_int(self.nt),
                            _p(self.D_v0), _p(self.D_v1), _p(self.D_e1), _p(self.dual_length),
                            _p(self.dual_area), _p(self.length), _p(self.area), _p(h))
        del m
        return h

    def inv_hodge_diag(self, k):
        h = self.hodge_diag(k)
        out = np.zeros(len(h))
        lib().om_inv_hodge_diag(c_int(k), c_int(len(h)), _p(h), _p(out))
        return out

    def geo_hodge1(self) -> CSC:
        m = self.m
        s = _OCSC()
        lib().om_geo_hodge1(c_int(m.nv), c_int(m.ne), c_int(m.nt), _p(m.v0), _p(m.v1),
                            _p(m.e0), _p(m.e1), _p(m.e2), _p(self.dual_point), ctypes.byref(s))
        return CSC._take(s)

    # --- derivative / boundary ---
    def derivbound(self, k, transpose=False, negate=False) -> CSC:
        m = self.m
        s = _OCSC()
        lib().om_derivbound(c_int(k), c_int(m.nv), c_int(m.ne), c_int(m.nt), _p(m.v0),
                            _p(m.v1), _p(m.e0), _p(m.e1), _p(m.e2), _p(self.eo), _p(self.to),
                            c_int(int(transpose)), c_int(int(negate)), ctypes.byref(s))
        return CSC._take(s)

    def differential(self, k):          # dec_differential (FastDEC.jl:468-469)
        return self.derivbound(k)

    def boundary(self, k):              # dec_boundary (FastDEC.jl:433-437): ∂_k = d_{k-1}ᵀ
        return self.derivbound(k - 1, transpose=True)

    def dual_derivative(self, k):       # dec_dual_derivative 2-D (FastDEC.jl:446-450)
        return self.derivbound(1, transpose=True) if k == 0 else \
            self.derivbound(0, transpose=True, negate=True)

    # --- wedges ---
    def wedge02_coeffs(self):
        verts = np.zeros(6 * self.nt, np.int32); c = np.zeros(6 * self.nt)
        lib().om_wedge02_coeffs(c_int(self.nv), c_int(self.ne), c_int(self.nt), _p(self.D_v1),
                                _p(self.D_e2), _p(self.dual_area), _p(self.area), _p(verts), _p(c))
        return verts, c

    def wedge11_coeffs(self):
        m = self.m
        e3 = np.zeros(3 * self.nt, np.int32); c = np.zeros(3 * self.nt)
        lib().om_wedge11_coeffs(c_int(self.nt), _p(m.e0), _p(m.e1), _p(m.e2), _p(self.eo),
                                _p(self.to), _p(self.dual_area), _p(self.area), _p(e3), _p(c))
        return e3, c

    def wedge01(self, f, a):
        """dec_wedge_product(0,1,sd)(f, a): f a 0-form, a a 1-form."""
        res = np.zeros(self.ne)
        lib().om_wedge01(c_int(self.ne), _p(self.m.v0), _p(self.m.v1), _p(_f64(f)), _p(_f64(a)), _p(res))
        return res

    def wedge02(self, f, a):
        verts, c = self.wedge02_coeffs()
        res = np.zeros(self.nt)
        lib().om_wedge02(c_int(self.nt), _p(verts), _p(c), _p(_f64(f)), _p(_f64(a)), _p(res))
        return res

    def wedge11(self, a, b):
        e3, c = self.wedge11_coeffs()
        res = np.zeros(self.nt)
        lib().om_wedge11(c_int(self.nt), _p(e3), _p(c), _p(_f64(a)), _p(_f64(b)), _p(res))
        return res

    def wedge_dd01_mat(self) -> CSC:
        s = _OCSC()
        lib().om_wedge_dd01_mat(c_int(self.nv), c_int(self.ne), c_int(self.nt), _p(self.D_v0),
                                _p(self.D_v1), _p(self.dual_length), ctypes.byref(s))
        return CSC._take(s)

    def wedge_pd01_mat(self) -> CSC:
        s = _OCSC()
        m = self.m
        lib().om_wedge_pd01_mat(c_int(self.nv), c_int(self.ne), c_int(self.nt), _p(m.v0), _p(m.v1), _p(m.e1),
                                _p(m.e2), _p(self.D_v0), _p(self.D_v1), _p(self.dual_length), ctypes.byref(s))
        return CSC._take(s)

    def wedge_pd01(self, g, f):
        """dec_wedge_product_pd(0,1,sd)(g,f) = (m*g) .* f, g primal 0-form, f dual 1-form (ne)."""
        return self.wedge_pd01_mat().matvec(g) * f

    def wedge_dp10(self, f, g):
        """dec_wedge_product_dp(1,0,sd)(f,g) = f .* (m*g)."""
        return f * self.wedge_pd01_mat().matvec(g)

    def wedge_dd01(self, f, g):
        """dec_wedge_product_dd(0,1,sd)(f,g) = (m*f) .* g, f dual 0-form (nt), g dual 1-form (ne)."""
        return self.wedge_dd01_mat().matvec(f) * g

    def wedge_dd10(self, f, g):
        """dec_wedge_product_dd(1,0,sd)(f,g) = f .* (m*g)."""
        return f * self.wedge_dd01_mat().matvec(g)

    # --- Laplacian ---
    def laplacian0(self, geometric=False) -> CSC:
        """fast_laplace_beltrami (test/Multigrid.jl:46-48) as an assembled CSC."""
        d0 = self.differential(0)
        ih0 = self.inv_hodge_diag(0)
        s = _OCSC()
        if geometric:
            H = self.geo_hodge1()
            lib().om_laplacian0(ctypes.byref(d0.struct()), None, ctypes.byref(H.struct()),
                                _p(ih0), ctypes.byref(s))
        else:
            h1 = self.hodge_diag(1)
            lib().om_laplacian0(ctypes.byref(d0.struct()), _p(h1), None, _p(ih0), ctypes.byref(s))
        return CSC._take(s)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


# --------------------------------------------------------------------------
# Multigrid
# --------------------------------------------------------------------------
CG, GS, JACOBI, BLOCK_GS = 0, 1, 2, 3


class Hierarchy:
    """MultigridData(s, BinarySubdivision(), levels, op, steps) in DirectMode
    (src/Multigrid.jl:752-797): operators finest first, R_i / S_i between
    level i and i+1 (P_i = transpose(S_i)).  galerkin=True is GalerkinMode
    (src/Multigrid.jl:740-748): only the finest level is dualized and
    A_{i+1} = (R_i A_i) P_i."""

    def __init__(self, base: Mesh, levels: int, project=False, geometric=False,
                 circumcenter=True, galerkin=False):
        meshes = [base]
        S = []
        m = base
        for _ in range(levels):
            S.append(subdivision_matrix(m))
            m = subdivide(m, project)
            meshes.append(m)
        meshes = meshes[::-1]
        S = S[::-1]
        self.meshes = meshes
        self.S = S
        self.R = [make_restriction(s) for s in S]
        self.A = []
        self.duals = []
        for k, mm in enumerate(meshes):
            if galerkin and k > 0:
                self.A.append((self.R[k - 1] @ self.A[k - 1]) @ self.S[k - 1].transpose())
                continue
            mesh = mm
            if k == len(meshes) - 1:
                # the base level is dualized from the caller's mesh, which may
                # carry orientations; refined levels are re-oriented by dualize.
                mesh = Mesh(mm.p, mm.v0, mm.v1, mm.e0, mm.e1, mm.e2, mm.eo, mm.to)
            d = Dual(mesh, circumcenter)
            self.duals.append(d)
            self.A.append(d.laplacian0(geometric))

    def _structs(self):
        nl = len(self.A)
        A = (_OCSC * nl)(*[a.struct() for a in self.A])
        R = (_OCSC * max(nl - 1, 1))(*[r.struct() for r in self.R])
        S = (_OCSC * max(nl - 1, 1))(*[s.struct() for s in self.S])
        return nl, A, R, S

    def _colors(self, colors):
        """colors: None (greedy colours per level) or one colour array per
        level, finest first."""
        if colors is None:
            return None, None
        flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.int32) for c in colors]), np.int32)
        assert flat.size == sum(a.m for a in self.A)
        return flat, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))

    def mu_cycles(self, u, b, cycles, alg=CG, mu=1, steps=3, coarse_direct=False, omega=2 / 3, colors=None,
                  blocks=None):
        """blocks (alg=BLOCK_GS): one block-index array per level, finest
        first (None: one block per level)."""
        nl, A, R, S = self._structs()
        st = np.full(nl, steps, np.int32) if np.isscalar(steps) else np.ascontiguousarray(steps, np.int32)
        u = _f64(u).copy()
        keep, cp = self._colors(colors)
        keepb, bp = self._colors(blocks)
        lib().om_mu_cycles(c_int(nl), A, R, S, _p(st), c_int(alg), c_int(mu), c_int(cycles),
                           c_int(int(coarse_direct)), c_double(omega), _p(u), _p(_f64(b)), cp, bp)
        return u

    def full_multigrid(self, b, cycles, alg=CG, mu=1, steps=3, coarse_direct=False, omega=2 / 3, colors=None,
                       blocks=None):
        nl, A, R, S = self._structs()
        st = np.full(nl, steps, np.int32) if np.isscalar(steps) else np.ascontiguousarray(steps, np.int32)
        u = np.zeros(self.A[0].m)
        keep, cp = self._colors(colors)
        keepb, bp = self._colors(blocks)
        lib().om_full_multigrid(c_int(nl), A, R, S, _p(st), c_int(alg), c_int(mu), c_int(cycles),
                                c_int(int(coarse_direct)), c_double(omega), _p(_f64(b)), _p(u), cp, bp)
        return u


def greedy_colors(A: CSC):
    col = np.zeros(A.n, np.int32)
    nc = lib().om_greedy_colors(ctypes.byref(A.struct()), _p(col))
    return nc, col


def coarse_inverse(A: CSC):
    Z = np.zeros((A.m, A.m))
    lib().om_coarse_inverse(ctypes.byref(A.struct()), _p(Z))
    return Z


def cg(A: CSC, b, x0, itmax):
    x = _f64(x0).copy()
    lib().om_cg(ctypes.byref(A.struct()), _p(_f64(b)), _p(x), c_int(itmax))
    return x


def pcg(A: CSC, b, x0, iters, precond):
    """Preconditioned conjugate gradients, the library's DEC_SOLVE_PCG_*
    variants (no reference counterpart; the reference's solves are V/W cycles
    and full multigrid, src/Multigrid.jl:814-855).  The textbook recurrence
    (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 9.1):
    z = M r, β = (r·z)/(r·z)_prev, p = z + β p, α = (r·z)/(p·A p),
    x += α p, r -= α A p; α = 0 when p·A p is not positive.  `precond(r)`
    returns z.  Runs `iters` iterations; returns (x, r)."""
    x = _f64(x0).copy()
    r = _f64(b) - A.matvec(x)
    p = np.zeros_like(x)
    rz_prev = 1.0
    for _ in range(iters):
        z = precond(r)
        rz = float(r @ z)
        p = z + (rz / rz_prev) * p
        q = A.matvec(p)
        pq = float(p @ q)
        alpha = rz / pq if pq > 0 else 0.0
        x = x + alpha * p
        r = r - alpha * q
        rz_prev = rz
    return x, r
