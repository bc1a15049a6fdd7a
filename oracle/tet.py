"""Tetrahedral complexes, their dual complex and the 3-D DEC operators,
restated in numpy and plain Python.

TEST INFRASTRUCTURE ONLY: what the HIP library is tested against.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

The dual complex is built simplex by simplex as the reference builds it, with
the same DualV / DualE / DualTri / DualTet numbering (0-based here), so that
elementary duals, subsimplices and primal vertices come from the reference's
own incidence definitions instead of a shortcut:

* glue_tetrahedron! / glue_sorted_tetrahedron!  SimplicialSets.jl:511-534, 323-364
* orient!(s, Val(3))                            SimplicialSets.jl:812-852
* tetrahedron_{triangles,edges,vertices}        SimplicialSets.jl:466-492
* triangle_vertices                             SimplicialSets.jl:304-309
* dual vertices / edges / triangles (1-D, 2-D)  DiscreteExteriorCalculus.jl:301-329, 545-606
* dual tetrahedra (24 per tetrahedron)          DiscreteExteriorCalculus.jl:1115-1245
* elementary_duals, subsimplices, primal_vertex DiscreteExteriorCalculus.jl:196-206, 462-470, 1034-1046
* geometry: volume, Circumcenter, Barycenter    SimplicialSets.jl:896-940, DiscreteExteriorCalculus.jl:94-128,
                                                385-422, 941-990, 1303-1335
* dec_differential 0..2, ∂ 1..3, d̃ 0..2 (3-D)   FastDEC.jl:433-437, 455-462, 468-548
* diagonal ⋆0..⋆3 and inverses (3-D)            FastDEC.jl:606-636, 721-727
* ∧ 0-k (wedge_product_zero)                    DiscreteExteriorCalculus.jl:2078-2093
* ∧ 1-1                                         DiscreteExteriorCalculus.jl:914-934
* ∧ 2-1 / 1-2                                   DiscreteExteriorCalculus.jl:1337-1381
* FastDEC's 0-3 and 2-1 coefficient tables      FastDEC.jl:64-76, 96-133, 173-202
* ∇²₀ with the diagonal ⋆                       DiscreteExteriorCalculus.jl:1836-1890
* p3_d3_interpolation, d0_p0_interpolation (3-D) DiscreteExteriorCalculus.jl:2038-2053, FastDEC.jl:924-926
* Δᵈ(Val(0)) on a 3-D complex                   FastDEC.jl:832-839

Circumcenters are taken, as in the reference, from the first row of the
inverse Cayley-Menger matrix (numpy's inverse here).

Pinned by the reference's 3-D tests (tests/test_oracle_tet.py): the cell
counts of one tetrahedron and of the square pyramid, the elementary-dual and
subsimplex incidences, d̃∘d̃ = 0, the regular tetrahedron's 24 equal dual
volumes and centroidal circumcenter, heat diffusion to the mean, equal
barycentric dual volumes of the six-tetrahedron cube, dXdY ∧ dZ = vol and
⋆3((dX ∧ dY) ∧ dZ) = 1, the 0-k wedges of constants, FastDEC's tables
against the pedagogical wedges, and orientation-flip invariance; and by
Σ_σ |σ||⋆σ| = C(3,k)·volume for circumcentric duals.
"""
from __future__ import annotations

import copy
import itertools

import numpy as np
import scipy.sparse as sp

UNSET = -1


class TetMesh:
    """EmbeddedDeltaSet3D: points and the boundary maps ∂v0, ∂v1 (edges),
    ∂e0..∂e2 (triangles), ∂t0..∂t3 (tetrahedra); orientations are int8
    (1 true, 0 false, -1 unset)."""

    def __init__(self, p, v0, v1, e0, e1, e2, t0, t1, t2, t3, eo=None, to=None, teo=None):
        self.p = np.asarray(p, np.float64).reshape(-1, 3)
        ar = lambda a: np.asarray(a, np.int64)
        self.v0, self.v1 = ar(v0), ar(v1)
        self.e0, self.e1, self.e2 = ar(e0), ar(e1), ar(e2)
        self.t0, self.t1, self.t2, self.t3 = ar(t0), ar(t1), ar(t2), ar(t3)
        self.nv, self.ne, self.nt, self.ntet = len(self.p), len(self.v0), len(self.e0), len(self.t0)

        def o(a, n):
            return np.full(n, UNSET, np.int8) if a is None else np.asarray(a, np.int8).copy()

        self.eo, self.to, self.teo = o(eo, self.ne), o(to, self.nt), o(teo, self.ntet)

    # SimplicialSets.jl:466-492
    def tetrahedron_triangles(self):
        return np.stack([self.t0, self.t1, self.t2, self.t3], 1)

    def tetrahedron_edges(self):
        return np.stack([self.e0[self.t0], self.e1[self.t0], self.e2[self.t0],
                         self.e1[self.t1], self.e2[self.t1], self.e2[self.t2]], 1)

    def tetrahedron_vertices(self):
        return np.stack([self.v1[self.e2[self.t2]], self.v0[self.e2[self.t2]],
                         self.v1[self.e0[self.t0]], self.v0[self.e0[self.t0]]], 1)

    # SimplicialSets.jl:304-309
    def triangle_vertices(self):
        return np.stack([self.v1[self.e1], self.v0[self.e2], self.v0[self.e1]], 1)


def glue_tetrahedra(points, tets, sorted_=True, tet_orientation=None) -> TetMesh:
    """glue_sorted_tetrahedron! (sorted_) or glue_tetrahedron! of each
    tetrahedron in order onto nv vertices (SimplicialSets.jl:511-534):
    get_triangle! reuses the triangle with directed edges v1→v2, v0→v2, v0→v1
    (as ∂e0, ∂e1, ∂e2), else glue_triangle! creates it, and get_edge! reuses
    the edge src→tgt (∂v1 = src, ∂v0 = tgt), else adds it."""
    edge, tri = {}, {}
    v0, v1, e0, e1, e2, T = [], [], [], [], [], []

    def get_edge(s, t):
        k = edge.get((s, t))
        if k is None:
            k = edge[(s, t)] = len(v0)
            v1.append(s)
            v0.append(t)
        return k

    def get_triangle(a, b, c):
        k = tri.get((a, b, c))
        if k is None:
            ea, eb, ec = get_edge(a, b), get_edge(b, c), get_edge(a, c)
            k = tri[(a, b, c)] = len(e0)
            e0.append(eb)
            e1.append(ec)
            e2.append(ea)
        return k

    for tv in np.asarray(tets, np.int64).reshape(-1, 4).tolist():
        a, b, c, d = sorted(tv) if sorted_ else tv
        T.append((get_triangle(b, c, d), get_triangle(a, c, d), get_triangle(a, b, d), get_triangle(a, b, c)))
    T = np.array(T, np.int64).reshape(-1, 4)
    teo = None if tet_orientation is None else np.full(len(T), int(tet_orientation), np.int8)
    return TetMesh(points, v0, v1, e0, e1, e2, T[:, 0], T[:, 1], T[:, 2], T[:, 3], teo=teo)


def orient3(m: TetMesh) -> bool:
    """orient!(s, Val(3)) (SimplicialSets.jl:812-852): depth-first from each
    unvisited tetrahedron in order; a neighbour y across face ∂t_i(x) = ∂t_j(y)
    gets -o(x) when i + j is even and o(x) otherwise.  Sets m.teo and returns
    True, or returns False (orientations unchanged) on a conflict."""
    faces = m.tetrahedron_triangles()
    cof = [[[] for _ in range(m.nt)] for _ in range(4)]                  # coface(3, j, s, tri), ascending
    for x in range(m.ntet):
        for j in range(4):
            cof[j][faces[x, j]].append(x)
    ors = np.zeros(m.ntet, np.int8)
    for seed in range(m.ntet):
        if ors[seed] != 0:
            continue
        stack = [seed]
        ors[seed] = 1
        while stack:
            x = stack.pop()
            ox = ors[x]
            nox = -ox
            for i in range(4):
                face = faces[x, i]
                for j in range(4):
                    same = (i + j) % 2 == 0
                    for y in cof[j][face]:
                        if y == x:
                            continue
                        oy = ors[y]
                        if oy == 0:
                            ors[y] = nox if same else ox
                            stack.append(y)
                        elif same and oy == ox:
                            return False
    m.teo = np.where(ors == 1, 1, 0).astype(np.int8)
    return True


# ---- meshes ---------------------------------------------------------------

def single_tetrahedron():
    """single_tetrahedron() (CombMeshes.jl:419-428): the origin and the unit
    basis vectors, glue_sorted_tetrahedron!, no orientations set."""
    p = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    return glue_tetrahedra(p, [(0, 1, 2, 3)])


def regular_tetrahedron():
    """The regular tetrahedron of edge 2√2 (test/DiscreteExteriorCalculus.jl:970-976)."""
    p = [(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]
    return glue_tetrahedra(p, [(0, 1, 2, 3)], sorted_=False)


def square_pyramid():
    """Two tetrahedra sharing triangle (1,2,4) (0-based), glued unsorted
    (test/DiscreteExteriorCalculus.jl:919-923); points are a placeholder."""
    p = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0.5, 0.5, 1)]
    return glue_tetrahedra(p, [(0, 1, 2, 4), (1, 2, 3, 4)], sorted_=False)


def tet_cube(tet_orientation=None):
    """glue_sorted_tet_cube! of vertices 0..7 (SimplicialSets.jl:536-557): six
    tetrahedra about the diagonal 1-7; the points are the cube labelled as the
    docstring's faces require (0,1,2,3 around the bottom, 4..7 above them)."""
    p = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    tets = [(0, 1, 3, 7), (1, 2, 3, 7), (0, 1, 4, 7), (1, 2, 6, 7), (1, 4, 5, 7), (1, 5, 6, 7)]
    return glue_tetrahedra(p, tets, sorted_=False, tet_orientation=tet_orientation)


def blessent_cube():
    """Six tetrahedra of equal volume forming the unit cube, Blessent 2009
    Table 3.1 (test/DiscreteExteriorCalculus.jl:994-1012): glue_sorted, every
    orientation set true, then orient!."""
    p = [(0, 1, 0), (0, 0, 0), (1, 1, 0), (1, 0, 0), (0, 1, 1), (0, 0, 1), (1, 1, 1), (1, 0, 1)]
    tets = [(3, 5, 4, 2), (7, 6, 8, 4), (5, 6, 7, 4), (3, 5, 7, 4), (5, 6, 4, 2), (1, 5, 3, 2)]
    m = glue_tetrahedra(p, np.array(tets) - 1)
    m.eo[:] = 1
    m.to[:] = 1
    m.teo[:] = 1
    assert orient3(m)
    return m


def kuhn_grid_arrays(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0, shear=(0.0, 0.0),
                     jitter_boundary=False):
    """Points and tetrahedra of an nx x ny x nz box of cubes, each cut into the
    six Kuhn tetrahedra 000 → +e_a → +e_a+e_b → 111 over the permutations
    (a, b, c) of the axes in lexicographic order; vertex (i, j, k) has id
    i + (nx+1)(j + (ny+1)k).  Interior points move by up to `jitter` cell
    widths (seeded), boundary points too with `jitter_boundary`; `shear` =
    (dx, dy) slides the top face as the reference's parallelepiped does.  A synthetic mesh: the reference makes
    its 3-D test meshes with TetGen, which is not available here."""
    gx, gy, gz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ijk = np.stack([gx.ravel(order="F"), gy.ravel(order="F"), gz.ravel(order="F")], 1)
    h = np.array([lx / nx, ly / ny, lz / nz])
    p = ijk * h
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((ijk > 0) & (ijk < [nx, ny, nz]), axis=1) | jitter_boundary
        p[inner] += (rng.random((inner.sum(), 3)) * 2 - 1) * jitter * h
    p[:, 0] += shear[0] * p[:, 2] / lz
    p[:, 1] += shear[1] * p[:, 2] / lz
    vid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    unit = np.eye(3, dtype=np.int64)
    tets = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                o = np.array([i, j, k])
                for a, b, c in itertools.permutations(range(3)):
                    q1 = o + unit[a]
                    q2 = q1 + unit[b]
                    q3 = q2 + unit[c]
                    tets.append((vid(*o), vid(*q1), vid(*q2), vid(*q3)))
    return p, np.array(tets, np.int64)


def kuhn_grid(nx, ny, nz, **kw) -> TetMesh:
    """EmbeddedDeltaSet3D(mesh) of the Kuhn grid: glue_sorted_tetrahedron! of
    each tetrahedron, then orient! (MeshInterop.jl:141-153)."""
    p, tets = kuhn_grid_arrays(nx, ny, nz, **kw)
    m = glue_tetrahedra(p, tets)
    assert orient3(m)
    return m


# ---- geometry ---------------------------------------------------------------

def _sqd(a, b):
    d = a - b
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def vol1(a, b):
    return np.sqrt(_sqd(a, b))


def vol2(a, b, c):
    u, w = b - a, c - a
    uu = (u * u).sum(-1)
    ww = (w * w).sum(-1)
    uw = (u * w).sum(-1)
    return np.sqrt(np.abs(uu * ww - uw ** 2)) / 2


def vol3(a, b, c, d):
    return np.abs(np.einsum("...k,...k->...", b - a, np.cross(c - a, d - a))) / 6


def _cayley_menger(P):
    """(..., n+2, n+2) Cayley-Menger matrices of (..., n+1, 3) points."""
    k = P.shape[-2]
    D = _sqd(P[..., :, None, :], P[..., None, :, :])
    CM = np.ones(P.shape[:-2] + (k + 1, k + 1))
    CM[..., 0, 0] = 0
    CM[..., 1:, 1:] = D
    return CM


def circumcenter(P):
    """geometric_center(points, Circumcenter()): barycentric coordinates from
    the first row of inv(cayley_menger(points...))."""
    lam = np.linalg.inv(_cayley_menger(P))[..., 0, 1:]
    return np.einsum("...i,...ik->...k", lam, P)


def barycenter(P):
    s = P[..., 0, :]
    for i in range(1, P.shape[-2]):
        s = s + P[..., i, :]
    return s / P.shape[-2]


class TetDual:
    """EmbeddedDeltaDualComplex3D(s) followed by subdivide_duals!(sd, center)."""

    # D_tetrahedron_schemas (DiscreteExteriorCalculus.jl:1184-1188): indices
    # into (v0..v3, e0..e5, t0..t3)
    SCHEMA = [(0, 8, 13), (0, 8, 11), (0, 7, 11), (0, 7, 12), (0, 9, 12), (0, 9, 13),
              (1, 9, 13), (1, 9, 12), (1, 5, 12), (1, 5, 10), (1, 6, 10), (1, 6, 13),
              (2, 6, 13), (2, 6, 10), (2, 4, 10), (2, 4, 11), (2, 8, 11), (2, 8, 13),
              (3, 7, 12), (3, 7, 11), (3, 4, 11), (3, 4, 10), (3, 5, 10), (3, 5, 12)]
    # D_triangle_schemas (v, e, ev) (DiscreteExteriorCalculus.jl:563)
    TRI_SCHEMA = [(0, 1, 1), (0, 2, 1), (1, 2, 0), (1, 0, 1), (2, 0, 0), (2, 1, 0)]

    def __init__(self, m: TetMesh, circumcenter_=True, geometry=True):
        # the dual complex holds its own copy of the primal parts (copy_parts!)
        self.m = m = copy.copy(m)
        m.eo, m.to, m.teo = m.eo.copy(), m.to.copy(), m.teo.copy()
        m.eo[m.eo < 0] = 0                    # zero(Bool) for unset edges (:311-321)
        m.to[m.to < 0] = 0                    # and triangles (:573-584) of a 3-D set
        if (m.teo < 0).any() and not orient3(m):
            raise ValueError("The 3-simplices of the given 3D delta set are non-orientable.")
        self.nv, self.ne, self.nt, self.ntet = m.nv, m.ne, m.nt, m.ntet
        self._build()
        if geometry:
            self._geometry(circumcenter_)

    # -- combinatorics -------------------------------------------------------
    def _build(self):
        m = self.m
        nv, ne, nt, ntet = self.nv, self.ne, self.nt, self.ntet
        self.vc, self.ec = np.arange(nv), nv + np.arange(ne)
        self.tc, self.tetc = nv + ne + np.arange(nt), nv + ne + nt + np.arange(ntet)
        self.ndv = nv + ne + nt + ntet
        Dv0, Dv1 = [], []
        De0, De1, De2 = [], [], []
        ekey, tkey = {}, {}

        def add_edges(tgt, src):
            first = len(Dv0)
            for s, t in zip(src.tolist(), tgt.tolist()):
                ekey.setdefault((s, t), len(Dv0))
                Dv0.append(t)
                Dv1.append(s)
            return first + np.arange(len(src))

        # 1-D: dual edges from ∂v_i(e) to the edge center (:301-309)
        E01 = [add_edges(self.ec, self.vc[m.v0]), add_edges(self.ec, self.vc[m.v1])]
        # 2-D: edge centers and vertex centers to the triangle center (:549-558)
        E2 = [m.e0, m.e1, m.e2]
        E12 = [add_edges(self.tc, self.ec[E2[j]]) for j in range(3)]
        tv = m.triangle_vertices()
        E02 = [add_edges(self.tc, self.vc[tv[:, j]]) for j in range(3)]
        for v, e, ev in self.TRI_SCHEMA:            # (:563-568), block k holds triangles k*nt + t
            a, b, c = E12[e], E02[v], E01[ev][E2[e]]
            for x, y, z in zip(a.tolist(), b.tolist(), c.tolist()):
                tkey.setdefault((x, y, z), len(De0))
                De0.append(x)
                De1.append(y)
                De2.append(z)

        def get_edge(s, t):                         # get_dual_edge! (:1118-1121)
            k = ekey.get((s, t))
            if k is None:
                k = ekey[(s, t)] = len(Dv0)
                Dv0.append(t)
                Dv1.append(s)
            return k

        def get_tri(a, b, c):                       # get_dual_triangle! (:1134-1150)
            ea, eb, ec = ekey.get((b, c)), ekey.get((a, c)), ekey.get((a, b))
            if ea is not None and eb is not None and ec is not None:
                k = tkey.get((ea, eb, ec))
                if k is not None:
                    return k
            x, y, z = get_edge(a, b), get_edge(b, c), get_edge(a, c)
            k = len(De0)
            tkey.setdefault((y, z, x), k)
            De0.append(y)
            De1.append(z)
            De2.append(x)
            return k

        Dt = []
        tvs, tes, tts = m.tetrahedron_vertices(), m.tetrahedron_edges(), m.tetrahedron_triangles()
        for tet in range(ntet):                     # (:1171-1195)
            dvs = np.r_[self.vc[tvs[tet]], self.ec[tes[tet]], self.tc[tts[tet]]].tolist()
            c = int(self.tetc[tet])
            for x, y, z in self.SCHEMA:             # glue_sorted_dual_tetrahedron!
                a, b, cc, d = sorted((dvs[x], dvs[y], dvs[z], c))
                Dt.append((get_tri(b, cc, d), get_tri(a, cc, d), get_tri(a, b, d), get_tri(a, b, cc)))
        self.D_v0, self.D_v1 = np.array(Dv0, np.int64), np.array(Dv1, np.int64)
        self.D_e0, self.D_e1, self.D_e2 = (np.array(x, np.int64) for x in (De0, De1, De2))
        Dt = np.array(Dt, np.int64).reshape(-1, 4)
        self.D_t0, self.D_t1, self.D_t2, self.D_t3 = Dt[:, 0], Dt[:, 1], Dt[:, 2], Dt[:, 3]
        self.nde, self.ndt, self.ndtet = len(Dv0), len(De0), len(Dt)

    @staticmethod
    def _incident(keys, value_count):
        """incident(s, x, path) for every x: ids grouped by key, ascending."""
        order = np.argsort(keys, kind="stable")
        ptr = np.searchsorted(keys[order], np.arange(value_count + 1))
        return [order[ptr[i]:ptr[i + 1]] for i in range(value_count)]

    def elementary_duals(self, k):
        """elementary_duals(Val(k), s, x) for every primal k-simplex x (:1040-1046)."""
        if k == 0:
            g = self._incident(self.D_v1[self.D_e1[self.D_t1]], self.ndv)
            return g[:self.nv]
        if k == 1:
            g = self._incident(self.D_v1[self.D_e1], self.ndv)
            return g[self.nv:self.nv + self.ne]
        if k == 2:
            g = self._incident(self.D_v1, self.ndv)
            return g[self.nv + self.ne:self.nv + self.ne + self.nt]
        return [np.array([c]) for c in self.tetc]

    def subsimplices(self, k):
        """subsimplices(Val(k), s, x) (:200-201, 462-463, 1034-1035)."""
        if k == 1:
            return self._incident(self.D_v0, self.ndv)[self.nv:self.nv + self.ne]
        if k == 2:
            return self._incident(self.D_v0[self.D_e1], self.ndv)[self.nv + self.ne:self.nv + self.ne + self.nt]
        return self._incident(self.D_v0[self.D_e1[self.D_t1]], self.ndv)[self.nv + self.ne + self.nt:]

    def primal_vertex(self, k, x):
        """primal_vertex(Val(k), s, x) (:203, 465-466, 1037-1038)."""
        x = np.asarray(x)
        if k >= 3:
            x = self.D_t1[x]
        if k >= 2:
            x = self.D_e2[x]
        return self.D_v1[x]

    def dual_triangle_vertices(self, dt):             # (:224-228)
        return np.stack([self.D_v1[self.D_e1[dt]], self.D_v1[self.D_e0[dt]], self.D_v0[self.D_e0[dt]]], -1)

    def dual_tetrahedron_vertices(self, dt):          # (:1053-1058)
        return np.stack([self.D_v1[self.D_e2[self.D_t2[dt]]], self.D_v0[self.D_e2[self.D_t2[dt]]],
                         self.D_v1[self.D_e0[self.D_t0[dt]]], self.D_v0[self.D_e0[self.D_t0[dt]]]], -1)

    # -- geometry -------------------------------------------------------------
    def _geometry(self, circ):
        m = self.m
        P = m.p
        center = circumcenter if circ else barycenter
        dp = np.zeros((self.ndv, 3))
        dp[:self.nv] = P                                               # (:389-391)
        dp[self.ec] = center(np.stack([P[m.v0], P[m.v1]], 1))          # edge_vertices = (∂v0, ∂v1)
        tv = m.triangle_vertices()
        dp[self.tc] = center(P[tv])
        self.tet_vertices = m.tetrahedron_vertices()
        dp[self.tetc] = center(P[self.tet_vertices])
        self.dual_point = dp
        self.length = vol1(P[m.v0], P[m.v1])
        self.area = vol2(P[tv[:, 0]], P[tv[:, 1]], P[tv[:, 2]])
        T = self.tet_vertices
        self.vol = vol3(P[T[:, 0]], P[T[:, 1]], P[T[:, 2]], P[T[:, 3]])
        self.dual_length = vol1(dp[self.D_v0], dp[self.D_v1])
        dtv = self.dual_triangle_vertices(np.arange(self.ndt))
        self.dual_area = vol2(dp[dtv[:, 0]], dp[dtv[:, 1]], dp[dtv[:, 2]])
        dtt = self.dual_tetrahedron_vertices(np.arange(self.ndtet))
        self.dual_vol = vol3(dp[dtt[:, 0]], dp[dtt[:, 1]], dp[dtt[:, 2]], dp[dtt[:, 3]])

    # -- signs ----------------------------------------------------------------
    def sign(self, k):
        o = [None, self.m.eo, self.m.to, self.m.teo][k]
        return np.where(o == 1, 1, -1).astype(np.int64)

    # -- matrices (exact integers) ----------------------------------------------
    def derivbound(self, k, transpose=False, negate=False):
        """dec_p_derivbound(Val(k)) (FastDEC.jl:470-548) as a CSR matrix."""
        m = self.m
        if k == 0:
            n, cols, s = self.ne, np.stack([m.v0, m.v1], 1), self.sign(1)
            V = np.stack([s, -s], 1)
            ncol = self.nv
        elif k == 1:
            n, cols = self.nt, np.stack([m.e0, m.e1, m.e2], 1)
            se, st = self.sign(1), self.sign(2)
            V = np.stack([se[m.e0] * st, -se[m.e1] * st, se[m.e2] * st], 1)
            ncol = self.ne
        else:
            n, cols = self.ntet, m.tetrahedron_triangles()
            st, sT = self.sign(2), self.sign(3)
            V = np.stack([st[cols[:, 0]] * sT, -st[cols[:, 1]] * sT, st[cols[:, 2]] * sT, -st[cols[:, 3]] * sT], 1)
            ncol = self.nt
        I = np.repeat(np.arange(n), cols.shape[1])
        J, V = cols.ravel(), V.ravel()
        if transpose:
            I, J, n, ncol = J, I, ncol, n
        if negate:
            V = -V
        return sp.csr_matrix((V.astype(np.float64), (I, J)), shape=(n, ncol))

    def differential(self, k):                         # FastDEC.jl:468-469
        return self.derivbound(k)

    def boundary(self, k):                             # FastDEC.jl:433-437
        return self.derivbound(k - 1, transpose=True)

    def dual_derivative(self, k):                      # FastDEC.jl:455-462
        return self.derivbound(2 - k, transpose=True, negate=(k != 1))

    # -- Hodge stars (FastDEC.jl:606-636, 721-727) ------------------------------
    def hodge_diag(self, k):
        if k == 0:
            return np.array([np.sum(self.dual_vol[d]) for d in self.elementary_duals(0)])
        if k == 1:
            return np.array([np.sum(self.dual_area[d]) for d in self.elementary_duals(1)]) / self.length
        if k == 2:
            return np.array([np.sum(self.dual_length[d]) for d in self.elementary_duals(2)]) / self.area
        return 1 / self.vol

    def inv_hodge_diag(self, k):
        mult = 1.0 if (k * (3 - k)) % 2 == 0 else -1.0
        return (1 / self.hodge_diag(k)) * mult

    # -- wedge products ----------------------------------------------------------
    def _dual_measure(self, k):
        return [self.dual_length, self.dual_area, self.dual_vol][k - 1]

    def _primal_measure(self, k):
        return [self.length, self.area, self.vol][k - 1]

    def wedge0k(self, k, f, a):
        """wedge_product_zero(Val(k), s, f, α, x) (DiscreteExteriorCalculus.jl:2087-2093)."""
        dm, pm = self._dual_measure(k), self._primal_measure(k)
        out = np.empty(len(pm))
        for x, subs in enumerate(self.subsimplices(k)):
            vs = self.primal_vertex(k, subs)
            out[x] = np.dot(dm[subs] / pm[x], f[vs]) * a[x]
        return out

    def wedge11(self, a, b):
        """∧(Val(1), Val(1), s, α, β, x) (DiscreteExteriorCalculus.jl:914-934)."""
        m = self.m
        s1, s2 = self.sign(1), self.sign(2)
        at_tc = self._incident(self.D_v0, self.ndv)
        co_e1 = self._incident(self.D_e1, self.nde)
        tv = m.triangle_vertices()
        out = np.empty(self.nt)
        for x in range(self.nt):
            dual_vs = self.vc[tv[x]]
            des = at_tc[self.tc[x]]
            # sort(..., by = e -> D_∂v1[e] .== dual_vs, rev=true)[1:3]
            keyed = sorted(des.tolist(), key=lambda e: tuple(self.D_v1[e] == dual_vs), reverse=True)[:3]
            ws = np.array([np.sum(self.dual_area[co_e1[e]]) for e in keyed]) / self.area[x]
            e0, e1, e2 = m.e0[x], m.e1[x], m.e2[x]
            a0, a1, a2 = a[[e0, e1, e2]]
            b0, b1, b2 = b[[e0, e1, e2]]
            form = s2[x] * np.dot(ws, [s1[e1] * s1[e2] * (b1 * a2 - a1 * b2),
                                       s1[e0] * s1[e2] * (b0 * a2 - a0 * b2),
                                       s1[e0] * s1[e1] * (b0 * a1 - a0 * b1)])
            out[x] = form / 2
        return out

    def _vertex_weights(self):
        """d_volume(d_tets ∩ elementary_duals(0, s, v)) / vol per vertex of each
        tetrahedron (DiscreteExteriorCalculus.jl:1338-1347)."""
        ed0 = self.elementary_duals(0)
        W = np.empty((self.ntet, 4))
        for x, subs in enumerate(self.subsimplices(3)):
            for i, v in enumerate(self.tet_vertices[x]):
                common = np.intersect1d(subs, ed0[v])
                W[x, i] = np.sum(self.dual_vol[common]) / self.vol[x]
        return W

    def wedge21(self, a, b):
        """∧(Val(2), Val(1), s, α, β, x): α a 2-form, β a 1-form (:1337-1378)."""
        m = self.m
        W = self._vertex_weights()
        s1, s2, s3 = self.sign(1), self.sign(2), self.sign(3)
        T, E = m.tetrahedron_triangles(), m.tetrahedron_edges()
        out = np.empty(self.ntet)
        for x in range(self.ntet):
            t0, t1, t2, t3 = T[x]
            e0, e1, e2, e3, e4, e5 = E[x]
            a0, a1, a2, a3 = a[[t0, t1, t2, t3]]
            b0, b1, b2, b3, b4, b5 = b[[e0, e1, e2, e3, e4, e5]]
            terms = [
                s1[e3] * s2[t3] * b3 * a3 + s1[e4] * s2[t2] * -b4 * a2 + s1[e5] * s2[t1] * b5 * a1,
                s1[e1] * s2[t3] * b1 * a3 + s1[e2] * s2[t2] * -b2 * a2 + s1[e5] * s2[t0] * b5 * a0,
                s1[e0] * s2[t3] * b0 * a3 + s1[e2] * s2[t1] * -b2 * a1 + s1[e4] * s2[t0] * b4 * a0,
                s1[e0] * s2[t2] * b0 * a2 + s1[e1] * s2[t1] * -b1 * a1 + s1[e3] * s2[t0] * b3 * a0]
            out[x] = s3[x] * np.dot(W[x], terms) / 3
        return out

    def wedge12(self, a, b):                           # (:1380-1381)
        return self.wedge21(b, a)

    # FastDEC's tables and kernels ------------------------------------------------
    def fast_coeffs03(self):
        """wedge_kernel_coeffs(Val(0), Val(3)) (FastDEC.jl:64-76): (verts, coeffs), 24 x ntet."""
        idx = 24 * np.arange(self.ntet)[None, :] + np.arange(24)[:, None]
        verts = self.D_v1[self.D_e2[self.D_t2[idx]]]
        return verts, self.dual_vol[idx] / self.vol[None, :]

    def fast_wedge03(self, f, a):
        """wedge_kernel_03! (FastDEC.jl:173-183): Σ_j α c_j f[p_j], j ascending."""
        p, c = self.fast_coeffs03()
        tmp = np.zeros(self.ntet)
        for j in range(24):
            tmp = tmp + a * c[j] * f[p[j]]
        return tmp

    def fast_coeffs21(self):
        """wedge_kernel_coeffs(Val(2), Val(1)) (FastDEC.jl:96-133): (ets 10 x ntet, coeffs 12 x ntet)."""
        m = self.m
        ws = self._vertex_weights() / 3
        E, T = m.tetrahedron_edges(), m.tetrahedron_triangles()
        s1, s2, s3 = self.sign(1), self.sign(2), self.sign(3)
        pairs = [(3, 3, 0), (4, 2, 0), (5, 1, 0), (1, 3, 1), (2, 2, 1), (5, 0, 1),
                 (0, 3, 2), (2, 1, 2), (4, 0, 2), (0, 2, 3), (1, 1, 3), (3, 0, 3)]
        C = np.stack([s3 * s1[E[:, e]] * s2[T[:, t]] * ws[:, v] for e, t, v in pairs], 0)
        return np.concatenate([E.T, T.T], 0), C

    def fast_wedge21(self, a, b):
        """wedge_kernel_21! (FastDEC.jl:194-208): α a 2-form, β a 1-form."""
        ets, c = self.fast_coeffs21()
        e, t = ets[:6], ets[6:]
        be = [b[e[i]] for i in range(6)]
        at = [a[t[i]] for i in range(4)]
        return (c[0] * be[3] * at[3] - c[1] * be[4] * at[2] + c[2] * be[5] * at[1] + c[3] * be[1] * at[3]
                - c[4] * be[2] * at[2] + c[5] * be[5] * at[0] + c[6] * be[0] * at[3] - c[7] * be[2] * at[1]
                + c[8] * be[4] * at[0] + c[9] * be[0] * at[2] - c[10] * be[1] * at[1] + c[11] * be[3] * at[0])

    # -- interpolation -------------------------------------------------------------
    def p3_d3_interpolation(self):
        """mat[v, x] += dual_vol / vol over x ascending and its 24 dual
        tetrahedra, v = D_∂v1[D_∂e2[D_∂t1[dual tetrahedron]]]
        (DiscreteExteriorCalculus.jl:2038-2053)."""
        acc = {}
        for x in range(self.ntet):
            V = self.vol[x]
            for dt in range(24 * x, 24 * x + 24):
                v = int(self.D_v1[self.D_e2[self.D_t1[dt]]])
                acc[(v, x)] = acc.get((v, x), 0.0) + self.dual_vol[dt] / V
        keys = np.array(list(acc.keys()), np.int64).reshape(-1, 2)
        return sp.csr_matrix((np.array(list(acc.values())), (keys[:, 0], keys[:, 1])), shape=(self.nv, self.ntet))

    def d0_p0_interpolation(self):
        """⋆0⁻¹ p3_d3 ⋆3⁻¹ (FastDEC.jl:924-926)."""
        return sp.diags(self.inv_hodge_diag(0)) @ self.p3_d3_interpolation() @ sp.diags(self.inv_hodge_diag(3))

    # -- Laplacians ---------------------------------------------------------------
    def laplacian0(self):
        """∇²(0, s) with the diagonal ⋆: -⋆0⁻¹ d̃2 ⋆1 d0 (DiscreteExteriorCalculus.jl:1836-1890)."""
        return -(sp.diags(self.inv_hodge_diag(0)) @ self.dual_derivative(2) @ sp.diags(self.hodge_diag(1))
                 @ self.differential(0))

    def dual_laplacian0(self):
        """Δᵈ(Val(0)) on a 3-D complex: ⋆3 d2 ⋆2⁻¹ d̃0 (FastDEC.jl:832-839)."""
        return (sp.diags(self.hodge_diag(3)) @ self.differential(2) @ sp.diags(self.inv_hodge_diag(2))
                @ self.dual_derivative(0))
