"""Write the golden fixtures of the DEC operators: the oracle's outputs on
small meshes for seeded forms, stored as .npz (no pickles).

    python tests/golden/make_fixtures.py

Each file holds the mesh (topology, orientations, points), the dual geometry
(lengths, areas; volumes for the tetrahedral file tet_*.npz), the seeded
input forms and every operator's output.  The
CPU tests check that the oracle still reproduces them and the GPU tests
compare the HIP library against them.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import interp as IP  # noqa: E402
import musical as MU  # noqa: E402
import oracle as O  # noqa: E402
import tet as TT  # noqa: E402


def outputs(mesh, circumcenter=True):
    d = O.Dual(mesh, circumcenter=circumcenter)
    nv, ne, nt = d.nv, d.ne, d.nt
    V, E, E2, T = (O.uniform_form(s, n) for s, n in ((1, nv), (2, ne), (3, ne), (4, nt)))
    DT, DE = O.uniform_form(5, nt), O.uniform_form(6, ne)   # dual 0-form (on triangles), dual 1-form
    out = {
        "v0": mesh.v0, "v1": mesh.v1, "e0": mesh.e0, "e1": mesh.e1, "e2": mesh.e2,
        "edge_orient": d.eo, "tri_orient": d.to, "points": mesh.p,
        "length": d.length, "area": d.area, "dual_length": d.dual_length, "dual_area": d.dual_area,
        "V": V, "E": E, "E2": E2, "T": T, "DT": DT, "DE": DE,
        "wedge01": d.wedge01(V, E), "wedge11": d.wedge11(E, E2), "wedge02": d.wedge02(V, T),
        "wedge_dd01": d.wedge_dd01(DT, DE), "wedge_dd10": d.wedge_dd10(DE, DT),
        "wedge_pd01": d.wedge_pd01(V, DE), "wedge_dp10": d.wedge_dp10(DE, V),
        "d0": d.differential(0).matvec(V), "d1": d.differential(1).matvec(E),
        "boundary1": d.boundary(1).matvec(E), "boundary2": d.boundary(2).matvec(T),
        "dual_d0": d.dual_derivative(0).matvec(DT), "dual_d1": d.dual_derivative(1).matvec(DE),
        "hodge0": d.hodge_diag(0) * V, "hodge1": d.hodge_diag(1) * E, "hodge2": d.hodge_diag(2) * T,
        "inv_hodge0": d.inv_hodge_diag(0) * V, "inv_hodge1": d.inv_hodge_diag(1) * E,
        "inv_hodge2": d.inv_hodge_diag(2) * T,
        "geo_hodge1": d.geo_hodge1().matvec(E),
        "laplacian0": d.laplacian0().matvec(V), "laplacian0_geometric": d.laplacian0(geometric=True).matvec(V),
    }
    # the musical isomorphisms and their composites (numpy oracle; ♯ in the
    # triangle's plane, as the library takes it on curved surfaces)
    ops = MU.Ops(d, project=True)
    out.update({
        "flat_sharp": ops.FS @ DE, "wedge_pd11": ops.wedge_pd11(E, DE), "wedge_dp11": ops.wedge_dp11(DE, E),
        "inv_hodge1_geometric": ops.inv_hodge1(E), "interior11": ops.interior11(DE, E2),
        "interior12": ops.interior12(DE, V), "lie": ops.lie(DE, E2),
        "dual_laplacian0": ops.dual_laplacian0(DT), "dual_laplacian1": ops.dual_laplacian1(DE),
    })
    return {k: np.ascontiguousarray(v) for k, v in out.items()}


def _mv(A, x):
    """A x summed over each row's columns in ascending order (the library's order)."""
    A = A.tocsr()
    A.sort_indices()
    return A @ x


def outputs3(mesh, circumcenter=False):
    """The 3-D operators of a tetrahedral complex (oracle/tet.py) on seeded forms."""
    d = TT.TetDual(mesh, circumcenter_=circumcenter)
    m = d.m
    V, E, T, X = (O.uniform_form(s, n) for s, n in ((11, d.nv), (12, d.ne), (13, d.nt), (14, d.ntet)))
    out = {
        "v0": m.v0, "v1": m.v1, "e0": m.e0, "e1": m.e1, "e2": m.e2, "t0": m.t0, "t1": m.t1, "t2": m.t2,
        "t3": m.t3, "edge_orient": m.eo, "tri_orient": m.to, "tet_orient": m.teo, "points": m.p,
        "vol": d.vol, "dual_vol": d.dual_vol, "V": V, "E": E, "T": T, "X": X,
        "wedge01": d.wedge0k(1, V, E), "wedge02": d.wedge0k(2, V, T), "wedge03": d.wedge0k(3, V, X),
        "wedge11": d.wedge11(E, E), "wedge21": d.wedge21(T, E),
        "d0": _mv(d.differential(0), V), "d1": _mv(d.differential(1), E), "d2": _mv(d.differential(2), T),
        "boundary3": _mv(d.boundary(3), X), "dual_d0": _mv(d.dual_derivative(0), X),
        "dual_d1": _mv(d.dual_derivative(1), T), "dual_d2": _mv(d.dual_derivative(2), E),
        "hodge0": d.hodge_diag(0) * V, "hodge1": d.hodge_diag(1) * E, "hodge2": d.hodge_diag(2) * T,
        "hodge3": d.hodge_diag(3) * X, "laplacian0": d.laplacian0() @ V, "dual_laplacian0": d.dual_laplacian0() @ X,
        "p3_d3": d.p3_d3_interpolation() @ X, "d0_p0": d.d0_p0_interpolation() @ X,
        "p0_d0": IP.p0_d0_interpolation(d) @ V,
    }
    return {k: np.ascontiguousarray(v) for k, v in out.items()}


# outputs computed through LAPACK / SuperLU (pinv, splu): compared to 1e-13
# relative rather than bit for bit when the oracle is re-run
LAPACK_KEYS = {"flat_sharp", "wedge_pd11", "wedge_dp11", "inv_hodge1_geometric", "interior11", "interior12", "lie",
               "dual_laplacian0", "dual_laplacian1"}


def main():
    np.savez(os.path.join(HERE, "icosphere2_circumcenter.npz"), **outputs(O.icosphere(2)))
    grid = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, compress=False)
    np.savez(os.path.join(HERE, "grid4_barycenter.npz"), **outputs(grid, circumcenter=False))
    km = TT.kuhn_grid(2, 2, 1, lx=2.0, ly=2.0, lz=1.0, jitter=0.2, seed=4)
    np.savez(os.path.join(HERE, "tet_kuhn221_barycenter.npz"), **outputs3(km, circumcenter=False))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
