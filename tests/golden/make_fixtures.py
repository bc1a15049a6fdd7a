"""Write the golden fixtures of the DEC operators: the oracle's outputs on
small meshes for seeded forms, stored as .npz (no pickles).

    python tests/golden/make_fixtures.py

Each file holds the mesh (topology, orientations, points), the dual geometry
(lengths, areas), the seeded input forms and every operator's output.  The
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

import musical as MU  # noqa: E402
import oracle as O  # noqa: E402


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


# outputs computed through LAPACK / SuperLU (pinv, splu): compared to 1e-13
# relative rather than bit for bit when the oracle is re-run
LAPACK_KEYS = {"flat_sharp", "wedge_pd11", "wedge_dp11", "inv_hodge1_geometric", "interior11", "interior12", "lie",
               "dual_laplacian0", "dual_laplacian1"}


def main():
    np.savez(os.path.join(HERE, "icosphere2_circumcenter.npz"), **outputs(O.icosphere(2)))
    grid = O.triangulated_grid(1, 1, 1 / 4, math.sqrt(3) / 2 * 1 / 4, compress=False)
    np.savez(os.path.join(HERE, "grid4_barycenter.npz"), **outputs(grid, circumcenter=False))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
