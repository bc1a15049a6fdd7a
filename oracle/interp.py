"""Interpolation matrices between primal and dual 0-forms / top forms of a
2-D complex, restated in numpy over an oracle.Dual (0-based indices).

TEST INFRASTRUCTURE ONLY: what the HIP library is tested against.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

* p2_d2_interpolation   src/DiscreteExteriorCalculus.jl:2016-2031
* p0_d0_interpolation   src/FastDEC.jl:928-941
* d0_p0_interpolation   src/FastDEC.jl:915-917 (⋆0⁻¹ p2_d2 ⋆2⁻¹, both stars diagonal)

Each loop runs in the reference's order and accumulates repeated entries with
`+=` in that order.  Pinned by the reference's own interpolation tests
(test/DiscreteExteriorCalculus.jl:780-888, tests/test_oracle_interp.py).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

import oracle as O


def _accumulate(entries, shape):
    acc = {}
    for key, w in entries:
        acc[key] = acc.get(key, 0.0) + w
    if not acc:
        return sp.csr_matrix(shape)
    keys = np.array(list(acc.keys()), np.int64)
    return sp.csr_matrix((np.array(list(acc.values())), (keys[:, 0], keys[:, 1])), shape=shape)


def p2_d2_interpolation(d: O.Dual):
    """mat[v, t] += dual_area / area over t ascending and its dual triangles
    t, t + nt, ..., v = D_∂v1[D_∂e1[dual triangle]]."""
    nt = d.nt

    def gen():
        for t in range(nt):
            A = d.area[t]
            for dt in range(t, 6 * nt, nt):
                yield (int(d.D_v1[d.D_e1[dt]]), t), d.dual_area[dt] / A

    return _accumulate(gen(), (d.nv, nt))


def p0_d0_interpolation(d: O.Dual):
    """m[t, v] += dual_area / area, v the first vertex of the dual triangle."""
    nt = d.nt

    def gen():
        for t in range(nt):
            A = d.area[t]
            for i in range(6):
                dt = t + i * nt
                v = d.dual_triangle_vertices(dt)[0]
                yield (t, int(v)), d.dual_area[dt] / A

    return _accumulate(gen(), (nt, d.nv))


def d0_p0_interpolation(d: O.Dual):
    return sp.diags(d.inv_hodge_diag(0)) @ p2_d2_interpolation(d) @ sp.diags(d.inv_hodge_diag(2))
