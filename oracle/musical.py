"""Musical isomorphisms and the operators built on them, restated in numpy.

TEST INFRASTRUCTURE ONLY: what the HIP library is tested against.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

Each function follows a reference routine of CombinatorialSpaces.jl step by
step on an oracle.Dual (0-based indices):

* ♯ (LLSDDSharp)        src/DiscreteExteriorCalculus.jl:868-909
* ♭ (DPPFlat)           src/DiscreteExteriorCalculus.jl:684-714
* ♭♯_mat                src/DiscreteExteriorCalculus.jl:1985-1992
* wedge pd/dp 1-1       src/FastDEC.jl:390-400
* geometric ⋆1⁻¹        src/FastDEC.jl:744-748 (a sparse LU solve of -⋆1)
* interior_product_dd   src/FastDEC.jl:758-786
* ℒ_dd                  src/FastDEC.jl:788-803
* Δᵈ (2-D)              src/FastDEC.jl:821-829, 835-860
* eval_constant_*_form  src/DiscreteExteriorCalculus.jl:2179-2193

pinv: the reference takes LinearAlgebra.pinv of the 3x3 matrix of dual-edge
vectors, keeping singular values above eps * 3 * σmax.  `project=False`
restates that literally.  The rows are tangent to the triangle, so on a flat
mesh the third singular value is exactly zero; on a curved surface rounding
leaves it near 1e-14 σmax, above that cut, and the literal pinv inverts it.
`project=True` first removes the normal component of each row (the
pseudo-inverse in the triangle's plane), which equals the literal form on
flat meshes and is what the HIP library computes everywhere.

Pinned by the reference's own tests on flat meshes (tests/test_oracle_musical.py):
♯ of a constant dual form returns the constant vector, ♭♯ of it returns the
constant primal form, the pd/dp 1-1 wedges are antisymmetric and vanish on
f ∧ f, d̃0 ι11(u, u) vanishes on interior edges for a constant flow, and
ι11(u, u) of u = ⋆(x dx) approximates x².
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle as O


def _csr(A: O.CSC):
    rp, ci, vv = A.csr()
    return sp.csr_matrix((vv, ci, rp), shape=(A.m, A.n))


def _tri_edges(d: O.Dual):
    m = d.m
    return np.stack([m.e0, m.e1, m.e2], axis=1).astype(np.int64)        # (nt, 3) slot j = ∂e_j


def _edge_dual_totals(d: O.Dual):
    """Σ dual_length over elementary_duals(1, s, e) (the B-block dual edges
    ending at the edge center), per edge."""
    nv, ne, nt = d.nv, d.ne, d.nt
    tot = np.zeros(ne)
    B = np.arange(2 * ne, 2 * ne + 3 * nt)
    np.add.at(tot, d.D_v1[B] - nv, d.dual_length[B])
    return tot


def _bdual(d: O.Dual):
    """(nt, 3) index of the dual edge from the center of t to the center of
    its edge in slot j (DualE block B_j, DiscreteExteriorCalculus.jl:551-554)."""
    nv, ne, nt = d.nv, d.ne, d.nt
    b = 2 * ne + np.arange(3)[None, :] * nt + np.arange(nt)[:, None]
    E = _tri_edges(d)
    assert np.array_equal(d.D_v0[b], np.broadcast_to(nv + ne + np.arange(nt)[:, None], b.shape))
    assert np.array_equal(d.D_v1[b], nv + E)
    return b


def _normals(d: O.Dual):
    m = d.m
    tv = m.triangle_vertices()
    p = m.p
    n = np.cross(p[tv[:, 1]] - p[tv[:, 0]], p[tv[:, 2]] - p[tv[:, 0]])
    return n / np.linalg.norm(n, axis=1)[:, None]


def edge_vectors(d: O.Dual):
    """(p[tgt] - p[src]) * sign(1, s, e) per edge."""
    m = d.m
    sgn = np.where(d.eo != 0, 1.0, -1.0)
    return (m.p[m.v0] - m.p[m.v1]) * sgn[:, None]


def sharp_lls(d: O.Dual, project=False):
    """♯_mat(s, LLSDDSharp()) as (edges, vectors): for triangle t the sorted
    edges E[t, i] and the 3-vectors ♯[t, E[t, i]]."""
    nv, ne, nt = d.nv, d.ne, d.nt
    E = _tri_edges(d)
    b = _bdual(d)
    order = np.argsort(E, axis=1, kind="stable")                          # sort(triangle_edges)
    Es = np.take_along_axis(E, order, axis=1)
    bs = np.take_along_axis(b, order, axis=1)
    sign = np.where(d.D_eo[bs] != 0, 1.0, -1.0)                           # de_sign
    X = sign[:, :, None] * (d.dual_point[d.D_v0[bs]] - d.dual_point[d.D_v1[bs]])
    tot = _edge_dual_totals(d)
    w = d.dual_length[bs] / tot[Es]
    if project:
        n = _normals(d)
        X = X - np.einsum("tik,tk->ti", X, n)[:, :, None] * n[:, None, :]
    L = np.linalg.pinv(X, rcond=np.finfo(float).eps * 3)                  # (nt, 3, 3)
    vec = np.transpose(L, (0, 2, 1)) * w[:, :, None]                      # LLS[:, i] * w_i
    return Es, vec


def flat_dpp(d: O.Dual):
    """♭_mat(s, DPPFlat()) as (edges, vectors): ♭[E[t, j], t] = del e_vec / dels_sum."""
    E = _tri_edges(d)
    b = _bdual(d)
    tot = _edge_dual_totals(d)
    ev = edge_vectors(d)
    vec = d.dual_length[b][:, :, None] * ev[E] / tot[E][:, :, None]
    return E, vec


def flat_sharp_mat(d: O.Dual, project=False):
    """♭♯_mat = only.(♭ * ♯): entry (e, e') = Σ_t ♭[e, t] · ♯[t, e'], summed
    over triangles in ascending order."""
    Ef, fl = flat_dpp(d)
    Es, sh = sharp_lls(d, project)
    nt = d.nt
    rows = np.repeat(Ef, 3, axis=1).ravel()
    cols = np.tile(Es, (1, 3)).ravel()
    vals = np.einsum("tik,tjk->tij", fl, sh).ravel()
    tri = np.repeat(np.arange(nt), 9)
    order = np.lexsort((tri, cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    key = rows * d.ne + cols
    start = np.r_[0, np.flatnonzero(np.diff(key)) + 1]
    summed = np.array([np.sum(vals[s:e]) for s, e in zip(start, np.r_[start[1:], len(vals)])])
    return sp.csr_matrix((summed, (rows[start], cols[start])), shape=(d.ne, d.ne))


def eval_constant_primal_form(d: O.Dual, alpha):
    return edge_vectors(d) @ np.asarray(alpha, float)


def eval_constant_dual_form(d: O.Dual, alpha):
    """hodge_star(1, s) * eval_constant_primal_form(s, (α2, -α1, α3)); the
    reference's ⋆1 default is the geometric one."""
    a = np.asarray(alpha, float)
    return _csr(d.geo_hodge1()) @ eval_constant_primal_form(d, (a[1], -a[0], a[2]))


class Ops:
    """The composite operators of one dual complex, with the sparse matrices
    and the ⋆1 factorisation built once."""

    def __init__(self, d: O.Dual, project=False):
        self.d = d
        self.FS = flat_sharp_mat(d, project)
        self.H1 = _csr(d.geo_hodge1())
        self.lu = sla.splu((-self.H1).tocsc())
        self.h2 = d.hodge_diag(2)
        self.ih0 = d.inv_hodge_diag(0)
        self.W01 = _csr(d.wedge_pd01_mat())
        self.dd0 = _csr(d.dual_derivative(0))
        self.dd1 = _csr(d.dual_derivative(1))
        self.d0 = _csr(d.differential(0))
        self.d1 = _csr(d.differential(1))

    def inv_hodge1(self, x):                      # x -> (-⋆1) \ x
        return self.lu.solve(np.asarray(x, float))

    def wedge_pd11(self, f, g):                   # (f, g) -> Λ11(f, ♭♯ g)
        return self.d.wedge11(f, self.FS @ g)

    def wedge_dp11(self, f, g):                   # (f, g) -> Λ11(♭♯ f, g)
        return self.d.wedge11(self.FS @ f, g)

    def interior11(self, f, g):                   # ⋆2 Λpd11(⋆1⁻¹ g, f)
        return self.h2 * self.wedge_pd11(self.inv_hodge1(g), f)

    def interior12(self, f, g):                   # ⋆1 ♭♯ ((Λ01 ⋆0⁻¹ g) .* f)
        return self.H1 @ (self.FS @ ((self.W01 @ (self.ih0 * g)) * f))

    def lie(self, f, g):                          # -(d̃0 ι11(f, g)) - ι12(f, d̃1 g)
        return -(self.dd0 @ self.interior11(f, g)) - self.interior12(f, self.dd1 @ g)

    def dual_laplacian0(self, x):                 # (⋆2 d1) ⋆1⁻¹(d̃0 x)
        return (sp.diags(self.h2) @ self.d1) @ self.inv_hodge1(self.dd0 @ x)

    def dual_laplacian1(self, x):                 # M x + N ⋆1⁻¹(x)
        M = self.H1 @ self.d0 @ sp.diags(self.ih0) @ self.dd1
        N = self.dd0 @ sp.diags(self.h2) @ self.d1
        return M @ x + N @ self.inv_hodge1(x)
