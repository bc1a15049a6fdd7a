"""Geometric multigrid on the GPU with CombinatorialSpaces.jl's Multigrid API.

Mirrors src/Multigrid.jl: `MultigridData(s, BinarySubdivision(), levels, op,
steps)` in DirectMode, `multigrid_vcycles`, `multigrid_wcycles`,
`full_multigrid`, and FastDEC's `dec_Δ⁻¹` (here `dec_Delta_inv`).  `op` is the
0-form Laplace-Beltrami operator (`fast_laplace_beltrami`) with the chosen
Hodge star for ⋆1.  `alg` selects the smoother: `cg` (Krylov.jl's cg, the
reference default), `gauss_seidel` (multicolour), `block_gauss_seidel`
(hybrid block GS) or `jacobi`.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int32, c_int64, c_void_p

import numpy as np

from . import _lib as L
from ._lib import check, lib, ptr
from .fastdec import Circumcenter, GeometricHodge, Mesh, _hodge, _stream


class BinarySubdivision:
    pass


class _Smoother:
    def __init__(self, name, code):
        self.name, self.code = name, code

    def __repr__(self):
        return self.name


cg = _Smoother("cg", L.DEC_SMOOTH_CG)
gauss_seidel = _Smoother("gauss_seidel", L.DEC_SMOOTH_GS)
jacobi = _Smoother("jacobi", L.DEC_SMOOTH_JACOBI)
# hybrid block Gauss-Seidel (this library's GPU smoother): multicolour GS
# inside blocks of the level, neighbours in other blocks from the start of the
# sweep; one launch per sweep
block_gauss_seidel = _Smoother("block_gauss_seidel", L.DEC_SMOOTH_BLOCK_GS)


class DirectMode:
    """Every level dualized and discretized (Multigrid.jl:566)."""
    code = L.DEC_MG_DIRECT


class GalerkinMode:
    """Only the finest level dualized; Aₖ₊₁ = Rₖ Aₖ Pₖ (Multigrid.jl:567, 740-748)."""
    code = L.DEC_MG_GALERKIN


class MultigridData:
    """MultigridData(s, BinarySubdivision(), levels, fast_laplace_beltrami, steps;
    alg=Circumcenter(), mode=DirectMode()) (Multigrid.jl:698-702, 752-797).
    Levels are finest first."""

    def __init__(self, s: Mesh, scheme=BinarySubdivision, levels: int = 1, hodge=GeometricHodge,
                 steps: int = 3, alg=Circumcenter, project: bool = False, mode=DirectMode):
        if scheme is not BinarySubdivision:
            raise ValueError("only BinarySubdivision is implemented")
        h = c_void_p()
        center = alg.code if hasattr(alg, "code") else int(alg)
        check(lib().dec_mg_create_mode(s.handle, levels, int(bool(project)), _hodge(hodge), center, steps,
                                       mode.code, ctypes.byref(h)))
        self._h = h
        self.steps = steps
        n = c_int32()
        check(lib().dec_mg_nlevels(h, ctypes.byref(n)))
        self.nlevels = n.value
        self.sizes = []
        for lv in range(self.nlevels):
            check(lib().dec_mg_level_size(h, lv, ctypes.byref(n)))
            self.sizes.append(n.value)
        self._alg = None

    @classmethod
    def from_matrices(cls, operators, restrictions, prolongations, steps: int = 3):
        """MultigridData(ops, rs, ps, steps) (Multigrid.jl:545-546): a hierarchy
        from host matrices, finest first.  Each matrix is a scipy sparse
        matrix or a CSR triple (rowptr, col, val), 0-based."""
        import ctypes as C

        def csr(M):
            if hasattr(M, "tocsr"):
                M = M.tocsr()
                return (np.ascontiguousarray(M.indptr, np.int32), np.ascontiguousarray(M.indices, np.int32),
                        np.ascontiguousarray(M.data, np.float64))
            rp, ci, vv = M
            return (np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
                    np.ascontiguousarray(vv, np.float64))

        A = [csr(M) for M in operators]
        R = [csr(M) for M in restrictions]
        P = [csr(M) for M in prolongations]
        nl = len(A)
        if len(R) != nl - 1 or len(P) != nl - 1:
            raise ValueError("need one restriction and one prolongation per pair of levels")
        sizes = np.array([len(a[0]) - 1 for a in A], np.int32)

        def arr(ms, j):
            return (C.c_void_p * max(len(ms), 1))(*[m[j].ctypes.data for m in ms])

        h = c_void_p()
        check(lib().dec_mg_create_from_csr(nl, ptr(sizes), arr(A, 0), arr(A, 1), arr(A, 2), arr(R, 0), arr(R, 1),
                                           arr(R, 2), arr(P, 0), arr(P, 1), arr(P, 2), 0, steps, ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.steps = steps
        self.nlevels = nl
        self.sizes = [int(x) for x in sizes]
        self._alg = None
        self._keep = (A, R, P)
        return self

    def __len__(self):
        return self.nlevels

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().dec_mg_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def finest_mesh(self) -> Mesh:
        h = c_void_p()
        check(lib().dec_mg_finest_mesh(self._h, ctypes.byref(h)))
        return Mesh(h, owned=False)

    def set_smoother(self, alg=cg, steps=None, omega=2 / 3, coarse_direct=False):
        steps = self.steps if steps is None else steps
        check(lib().dec_mg_set_smoother(self._h, alg.code, steps, c_double(omega), int(bool(coarse_direct))))
        self._alg = (alg, steps, omega, coarse_direct)

    def export(self, level: int, which: str = "A"):
        """CSR (rowptr, col, val) of operators[level] ('A'), restrictions[level]
        ('R') or prolongations[level] ('P')."""
        w = {"A": 0, "R": 1, "P": 2}[which]
        nnz = c_int64()
        check(lib().dec_mg_export_level(self._h, level, w, ctypes.byref(nnz), None, None, None))
        n = self.sizes[level]
        rows = {"A": n, "R": self.sizes[level + 1] if w == 1 else 0, "P": n}[which]
        rp = np.zeros(rows + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        vv = np.zeros(nnz.value)
        check(lib().dec_mg_export_level(self._h, level, w, None, ptr(rp), ptr(ci), ptr(vv)))
        return rp, ci, vv

    def colors(self, level: int):
        """Gauss-Seidel colour of each vertex of operators[level] and the
        number of colours."""
        c = np.zeros(self.sizes[level], np.int32)
        k = c_int32()
        check(lib().dec_mg_export_colors(self._h, level, ptr(c), ctypes.byref(k)))
        return c, k.value

    def blocks(self, level: int):
        """Block of each vertex of operators[level] for the hybrid block
        Gauss-Seidel smoother, and the number of blocks."""
        c = np.zeros(self.sizes[level], np.int32)
        k = c_int32()
        check(lib().dec_mg_export_blocks(self._h, level, ptr(c), ctypes.byref(k)))
        return c, k.value


MGData = MultigridData


def _dev(x, n, name):
    import torch

    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float64:
        raise TypeError(f"{name} must be a float64 GPU tensor")
    if x.numel() != n or not x.is_contiguous():
        raise ValueError(f"{name} must be contiguous with {n} entries")
    return c_void_p(x.data_ptr())


def _prepare(md, alg, steps=None, coarse_direct=None, omega=None):
    cur = md._alg
    want = (alg, md.steps if steps is None else steps,
            (cur[2] if cur else 2 / 3) if omega is None else omega,
            (cur[3] if cur else False) if coarse_direct is None else coarse_direct)
    if cur != want:
        md.set_smoother(*want)


def multigrid_mu_cycles(u, b, md: MultigridData, cycles, alg=cg, mu=1, stream=None, **kw):
    _prepare(md, alg, **kw)
    n = md.sizes[0]
    check(lib().dec_mg_cycles(md.handle, mu, cycles, _dev(u, n, "u"), _dev(b, n, "b"), _stream(stream)))
    return u


def multigrid_vcycles(u, b, md: MultigridData, cycles, alg=cg, **kw):
    """multigrid_vcycles(u, b, md, cycles, alg=cg) (Multigrid.jl:814); u is updated in place."""
    return multigrid_mu_cycles(u, b, md, cycles, alg, 1, **kw)


def multigrid_wcycles(u, b, md: MultigridData, cycles, alg=cg, **kw):
    """multigrid_wcycles (Multigrid.jl:817)."""
    return multigrid_mu_cycles(u, b, md, cycles, alg, 2, **kw)


def full_multigrid(b, md: MultigridData, cycles, alg=cg, mu=1, stream=None, **kw):
    """full_multigrid(b, md, cycles, alg, μ) (Multigrid.jl:830-837)."""
    import torch

    _prepare(md, alg, **kw)
    n = md.sizes[0]
    u = torch.zeros(n, dtype=torch.float64, device=b.device)
    check(lib().dec_mg_full_multigrid(md.handle, mu, cycles, _dev(b, n, "b"), _dev(u, n, "u"), _stream(stream)))
    return u


def dec_Delta_inv(md: MultigridData, steps=3, cycles=5, alg=cg, mu=2):
    """dec_Δ⁻¹(Val(0), series; steps=3, cycles=5, alg=cg, μ=2) (FastDEC.jl:866-869):
    b -> full_multigrid(b, MGData(series, ∇²₀, steps), cycles, alg, μ)."""
    return lambda b: full_multigrid(b, md, cycles, alg, mu, steps=steps)


def solve(md: MultigridData, b, u=None, tol=1e-8, max_cycles=200, variant="plain", alg=block_gauss_seidel,
          steps=2, coarse_direct=True, stream=None):
    """Solve A u = b to ‖b - A u‖₂/‖b‖₂ ≤ tol (checked in float64) with float64
    V-cycles ("plain"), float32 V-cycles inside float64 iterative refinement
    ("mixed"), or float64 conjugate gradients preconditioned by one float64
    ("pcg") or float32 ("pcg_mixed") V-cycle.  Returns (u, cycles or CG
    iterations, relative residual)."""
    import torch

    _prepare(md, alg, steps=steps, coarse_direct=coarse_direct)
    n = md.sizes[0]
    if u is None:
        u = torch.zeros(n, dtype=torch.float64, device=b.device)
    v = {"plain": L.DEC_SOLVE_PLAIN_F64, "mixed": L.DEC_SOLVE_MIXED_IR, "pcg": L.DEC_SOLVE_PCG_F64,
         "pcg_mixed": L.DEC_SOLVE_PCG_MIXED}[variant]
    c = c_int32()
    r = c_double()
    check(lib().dec_mg_solve(md.handle, v, c_double(tol), max_cycles, _dev(u, n, "u"), _dev(b, n, "b"),
                             ctypes.byref(c), ctypes.byref(r), _stream(stream)))
    return u, c.value, r.value
