"""Learned multigrid smoother: a 3-layer edge message-passing network trained
on the GPU with the library's own forward / backward / Adam kernels.

The network is bound to a MultigridData hierarchy (one operator per level);
its parameters are level independent, so a network trained on one hierarchy
can be loaded into another with `set_params`.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import c_double, c_int32, c_void_p

import numpy as np

from ._lib import check, lib, ptr
from .fastdec import _stream
from .multigrid import MultigridData


class LearnedSmoother:
    def __init__(self, md: MultigridData, seed: int = 0):
        self.md = md
        h = c_void_p()
        check(lib().dec_nn_create(md.handle, seed, ctypes.byref(h)))
        self._h = h
        n = c_int32()
        check(lib().dec_nn_nparams(h, ctypes.byref(n)))
        self.nparams = n.value

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().dec_mg_set_learned(self.md.handle, None, 0, 0, 0, 1)
                lib().dec_nn_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        check(lib().dec_nn_get_params(self._h, ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(lib().dec_nn_set_params(self._h, ptr(p)))

    def _R(self, R, level):
        import torch

        n = self.md.sizes[level]
        if not (isinstance(R, torch.Tensor) and R.is_cuda and R.dtype == torch.float32 and R.is_contiguous()):
            raise TypeError("residuals must be a contiguous float32 GPU tensor")
        if R.numel() % n:
            raise ValueError(f"residuals must hold a whole number of level-{level} vectors ({n} entries)")
        return R.numel() // n

    def forward(self, level, R, precise=False, stream=None):
        import torch

        B = self._R(R, level)
        c = torch.empty_like(R)
        check(lib().dec_nn_forward(self._h, level, B, c_void_p(R.data_ptr()), c_void_p(c.data_ptr()),
                                   int(bool(precise)), _stream(stream)))
        return c

    def smooth(self, level, r, u, stream=None):
        """u += c(r) for one residual (float32 GPU vectors of the level's
        size), with the inference kernels of the V-cycles."""
        n = self.md.sizes[level]
        for t in (r, u):
            if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 4 and t.is_contiguous()
                    and t.numel() == n):
                raise TypeError(f"r and u must be contiguous float32 GPU vectors of {n} entries")
        check(lib().dec_nn_smooth(self._h, level, c_void_p(r.data_ptr()), c_void_p(u.data_ptr()), _stream(stream)))
        return u

    def _E(self, E, R):
        if E is None:
            return None
        if not (E.is_cuda and E.dtype == R.dtype and E.is_contiguous() and E.shape == R.shape):
            raise TypeError("errors must be a contiguous float32 GPU tensor shaped like the residuals")
        return c_void_p(E.data_ptr())

    def loss_grad(self, level, R, E=None, precise=False, stream=None):
        """Loss and gradient on the residuals R (residual loss) or on the
        (residual, error) pairs (R, E) (error loss, high-pass filtered)."""
        B = self._R(R, level)
        loss = c_double()
        g = np.zeros(self.nparams, np.float32)
        check(lib().dec_nn_loss_grad(self._h, level, B, c_void_p(R.data_ptr()), self._E(E, R), int(bool(precise)),
                                     ctypes.byref(loss), ptr(g), _stream(stream)))
        return loss.value, g

    def train(self, level, R, steps, E=None, batch=16, lr=1e-3, seed=0, stream=None):
        npairs = self._R(R, level)
        hist = np.zeros(max(steps, 1))
        check(lib().dec_nn_train(self._h, level, npairs, c_void_p(R.data_ptr()), self._E(E, R), steps, batch,
                                 float(lr), seed, ptr(hist), _stream(stream)))
        return hist[:steps]

    def attach(self, enable=True, min_rows=16385, max_rows=0, applications=1):
        """Use the network as the smoother of the float32 V-cycles of md on the
        levels of min_rows to max_rows rows (max_rows 0: no upper bound)."""
        check(lib().dec_mg_set_learned(self.md.handle, self._h if enable else None, int(bool(enable)), min_rows,
                                       max_rows, applications))


def learned_stats(md: MultigridData):
    f, c = c_int32(), c_int32()
    check(lib().dec_mg_learned_stats(md.handle, ctypes.byref(f), ctypes.byref(c)))
    return {"fallbacks": f.value, "learned_cycles": c.value}


def sh_basis(points, lmax=12):
    """Real orthonormal spherical harmonics Y_lm (1 ≤ l ≤ lmax) at the points."""
    from scipy.special import sph_harm_y

    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    cols, lm = [], []
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            Y = sph_harm_y(l, abs(m), theta, phi)
            if m > 0:
                cols.append(math.sqrt(2) * (-1) ** m * Y.real)
            elif m < 0:
                cols.append(math.sqrt(2) * (-1) ** m * Y.imag)
            else:
                cols.append(Y.real)
            lm.append((l, m))
    return np.stack(cols, axis=1), lm


def make_pairs(md: MultigridData, npairs, cycles_per_rhs=8, nterms=16, seed=0, random_start=True):
    """(residual, error) pairs from plain float64 V(2,2) Gauss-Seidel cycles on
    seeded spherical-harmonic right-hand sides (16 terms, l ≤ 12, ⋆0-weighted
    mean zero) on the finest level of md, one pair per iterate.  The cycles
    start from a seeded uniform random iterate scaled to the solution's RMS
    (random_start) or from zero; a random start gives the early iterates
    errors of every frequency, as a smoother meets them.  Returns float32
    (R, E) of shape (npairs, n) on the GPU."""
    import torch

    from . import fastdec as fd
    from . import multigrid as mgm

    fine = md.finest_mesh()
    n = fine.nv
    L = fd.fast_laplace_beltrami(fine, fd.DiagonalHodge)
    ones = torch.ones(n, dtype=torch.float64, device="cuda")
    h0 = (fd.dec_hodge_star(0, fine, fd.DiagonalHodge) @ ones)
    Y, _ = sh_basis(fine.export()["points"])
    Yd = torch.tensor(Y, device="cuda")
    rng = np.random.default_rng(seed)
    R = torch.empty((npairs, n), dtype=torch.float32, device="cuda")
    E = torch.empty((npairs, n), dtype=torch.float32, device="cuda")
    k = 0
    while k < npairs:
        idx = rng.choice(Y.shape[1], nterms, replace=False)
        c = torch.tensor(rng.uniform(-1, 1, nterms), device="cuda")
        f = Yd[:, idx] @ c
        f = f - (h0 @ f) / h0.sum()
        ustar, _, _ = mgm.solve(md, f, tol=1e-11, variant="plain", max_cycles=100)
        u = torch.zeros(n, dtype=torch.float64, device="cuda")
        if random_start:
            fd.fill_uniform(u, int(rng.integers(1 << 30)))
            u *= float(ustar.norm()) / float(u.norm())
        for _ in range(cycles_per_rhs):
            if k >= npairs:
                break
            R[k] = (f - L @ u).float()
            E[k] = (ustar - u).float()
            k += 1
            mgm.multigrid_vcycles(u, f, md, 1, mgm.gauss_seidel, steps=2, coarse_direct=True)
    return R, E
