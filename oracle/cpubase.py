"""CPU baseline of the operator set (benchmark and test infrastructure only):
the oracle's operator loops built with -O3 and OpenMP (cpu_baseline.c), run
at a chosen thread count on the oracle's own tables and forms."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int

import numpy as np

import oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpubase.so")
_lib = None

_I, _D = POINTER(c_int), POINTER(c_double)


class _OpSet(ctypes.Structure):
    _fields_ = [("nv", c_int), ("ne", c_int), ("nt", c_int),
                ("v0", _I), ("v1", _I), ("e11", _I), ("c11", _D), ("v02", _I), ("c02", _D),
                ("d0p", _I), ("d0c", _I), ("d0v", _D), ("d1p", _I), ("d1c", _I), ("d1v", _D),
                ("Lp", _I), ("Lc", _I), ("Lv", _D), ("h0", _D), ("h1", _D), ("h2", _D),
                ("V", _D), ("E", _D), ("E2", _D), ("T", _D), ("oV", _D), ("oE", _D), ("oT", _D)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            O.build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.cb_run.restype = c_double
    return _lib


NAMES = ["wedge01", "wedge11", "wedge02", "d0", "d1", "hodge0", "hodge1", "hodge2", "laplacian0"]


class OpSet:
    """The nine operators of the benchmark's step on the oracle's dual complex
    `d`, with the seeded forms V, E, E2, T (seeds 1-4, as the GPU leg uses)."""

    def __init__(self, d: O.Dual):
        self.d = d
        nv, ne, nt = d.nv, d.ne, d.nt
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data_as(_I if dt == np.int32 else _D)

        def csr(A: O.CSC):
            rp, ci, vv = A.csr()
            return arr(rp, np.int32), arr(ci, np.int32), arr(vv, np.float64)

        v02, c02 = d.wedge02_coeffs()
        e11, c11 = d.wedge11_coeffs()
        self.V, self.E, self.E2, self.T = (O.uniform_form(s, n) for s, n in ((1, nv), (2, ne), (3, ne), (4, nt)))
        self.oV, self.oE, self.oT = np.zeros(nv), np.zeros(ne), np.zeros(nt)
        self.L = d.laplacian0()
        s = _OpSet(nv, ne, nt, arr(d.m.v0, np.int32), arr(d.m.v1, np.int32), arr(e11, np.int32), arr(c11, np.float64),
                   arr(v02, np.int32), arr(c02, np.float64), *csr(d.differential(0)), *csr(d.differential(1)),
                   *csr(self.L), *(arr(d.hodge_diag(k), np.float64) for k in range(3)),
                   *(arr(x, np.float64) for x in (self.V, self.E, self.E2, self.T)),
                   self.oV.ctypes.data_as(_D), self.oE.ctypes.data_as(_D), self.oT.ctypes.data_as(_D))
        self._keep = keep
        self._s = s

    def run(self, threads: int, passes: int) -> float:
        """Wall seconds for `passes` passes of the nine operators."""
        return lib().cb_run(ctypes.byref(self._s), c_int(threads), c_int(passes))

    def apply(self, name: str, threads: int = 1) -> np.ndarray:
        w = NAMES.index(name)
        lib().cb_apply(ctypes.byref(self._s), c_int(threads), c_int(w))
        out = (self.oE, self.oT, self.oT, self.oE, self.oT, self.oV, self.oE, self.oT, self.oV)[w]
        return out.copy()
