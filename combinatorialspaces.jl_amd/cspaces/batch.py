"""Batched explicit heat steps over many meshes (one persistent launch per
step, optionally replayed from a hipGraph; launch shape from an autotuner)."""
from __future__ import annotations

import ctypes
from ctypes import c_int32, c_int64, c_void_p

import numpy as np

from . import _lib as L
from ._lib import check, lib, ptr
from .fastdec import _dtype_code, _stream, _torch_dtype


class HeatBatch:
    def __init__(self, meshes, cast_float=None):
        self.meshes = list(meshes)            # keep the meshes alive
        self.dtype_code = _dtype_code(cast_float)
        self.dtype = _torch_dtype(self.dtype_code)
        arr = (c_void_p * len(self.meshes))(*[m.handle.value for m in self.meshes])
        h = c_void_p()
        check(lib().dec_batch_create(len(self.meshes), arr, self.dtype_code, ctypes.byref(h)))
        self._h = h
        total = c_int64()
        self.offsets = np.zeros(len(self.meshes) + 1, np.int64)
        self.dt = np.zeros(len(self.meshes))
        check(lib().dec_batch_layout(h, ctypes.byref(total), ptr(self.offsets), ptr(self.dt)))
        self.total = total.value

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().dec_batch_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def _u(self, u):
        import torch

        if not isinstance(u, torch.Tensor) or not u.is_cuda or u.dtype != self.dtype or \
                u.numel() != self.total or not u.is_contiguous():
            raise ValueError(f"state must be a contiguous {self.dtype} GPU tensor of {self.total} entries")
        return c_void_p(u.data_ptr())

    def tune(self, u, cache_path=None):
        b, r, g = c_int32(), c_int32(), c_int32()
        check(lib().dec_batch_tune(self._h, self._u(u), cache_path.encode() if cache_path else None,
                                   ctypes.byref(b), ctypes.byref(r), ctypes.byref(g)))
        return {"block": b.value, "rows_per_thread": r.value, "grid_mult": g.value}

    def steps(self, u, nsteps, graph=True, stream=None):
        check(lib().dec_batch_heat_steps(self._h, self._u(u), nsteps, int(bool(graph)), _stream(stream)))
        return u

    def steps_noalloc(self, u, nsteps, stream=None):
        """The same steps with the library's fixed scratch array as the
        partner buffer: no device allocation (states up to 32 MiB)."""
        check(lib().dec_batch_heat_steps_noalloc(self._h, self._u(u), nsteps, _stream(stream)))
        return u

    def instance(self, u, i):
        return u[self.offsets[i]:self.offsets[i + 1]]
