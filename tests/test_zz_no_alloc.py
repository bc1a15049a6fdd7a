"""dec_batch_heat_steps_noalloc (the entry point that allocates nothing)
against dec_batch_heat_steps: same kernels, so the states must be equal bit
for bit, on two calls in a row with the second batch smaller than the first."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(cs, levels, dtype, seed):
    import torch

    meshes = {k: cs.dualize(cs.icosphere(k)) for k in sorted(set(levels))}
    b = cs.HeatBatch([meshes[k] for k in levels], dtype)
    u0 = torch.empty(b.total, dtype=dtype, device="cuda")
    for i in range(len(levels)):
        cs.fill_uniform(b.instance(u0, i), seed + i)
    return b, u0


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_noalloc_matches_heat_steps(cs, dtype_name):
    import torch

    from cspaces._lib import check, lib

    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(11)
    big = [int(x) for x in rng.choice([4, 5, 6, 7], size=20)]
    small = [int(x) for x in rng.choice([2, 3, 4], size=5)]
    for levels, nsteps, seed in ((big, 7, 100), (small, 4, 200)):   # second call: smaller input
        b, u0 = _batch(cs, levels, dtype, seed)
        ref = u0.clone()
        check(lib().dec_batch_heat_steps(b._h, ctypes.c_void_p(ref.data_ptr()), nsteps, 0, None))
        u = u0.clone()
        check(lib().dec_batch_heat_steps_noalloc(b._h, ctypes.c_void_p(u.data_ptr()), nsteps, None))
        torch.cuda.synchronize()
        assert torch.equal(u, ref), f"levels {levels}, {nsteps} steps"
        assert not torch.equal(u, u0)


def test_noalloc_rejects_state_above_scratch(cs):
    import torch

    from cspaces._lib import DEC_ERR_ARG, lib

    m = cs.dualize(cs.icosphere(8))
    b = cs.HeatBatch([m] * 7, torch.float64)          # 7 x 655,362 rows x 8 bytes > 32 MiB
    assert b.total * 8 > 32 << 20
    u = torch.zeros(b.total, dtype=torch.float64, device="cuda")
    rc = lib().dec_batch_heat_steps_noalloc(b._h, ctypes.c_void_p(u.data_ptr()), 1, None)
    assert rc == DEC_ERR_ARG
    assert b"scratch" in lib().dec_last_error()
