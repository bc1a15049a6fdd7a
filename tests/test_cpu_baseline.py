"""The CPU baseline (the oracle's operator loops built with -O3 and OpenMP,
oracle/cpu_baseline.c) computes what the oracle computes, bit for bit, at one
thread and at several: the benchmark times the same arithmetic it checks."""
import numpy as np
import pytest


@pytest.mark.parametrize("threads", [1, 4])
def test_cpu_baseline_equals_the_oracle(O, threads):
    import cpubase

    d = O.Dual(O.icosphere(3))
    s = cpubase.OpSet(d)
    V, E, E2, T = s.V, s.E, s.E2, s.T
    ref = {
        "wedge01": d.wedge01(V, E), "wedge11": d.wedge11(E, E2), "wedge02": d.wedge02(V, T),
        "d0": d.differential(0).matvec(V), "d1": d.differential(1).matvec(E),
        "hodge0": d.hodge_diag(0) * V, "hodge1": d.hodge_diag(1) * E, "hodge2": d.hodge_diag(2) * T,
        "laplacian0": d.laplacian0().matvec(V),
    }
    for name in cpubase.NAMES:
        assert np.array_equal(s.apply(name, threads), ref[name]), name
    assert s.run(threads, 3) > 0
