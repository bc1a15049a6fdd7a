#!/usr/bin/env python3
"""Benchmark of the DEC hot path on one MI355X.

A step is one application of each operator of the reference operator set
(wedge 0-1, 1-1, 0-2; d0, d1; ⋆0, ⋆1, ⋆2; one 0-form Laplacian matvec) on the
icosphere with 8 binary subdivisions (655,362 V / 1,966,080 E / 1,310,720 T),
float64 forms, int32 incidence tables, inputs resident in HBM.  Each step uses
the next of several complete table + buffer sets so that the working set
cycles through far more memory than the 256 MB Infinity Cache holds: the
figures are HBM figures.  `value` = algorithmic bytes of the whole step /
measured time (GB/s).  The same JSON line carries the float32 operator set,
the Poisson solve (levels 3→8, plain float64 V(2,2) vs float32-in-float64
iterative refinement), the roofline of the dominant kernel and the CPU
oracle's throughput on a bounded sample of the same operators.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "combinatorialspaces.jl_amd"))

METRIC = "DEC op GB/s vs HBM peak; Poisson solve s to 1e-8 resid, learned+FP32 vs plain MG"
HBM_PEAK_GBPS = 8000.0
OPSET = ["wedge01", "wedge11", "wedge02", "d0", "d1", "hodge0", "hodge1", "hodge2", "laplacian0"]


def alg_bytes(name, nv, ne, nt, s, lap_nnz=None, ntet=0):
    """Compulsory bytes of one application: every table / input element read
    once, every output written once (int32 indices, s-byte values)."""
    return {
        "wedge03": ntet * (16 + 6 * s) + nv * s,
        "wedge21": ntet * (42 + 5 * s) + (nt + ne) * s,
        "d2": ntet * (20 + s) + nt * s,
        "hodge3": 3 * s * ntet,
        "wedge01": ne * (8 + 2 * s) + nv * s,
        "wedge11": nt * (12 + 4 * s) + 2 * ne * s,
        "wedge02": nt * (12 + 5 * s) + nv * s,
        "wedge_dd01": ne * (8 + 4 * s) + nt * s,
        "d0": ne * (9 + s) + nv * s,
        "d1": nt * (13 + s) + ne * s,
        "boundary1": 2 * ne * 5 + nv * s + ne * s,
        "boundary2": 3 * nt * 5 + ne * s + nt * s,
        "hodge0": 3 * s * nv,
        "hodge1": 3 * s * ne,
        "hodge2": 3 * s * nt,
        "laplacian0": (lap_nnz if lap_nnz is not None else nv + 2 * ne) * (4 + s) + 2 * nv * s,
    }[name]


def make_op(cs, mesh, name, td):
    if name == "wedge01":
        return cs.dec_wedge_product(0, 1, mesh, cast_float=td).operator
    if name == "wedge11":
        return cs.dec_wedge_product(1, 1, mesh, cast_float=td).operator
    if name == "wedge02":
        return cs.dec_wedge_product(0, 2, mesh, cast_float=td).operator
    if name == "wedge_dd01":
        return cs.dec_wedge_product_dd(0, 1, mesh, cast_float=td).operator
    if name == "d0":
        return cs.dec_differential(0, mesh, cast_float=td)
    if name == "d1":
        return cs.dec_differential(1, mesh, cast_float=td)
    if name == "wedge03":
        return cs.dec_wedge_product(0, 3, mesh, cast_float=td).operator
    if name == "wedge21":
        return cs.dec_wedge_product(2, 1, mesh, cast_float=td).operator
    if name == "d2":
        return cs.dec_differential(2, mesh, cast_float=td)
    if name == "boundary1":
        return cs.dec_boundary(1, mesh, cast_float=td)
    if name == "boundary2":
        return cs.dec_boundary(2, mesh, cast_float=td)
    if name.startswith("hodge"):
        return cs.dec_hodge_star(int(name[-1]), mesh, cs.DiagonalHodge, cast_float=td)
    if name == "laplacian0":
        return cs.fast_laplace_beltrami(mesh, cs.DiagonalHodge, cast_float=td)
    raise KeyError(name)


def time_opset(torch, cs, mesh, names, td, sets, steps, warmup, cache_resident=True, per_op=True):
    """Per-op mean launch duration (ms): `steps` back-to-back launches of the
    op over `sets` independent table/buffer sets, between two HIP events
    recorded on the stream the kernels run on.  Then the whole step (every op
    once per step, rotating sets) timed the same way, once as one launch per
    op and once as one grouped launch (dec_op_apply_many) per step."""
    stream = torch.cuda.current_stream()
    ops = {n: [make_op(cs, mesh, n, td) for _ in range(sets)] for n in names}
    bufs = {}
    seed = 100
    for n in names:
        bufs[n] = []
        for s in range(sets):
            op = ops[n][s]
            x = cs.fill_uniform(torch.empty(op.cols, dtype=td, device="cuda"), seed); seed += 1
            y = (cs.fill_uniform(torch.empty(op.cols2, dtype=td, device="cuda"), seed) if op.cols2 else None)
            seed += 1
            out = torch.empty(op.rows, dtype=td, device="cuda")
            bufs[n].append((x, y, out))

    def launch(n, s):
        x, y, out = bufs[n][s]
        ops[n][s].apply(x, y, out=out, stream=stream)

    for w in range(warmup):
        for n in names:
            launch(n, w % sets)
    torch.cuda.synchronize()
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    per = {}
    for n in (names if per_op else []):
        a, b = ev(), ev()
        a.record(stream)
        for k in range(steps):
            launch(n, k % sets)
        b.record(stream)
        torch.cuda.synchronize()
        per[n] = a.elapsed_time(b) / steps
    t0, t1 = ev(), ev()
    t0.record(stream)
    for k in range(steps):
        for n in names:
            launch(n, k % sets)
    t1.record(stream)
    torch.cuda.synchronize()
    total = t0.elapsed_time(t1)
    groups = [[(ops[n][k], *bufs[n][k]) for n in names] for k in range(sets)]
    for k in range(warmup):
        cs.apply_many(groups[k % sets], stream=stream)
    torch.cuda.synchronize()
    t0, t1 = ev(), ev()
    t0.record(stream)
    for k in range(steps):
        cs.apply_many(groups[k % sets], stream=stream)
    t1.record(stream)
    torch.cuda.synchronize()
    grouped = t0.elapsed_time(t1)
    lap_nnz = None
    if "laplacian0" in names:
        lap_nnz = len(ops["laplacian0"][0].csr()[1])
    cache = {}
    if cache_resident:   # one set, back-to-back: the Infinity-Cache-resident figure
        for n in names:
            for _ in range(3):
                launch(n, 0)
            a, b = ev(), ev()
            a.record(stream)
            for _ in range(20):
                launch(n, 0)
            b.record(stream)
            torch.cuda.synchronize()
            cache[n] = a.elapsed_time(b) / 20
    del ops, bufs, groups
    torch.cuda.empty_cache()
    return per, total, grouped, cache, lap_nnz


def musical_ops(torch, cs, mesh, reps=20):
    """The composite operators on the headline mesh, FP64: milliseconds per
    application, wall clock over `reps` back-to-back applications (the ⋆1⁻¹
    solve inside some of them reads its stop flag from the host).  ♭♯ is a
    plain sparse-matrix product and is also given in GB/s of its compulsory
    traffic (nnz int32 columns and values, input and output once)."""
    td = torch.float64
    ne = mesh.ne
    fs = cs.flat_sharp_mat(mesh, cast_float=td)
    nnz = len(fs.csr()[1])
    ops = {
        "flat_sharp": (fs, 1),
        "inv_hodge1_geometric": (cs.dec_inv_hodge_star(1, mesh, cs.GeometricHodge, cast_float=td), 1),
        "wedge_pd11": (cs.dec_wedge_product_pd(1, 1, mesh, cast_float=td).operator, 2),
        "interior_product_dd11": (cs.interior_product_dd(1, 1, mesh, cast_float=td).operator, 2),
        "interior_product_dd12": (cs.interior_product_dd(1, 2, mesh, cast_float=td).operator, 2),
        "lie_derivative_dd11": (cs.lie_derivative_dd(1, 1, mesh, cast_float=td).operator, 2),
        "dual_laplacian0": (cs.dual_laplacian(0, mesh, cast_float=td), 1),
        "dual_laplacian1": (cs.dual_laplacian(1, mesh, cast_float=td), 1),
    }
    res = {}
    for name, (op, nin) in ops.items():
        x = cs.fill_uniform(torch.empty(op.cols, dtype=td, device="cuda"), 501)
        y = cs.fill_uniform(torch.empty(op.cols2, dtype=td, device="cuda"), 502) if nin == 2 else None
        out = torch.empty(op.rows, dtype=td, device="cuda")
        op.apply(x, y, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            op.apply(x, y, out=out)
        torch.cuda.synchronize()
        res[name] = {"ms": (time.perf_counter() - t0) / reps * 1e3}
    res["flat_sharp"]["GBps"] = (nnz * 12 + 2 * ne * 8) / (res["flat_sharp"]["ms"] * 1e-3) / 1e9
    res["flat_sharp"]["nnz"] = nnz
    return res


OPSET3 = ["wedge01", "wedge11", "wedge02", "wedge03", "wedge21", "d0", "d1", "d2", "hodge0", "hodge1", "hodge2",
          "hodge3", "laplacian0"]


def ops_3d(torch, cs, n=64, sets=3, steps=30, warmup=3):
    """The 3-D operator set on an n x n x n Kuhn box (6 n³ tetrahedra) with
    its barycentric dual, float64: one application each of the wedges 0-1,
    1-1, 0-2, 0-3 and 2-1, d0, d1, d2, ⋆0..⋆3 and the 0-form Laplacian per
    step, timed as in the 2-D step (rotating sets, HIP events on the kernels'
    stream); GB/s of the compulsory bytes."""
    t = time.perf_counter()
    mesh = cs.kuhn_box(n, n, n)
    cs.dualize(mesh, cs.Barycenter)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t
    nv, ne, nt, ntet = mesh.nv, mesh.ne, mesh.nt, mesh.ntet
    per, total, grouped, _, lap_nnz = time_opset(torch, cs, mesh, OPSET3, torch.float64, sets, steps, warmup,
                                                 cache_resident=False)
    B = {k: alg_bytes(k, nv, ne, nt, 8, lap_nnz, ntet) for k in OPSET3}
    step = sum(B.values())
    res = {"mesh": f"Kuhn box {n}^3 cubes ({nv} V / {ne} E / {nt} T / {ntet} Tet), barycentric dual",
           "setup_s": setup, "step_MB": step / 1e6,
           "grouped_GBps": step / (grouped / steps * 1e-3) / 1e9,
           "one_launch_per_op_GBps": step / (total / steps * 1e-3) / 1e9,
           "ops": {k: {"ms": per[k], "GBps": B[k] / (per[k] * 1e-3) / 1e9} for k in OPSET3}}
    del mesh
    torch.cuda.empty_cache()
    return res


def stream_gbps(torch):
    """Hand-written STREAM copy (b = a) and triad (a = b + s c) kernels
    (libdec_stream.so), float64, at traffic sizes of the operators (16-120 MB
    per launch: read + written bytes) and at 2 GiB.  Below 2 GiB the arrays
    rotate over enough sets that one rotation exceeds 1 GiB, four times the
    256 MB Infinity Cache, as the operator legs do.  Each figure is the best
    of plain and nontemporal stores."""
    lib = ctypes.CDLL(os.path.join(ROOT, "combinatorialspaces.jl_amd", "libdec_stream.so"))
    lib.dec_stream_run.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_float)]
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for kind, name, arrays in ((0, "copy", 2), (1, "triad", 3)):
        res = {}
        for mb in (16, 32, 64, 120, 2048):
            nbytes = mb * (1 << 20) if mb == 2048 else mb * 10 ** 6     # traffic per launch
            n = (nbytes // (8 * arrays)) // 2 * 2
            sets = 1 if mb == 2048 else max(1, -(-(1 << 30) // (n * 8 * arrays)))
            bufs = [torch.ones(n, dtype=torch.float64, device="cuda") for _ in range(arrays * sets)]
            ptrs = (ctypes.c_void_p * len(bufs))(*[b.data_ptr() for b in bufs])
            best = None
            for nt in (0, 1):
                ms = ctypes.c_float()
                reps = 20 if mb == 2048 else 200
                rc = lib.dec_stream_run(kind, nt, n, sets, ptrs, 2 * sets, reps, stream, ctypes.byref(ms))
                if rc != 0:
                    raise RuntimeError(f"dec_stream_run: hip error {rc}")
                gbps = arrays * n * 8 / (ms.value * 1e-3) / 1e9
                best = gbps if best is None else max(best, gbps)
            res[f"{mb}MB" if mb != 2048 else "2GiB"] = best
            del bufs
            torch.cuda.empty_cache()
        out[name] = res
    return out


def sh_rhs(points, h0, seed, nterms=16, lmax=12):
    """f = Σ_j c_j Y_{l_j m_j} (real orthonormal spherical harmonics, l ≤ 12),
    (l, m, c) drawn from a seeded stream, projected to ⋆0-weighted mean zero."""
    import numpy as np
    from scipy.special import sph_harm_y

    rng = np.random.default_rng(seed)
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    f = np.zeros(len(points))
    for _ in range(nterms):
        l = int(rng.integers(1, lmax + 1))
        m = int(rng.integers(-l, l + 1))
        c = rng.uniform(-1, 1)
        Y = sph_harm_y(l, abs(m), theta, phi)
        if m > 0:
            Yr = math.sqrt(2) * (-1) ** m * Y.real
        elif m < 0:
            Yr = math.sqrt(2) * (-1) ** m * Y.imag
        else:
            Yr = Y.real
        f += c * Yr
    return f - (h0 @ f) / h0.sum()


def poisson(torch, cs, base_k=3, levels=5):
    import numpy as np

    t = time.perf_counter()
    md = cs.MultigridData(cs.icosphere(base_k), cs.BinarySubdivision, levels, cs.DiagonalHodge, 2, project=True)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t
    fine = md.finest_mesh()
    pts = fine.export()["points"]
    h0 = cs.dec_hodge_star(0, fine, cs.DiagonalHodge) @ torch.ones(fine.nv, dtype=torch.float64, device="cuda")
    f = sh_rhs(pts, h0.cpu().numpy(), seed=2024)
    fd = torch.tensor(f, device="cuda")
    out = {"setup_s": setup, "levels": f"{base_k}->{base_k + levels}", "n_fine": fine.nv}
    cyc_bytes = {"plain": vcycle_bytes(md, 8), "mixed": vcycle_bytes(md, 4)}
    out["bytes_per_cycle_f64"] = cyc_bytes["plain"]
    out["bytes_per_cycle_f32"] = cyc_bytes["mixed"]
    for alg, tag in ((cs.block_gauss_seidel, ""), (cs.gauss_seidel, "_multicolour_gs")):
        for variant in ("plain", "mixed"):
            cs.solve(md, fd, tol=1e-8, variant=variant, alg=alg)   # warm-up (allocations, graph capture)
            torch.cuda.synchronize()
            dt = None
            for _ in range(3):
                t = time.perf_counter()
                u, cyc, rr = cs.solve(md, fd, tol=1e-8, variant=variant, alg=alg)
                torch.cuda.synchronize()
                e = time.perf_counter() - t
                dt = e if dt is None else min(dt, e)
            key = ("plain_f64" if variant == "plain" else "mixed_f32_ir") + tag
            gbps = cyc * cyc_bytes[variant] / dt / 1e9
            out[key] = {"seconds": dt, "cycles": cyc, "relres_f64": rr, "smoother": alg.name,
                        "cycle_GBps": gbps, "frac_hbm_peak": gbps / 8000.0}
    out["speedup_mixed_vs_plain"] = out["plain_f64"]["seconds"] / out["mixed_f32_ir"]["seconds"]
    # conjugate gradients preconditioned by one V-cycle (float64 or float32),
    # the hybrid block Gauss-Seidel smoother
    extra = pcg_extra_bytes(md)
    for variant, key, cyc in (("pcg", "pcg_f64", "plain"), ("pcg_mixed", "pcg_mixed", "mixed")):
        cs.solve(md, fd, tol=1e-8, variant=variant)
        torch.cuda.synchronize()
        dt = None
        for _ in range(3):
            t = time.perf_counter()
            u, it, rr = cs.solve(md, fd, tol=1e-8, variant=variant)
            torch.cuda.synchronize()
            e = time.perf_counter() - t
            dt = e if dt is None else min(dt, e)
        per_it = cyc_bytes[cyc] + extra[cyc]
        gbps = it * per_it / dt / 1e9
        out[key] = {"seconds": dt, "iterations": it, "relres_f64": rr, "smoother": "block_gauss_seidel",
                    "bytes_per_iteration": per_it, "iteration_GBps": gbps, "frac_hbm_peak": gbps / 8000.0}
    return out, md, fd


def pcg_extra_bytes(md):
    """Algorithmic bytes a CG iteration adds to its V-cycle on the finest level
    (float64 vectors, int32 indices): zeroing z; the dot r·z (r, z); the
    direction launch (the operator's entries and row index, z and p_old, p and
    A p written); the update (x, r, p, A p read, x, r, p_old written); and for
    the float32 preconditioner the conversions of r and z (8 + 4 bytes each)."""
    n = md.sizes[0]
    nnz = int(md.export(0, "A")[0][-1])
    base = n * 8 + 2 * n * 8 + (nnz * 12 + n * 4 + 2 * n * 8 + 2 * n * 8) + 7 * n * 8
    return {"plain": base, "mixed": base + 2 * n * 12 + n * 4}


def vcycle_bytes(md, s):
    """Algorithmic bytes of one V(2,2) cycle on levels above the dense coarse
    solve, s = 8 (float64) or 4 (float32), int32 indices: per level four
    sweeps (matrix entries, b, the iterate in and out, the row index), the
    residual (matrix, b, u, r, row index), the restriction (R entries, r, b_c,
    the zeroed u_c) and the prolongation (P entries, z_c, u in and out); the
    coarsest level's dense solve (n² entries, r, u)."""
    total = 0
    nl = md.nlevels
    for lv in range(nl):
        n = md.sizes[lv]
        rp, _, _ = md.export(lv, "A")
        nnz = int(rp[-1])
        if lv == nl - 1:
            total += n * n * s + 2 * n * s
            continue
        nc = md.sizes[lv + 1]
        rnnz = int(md.export(lv, "R")[0][-1])
        pnnz = int(md.export(lv, "P")[0][-1])
        sweep = nnz * (4 + s) + n * (3 * s + 4)
        total += 4 * sweep
        total += nnz * (4 + s) + n * (3 * s + 4)
        total += rnnz * (4 + s) + n * s + 2 * nc * s
        total += pnnz * (4 + s) + nc * s + 2 * n * s
    return total


def mg_reference_calls(torch, cs, fd, base_k=3, levels=5, reps=3):
    """The reference's own multigrid calls on the level 3->8 hierarchy, with its
    default Krylov CG smoother (3 steps): multigrid_vcycles / wcycles (5
    cycles, Multigrid.jl:814-817), full_multigrid (5 cycles, μ = 1 and the μ = 2
    of dec_Δ⁻¹, Multigrid.jl:830-837, FastDEC.jl:866-869), and V-cycles on a
    Galerkin hierarchy (Aₖ₊₁ = Rₖ Aₖ Pₖ).  Wall-clock ms per call (best of
    `reps`, after a warm-up call) and the FP64 relative residual reached.  The
    hierarchies are built here (DirectMode and GalerkinMode, 3 smoothing
    steps), so no other leg's smoother settings carry over; fd is the
    right-hand side of the Poisson leg on the same finest mesh."""
    md = cs.MultigridData(cs.icosphere(base_k), cs.BinarySubdivision, levels, cs.DiagonalHodge, 3, project=True)
    fine = md.finest_mesh()
    L = cs.fast_laplace_beltrami(fine, cs.DiagonalHodge)
    bn = torch.linalg.norm(fd).item()

    def relres(u):
        return torch.linalg.norm(fd - L @ u).item() / bn

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best, u = None, None
        for _ in range(reps):
            t = time.perf_counter()
            u = fn()
            torch.cuda.synchronize()
            e = time.perf_counter() - t
            best = e if best is None else min(best, e)
        return {"ms": best * 1e3, "relres_f64": relres(u)}

    zero = lambda: torch.zeros_like(fd)  # noqa: E731
    out = {
        "vcycles_cg_5": timed(lambda: cs.multigrid_vcycles(zero(), fd, md, 5, cs.cg, steps=3)),
        "wcycles_cg_5": timed(lambda: cs.multigrid_wcycles(zero(), fd, md, 5, cs.cg, steps=3)),
        "full_multigrid_cg_5": timed(lambda: cs.full_multigrid(fd, md, 5, cs.cg, 1, steps=3)),
        "dec_Delta_inv_fmg_mu2_5": timed(lambda: cs.full_multigrid(fd, md, 5, cs.cg, 2, steps=3)),
    }
    mg = cs.MultigridData(cs.icosphere(base_k), cs.BinarySubdivision, levels, cs.DiagonalHodge, 3, project=True,
                          mode=cs.GalerkinMode)
    out["galerkin_vcycles_cg_5"] = timed(lambda: cs.multigrid_vcycles(zero(), fd, mg, 5, cs.cg, steps=3))
    del mg, md
    return out


def interpolation_ops(torch, cs, mesh, reps=20):
    """The interpolation matrices on the headline mesh, FP64: µs per
    application and GB/s of their compulsory bytes (nnz int32 columns and
    float64 values, input and output once)."""
    td = torch.float64
    res = {}
    for name, op in (("p2_d2", cs.p2_d2_interpolation(mesh, cast_float=td)),
                     ("d0_p0", cs.d0_p0_interpolation(mesh, cast_float=td)),
                     ("p0_d0", cs.p0_d0_interpolation(mesh, cast_float=td))):
        nnz = len(op.csr()[1])
        x = cs.fill_uniform(torch.empty(op.cols, dtype=td, device="cuda"), 601)
        out = torch.empty(op.rows, dtype=td, device="cuda")
        op.apply(x, out=out)
        stream = torch.cuda.current_stream()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            op.apply(x, out=out, stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        byt = nnz * 12 + (op.rows + op.cols) * 8
        res[name] = {"us": ms * 1e3, "GBps": byt / (ms * 1e-3) / 1e9, "nnz": nnz}
    return res


def learned(torch, cs, md, fd, npairs=10000, train_levels=3, steps=2000, batch=16):
    """The learned smoother in the level 3->8 mixed solve.  Training pairs come
    from plain V-cycles on seeded spherical-harmonic right-hand sides on the
    3->(3+train_levels) hierarchy; the trained parameters (level independent)
    are loaded into the network bound to md."""
    t = time.perf_counter()
    mdt = cs.MultigridData(cs.icosphere(3), cs.BinarySubdivision, train_levels, cs.DiagonalHodge, 2, project=True)
    R, E = cs.make_pairs(mdt, npairs, cycles_per_rhs=6, seed=11)
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t
    nn_t = cs.LearnedSmoother(mdt, seed=5)
    val, evl = R[-512:].contiguous(), E[-512:].contiguous()
    l0, _ = nn_t.loss_grad(0, val, evl)
    r0, _ = nn_t.loss_grad(0, val)
    t = time.perf_counter()
    # error loss: the high-frequency part of the error left after smoothing
    hist = nn_t.train(0, R[:-512], steps=steps, E=E[:-512], batch=batch, lr=2e-3, seed=6)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    l1, _ = nn_t.loss_grad(0, val, evl)
    r1, _ = nn_t.loss_grad(0, val)
    params = nn_t.get_params()
    del R, E, val, evl
    nn = cs.LearnedSmoother(md, seed=5)
    nn.set_params(params)
    out = {"pairs": npairs, "train_mesh_vertices": mdt.sizes[0], "pairs_s": t_pairs, "train_steps": steps,
           "train_batch": batch, "train_s": t_train, "loss": "error (high-pass)",
           "val_error_loss_untrained": l0, "val_error_loss_trained": l1,
           "val_residual_loss_untrained": r0, "val_residual_loss_trained": r1,
           "loss_first": float(hist[0]), "loss_last": float(hist[-1])}
    # one smoothing application at the finest level with the inference kernels
    # (BF16 MFMA dense layers): time per application, and the dense layers'
    # rate (2 layers x 2 x 64 x 32 flops per row)
    n0 = md.sizes[0]
    r0 = torch.randn(n0, dtype=torch.float32, device="cuda")
    u0 = torch.zeros(n0, dtype=torch.float32, device="cuda")
    for _ in range(3):
        nn.smooth(0, r0, u0)
    stream = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    reps = 20
    for _ in range(reps):
        nn.smooth(0, r0, u0)
    b.record(stream)
    torch.cuda.synchronize()
    app_ms = a.elapsed_time(b) / reps
    out["inference_ms_per_application_fine"] = app_ms
    out["dense_layer_TFLOPs_fine"] = 2 * 2 * 64 * 32 * n0 / (app_ms * 1e-3) / 1e12
    # algorithmic bytes of the application, per row with w slots (icosphere
    # vertices have 5 or 6 neighbours, w = 6): the scale's read of r (4);
    # layer 1: a one-byte index and a weight per slot, r, x and h1 (32 BF16)
    # written; layers 2 and 3: column and weight per slot and the own h row,
    # h2 written; x, D and u read, u written; perm read in layers 1 and 3 (the
    # vertex-order call)
    w = 6
    per_row = 4 + (5 * w + 4 + 4 + 64) + (8 * w + 64 + 64) + (8 * w + 64 + 4 + 4 + 8) + 8
    out["inference_alg_MB_fine"] = per_row * n0 / 1e6
    out["inference_GBps_fine"] = per_row * n0 / (app_ms * 1e-3) / 1e9
    out["inference_frac_hbm_peak_fine"] = out["inference_GBps_fine"] / 8000.0
    # levels that use the network: all (one or two applications); the two
    # finest; the three below the finest (the finest keeps Gauss-Seidel)
    # one application per level at each level size (internal numbering, as in the cycles)
    out["inference_ms_per_application"] = {}
    for lv, n in enumerate(md.sizes[:-1]):
        rl = torch.randn(n, dtype=torch.float32, device="cuda")
        ul = torch.zeros(n, dtype=torch.float32, device="cuda")
        for _ in range(3):
            nn.smooth(lv, rl, ul)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(reps):
            nn.smooth(lv, rl, ul)
        b.record(stream)
        torch.cuda.synchronize()
        out["inference_ms_per_application"][str(n)] = a.elapsed_time(b) / reps
    for tag, lo, hi, apps in (("all_levels", 0, 0, 1), ("all_levels_2_applications", 0, 0, 2),
                              ("levels_ge_40963", 40963, 0, 1), ("levels_10242_to_163842", 10242, 163842, 1),
                              ("levels_10242_to_40962", 10242, 40962, 1), ("level_10242", 10242, 10242, 1),
                              ("level_40962", 40962, 40962, 1)):
        nn.attach(True, min_rows=lo, max_rows=hi, applications=apps)
        cs.solve(md, fd, tol=1e-8, variant="mixed")
        torch.cuda.synchronize()
        best = None
        for _ in range(3):
            t = time.perf_counter()
            u, cyc, rr = cs.solve(md, fd, tol=1e-8, variant="mixed")
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        st = cs.learned_stats(md)
        out[f"mixed_learned_{tag}"] = {"seconds": best, "cycles": cyc, "relres_f64": rr, "fallbacks": st["fallbacks"],
                                       "learned_cycles": st["learned_cycles"]}
    nn.attach(False)
    return out


def _host_cpus():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return avail, model


def cpu_baseline(k=8, budget_s=8.0):
    """The oracle's operator loops (C, -O3, OpenMP; oracle/cpu_baseline.c) on
    the operator set of the GPU step, icosphere k, float64: passes for about
    budget_s seconds at one thread and again at all host cores the process may
    use (at most 16, the box's CPU share of one GPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import cpubase

    t = time.perf_counter()
    d = O.Dual(O.icosphere(k))
    s = cpubase.OpSet(d)
    setup = time.perf_counter() - t
    nbytes = sum(alg_bytes(n, d.nv, d.ne, d.nt, 8, s.L.nnz) for n in OPSET)
    avail, model = _host_cpus()
    threads = max(1, min(avail, 16))
    rates = {}
    for th in sorted({1, threads}):
        dt1 = s.run(th, 1)
        passes = max(2, int(budget_s / max(dt1, 1e-4)))
        dt = s.run(th, passes)
        rates[th] = (passes * nbytes / dt / 1e9, passes, dt)
    v, passes, dt = rates[threads]
    return {"value": v, "unit": "GB/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpus_available": avail, "cpu_model": model,
            "one_thread_GBps": rates[1][0],
            "sample": f"the oracle's operator loops (C, -O3, OpenMP) on the same nine operators, float64, "
                      f"icosphere k={k} ({d.nv} V): {passes} passes in {dt:.1f} s at {threads} threads, "
                      f"{rates[1][1]} passes in {rates[1][2]:.1f} s at 1 thread; set-up {setup:.1f} s"}


def heat_instances(torch, cs, ninst=64, seed=7):
    """The batch of the batched leg: a seeded uneven mix of icosphere levels
    4-7, each instance on its own mesh (the level's points moved tangentially
    by a seeded jitter of up to 5% of the mean edge length and put back on the
    unit sphere), so no two instances share a Laplacian; seeded uniform states."""
    import numpy as np

    rng = np.random.default_rng(seed)
    levels = [int(x) for x in rng.choice([4, 5, 6, 7], size=ninst, p=[0.4, 0.3, 0.2, 0.1])]
    base = {k: cs.icosphere(k).export() for k in sorted(set(levels))}
    inst = []
    for i, k in enumerate(levels):
        e = base[k]
        p = e["points"].copy()
        h = np.linalg.norm(p[e["v0"]] - p[e["v1"]], axis=1).mean()
        p += np.random.default_rng(1000 + i).uniform(-0.05, 0.05, p.shape) * h
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        m = cs.mesh_from_arrays(p, e["v0"], e["v1"], e["e0"], e["e1"], e["e2"], e["edge_orientation"],
                                e["tri_orientation"])
        inst.append(cs.dualize(m))
    b32 = cs.HeatBatch(inst, torch.float32)
    b64 = cs.HeatBatch(inst, torch.float64)
    u64 = torch.empty(b64.total, dtype=torch.float64, device="cuda")
    for i in range(ninst):
        cs.fill_uniform(b64.instance(u64, i), 500 + i)
    return levels, inst, b32, b64, u64


def heat_step_bytes(levels, s):
    """Algorithmic bytes of one batched step: each instance's Laplacian rows
    (nnz = nv + 2 ne entries of int32 column + value) and its state read and
    written once."""
    total = 0
    for k in levels:
        nv, ne = 10 * 4 ** k + 2, 30 * 4 ** k
        total += (nv + 2 * ne) * (4 + s) + 2 * nv * s
    return total


def batched_heat(torch, cs, ninst=64, nsteps=100, seed=7, counters=True):
    """64 explicit heat instances u ← u - dt L u on a seeded uneven mix of
    icosphere levels 4-7 (one mesh per instance), FP32, in one persistent
    launch per step replayed from a hipGraph, launch shape from the autotuner;
    checked against the same steps in FP64.  The working set (tables and
    state, about 104 MB) fits the 256 MB Infinity Cache, so across steps the
    tables come back from that cache: the rate is of L2-miss traffic, and the
    counter pass gives the bytes the L2 actually fetched and wrote."""
    levels, inst, b32, b64, u64 = heat_instances(torch, cs, ninst, seed)
    u32 = u64.float()
    t = time.perf_counter()
    shape = b32.tune(u32.clone())
    tune_s = time.perf_counter() - t
    u = u32.clone()
    b32.steps(u, nsteps, graph=True)       # captures the graph for (u, nsteps)
    u.copy_(u32)
    stream = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    b32.steps(u, nsteps, graph=True)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    v = u64.clone()
    b64.steps(v, nsteps, graph=True)
    err = max(float((b32.instance(u, i).double() - b64.instance(v, i)).norm() / b64.instance(v, i).norm())
              for i in range(ninst))
    step_bytes = heat_step_bytes(levels, 4)
    out = {"instances": ninst, "meshes": "one jittered icosphere per instance (no shared tables)",
           "levels": {str(k): levels.count(k) for k in sorted(set(levels))},
           "total_vertices": int(b32.total), "steps": nsteps, "ms_per_step": ms / nsteps,
           "instance_steps_per_s": ninst * nsteps / (ms * 1e-3), "step_bytes": step_bytes,
           "GBps": step_bytes / (ms / nsteps * 1e-3) / 1e9, "working_set_MB": step_bytes / 1e6,
           "launch_shape": shape, "tune_s": tune_s, "max_rel_err_f32_vs_f64": err}
    if counters:
        try:
            t = traffic("heat", "k_heat_step", 0, 1)
            out["counter_bytes_per_step"] = t["bytes"]
            out["counter_GBps"] = t["bytes"] / (ms / nsteps * 1e-3) / 1e9
            out["counter_over_algorithmic"] = t["bytes"] / step_bytes
        except RuntimeError as e:
            out["counter_error"] = str(e)[:200]
    return out


COPY_MIB = 256  # calibration kernel (elementwise negation): 256 MiB read + 256 MiB written


def pmc_probe(op_name, k, sets, launches):
    """The program that the counter passes profile: one plain copy of known
    size (calibration) and `launches` back-to-back launches of one operator
    over `sets` rotating table/buffer sets, as in the timed run."""
    import torch
    import cspaces as cs

    torch.cuda.init()
    x = torch.ones(COPY_MIB * 2 ** 20 // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)     # an elementwise kernel: reads and writes COPY_MIB MiB
    torch.cuda.synchronize()
    del x, y
    if op_name == "heat":
        levels, inst, b32, b64, u64 = heat_instances(torch, cs)
        u = u64.float()
        b32.tune(u.clone())
        b32.steps(u, launches, graph=False)
        torch.cuda.synchronize()
        return
    mesh = cs.icosphere(k)
    cs.dualize(mesh)
    if op_name == "group":
        time_opset(torch, cs, mesh, OPSET, torch.float64, sets, launches, 2, False, per_op=False)
    else:
        time_opset(torch, cs, mesh, [op_name], torch.float64, sets, launches, 2, False)
    torch.cuda.synchronize()


def pmc_counter(counter, op_name, kernel, k, sets, launches, timeout_s=240):
    """Mean per-dispatch value (KiB) of one rocprofv3 counter for `kernel` and
    for the calibration copy, from a child `rocprofv3 --pmc <counter>` run of
    pmc_probe.  Returns (kernel KiB, copy KiB) or raises RuntimeError."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="dec_pmc_", dir="/tmp")
    cmd = ["timeout", "-s", "KILL", str(timeout_s), exe, "--pmc", counter, "--output-format", "csv", "-d", out,
           "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--pmc-probe", op_name, "--k", str(k),
           "--sets", str(sets), "--steps", str(launches)]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {counter} exited {r.returncode}: {r.stdout[-400:]}")
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file")
    kv, cv = [], []
    for r in csv.DictReader(open(files[0])):
        if r.get("Counter_Name") != counter:
            continue
        name, v = r["Kernel_Name"], float(r["Counter_Value"])
        if kernel + "<" in name or kernel + "(" in name:
            kv.append(v)
        elif "elementwise_kernel" in name and "neg" in name.lower():
            cv.append(v)
    shutil.rmtree(out, ignore_errors=True)
    if not kv or not cv:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: kernel rows {len(kv)}, copy rows {len(cv)}")
    return sum(kv) / len(kv), sum(cv) / len(cv)


def traffic(op_name, kernel, k, sets):
    """HBM bytes per launch of the dominant kernel from two counter passes
    (FETCH_SIZE, then WRITE_SIZE), each scaled by what the counter reads for
    the calibration copy (COPY_MIB MiB each way)."""
    fk, fc = pmc_counter("FETCH_SIZE", op_name, kernel, k, sets, 24)
    wk, wc = pmc_counter("WRITE_SIZE", op_name, kernel, k, sets, 24)
    true_kib = COPY_MIB * 1024.0
    f_fetch, f_write = fc / true_kib, wc / true_kib
    return {"bytes": fk * 1024 / f_fetch + wk * 1024 / f_write, "fetch_KiB": fk, "write_KiB": wk,
            "fetch_factor": f_fetch, "write_factor": f_write, "copy_MiB": COPY_MIB}


def traffic_per_op(k, sets):
    """HBM traffic of each operator of the step launched on its own, against
    its algorithmic bytes; one JSON line per operator."""
    nv, ne, nt = 10 * 4 ** k + 2, 30 * 4 ** k, 20 * 4 ** k
    kernels = {"wedge01": "k_wedge01", "wedge11": "k_wedge11", "wedge02": "k_wedge02", "d0": "k_d0", "d1": "k_d1",
               "hodge0": "k_diag", "hodge1": "k_diag", "hodge2": "k_diag", "laplacian0": "k_ell"}
    for name in OPSET:
        t = traffic(name, kernels[name], k, sets)
        alg = alg_bytes(name, nv, ne, nt, 8, nv + 2 * ne)
        print(json.dumps({"op": name, "traffic_MB": t["bytes"] / 1e6, "algorithmic_MB": alg / 1e6,
                          "ratio": t["bytes"] / alg,
                          "fetch_MB": t["fetch_KiB"] * 1024 / t["fetch_factor"] / 1e6,
                          "write_MB": t["write_KiB"] * 1024 / t["write_factor"] / 1e6}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--sets", type=int, default=6)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f32", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the ops outside the headline set")
    ap.add_argument("--no-cache-resident", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--learned-pairs", type=int, default=10000)
    ap.add_argument("--learned-steps", type=int, default=2000)
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--pmc-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--composite-only", action="store_true",
                    help="time only the composite operators (ms per application) and print their JSON")
    ap.add_argument("--no-3d", action="store_true", help="skip the operator set on a tetrahedral mesh")
    ap.add_argument("--ops3d-only", action="store_true",
                    help="time only the operator set on a tetrahedral mesh and print its JSON")
    ap.add_argument("--solve-only", action="store_true",
                    help="time only the Poisson solves on levels 3->8 and print their JSON")
    ap.add_argument("--reference-calls-only", action="store_true",
                    help="time only the reference's own multigrid calls (CG smoother) and print their JSON")
    ap.add_argument("--n3d", type=int, default=64, help="cubes per side of the Kuhn box of the 3-D leg")
    ap.add_argument("--traffic-per-op", action="store_true",
                    help="measure the HBM traffic of each operator of the step (two counter passes each)")
    args = ap.parse_args()
    if args.pmc_probe:
        pmc_probe(args.pmc_probe, args.k, args.sets, args.steps)
        return
    if args.traffic_per_op:
        traffic_per_op(args.k, args.sets)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; run it with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import cspaces as cs

    torch.cuda.init()
    if args.ops3d_only:
        print(json.dumps({"ops_3d_f64": ops_3d(torch, cs, args.n3d, steps=args.steps, warmup=args.warmup)}),
              flush=True)
        return
    if args.solve_only:
        print(json.dumps({"poisson": poisson(torch, cs)[0]}), flush=True)
        return
    if args.reference_calls_only:
        _, md, fd = poisson(torch, cs)
        del md
        print(json.dumps({"multigrid_reference_calls": mg_reference_calls(torch, cs, fd)}), flush=True)
        return
    t = time.perf_counter()
    mesh = cs.icosphere(args.k)
    cs.dualize(mesh)
    torch.cuda.synchronize()
    setup_mesh = time.perf_counter() - t
    nv, ne, nt = mesh.nv, mesh.ne, mesh.nt
    print(f"# icosphere k={args.k}: {nv} V {ne} E {nt} T, mesh+dual setup {setup_mesh:.2f} s", file=sys.stderr,
          flush=True)
    if args.composite_only:
        print(json.dumps({"composite_ops_f64": musical_ops(torch, cs, mesh),
                          "interpolation_ops_f64": interpolation_ops(torch, cs, mesh), "mesh_vertices": nv}),
              flush=True)
        return

    per, total, grouped, cache, lap_nnz = time_opset(torch, cs, mesh, OPSET, torch.float64, args.sets, args.steps,
                                                     args.warmup, not args.no_cache_resident)
    B = {n: alg_bytes(n, nv, ne, nt, 8, lap_nnz) for n in OPSET}
    step_bytes = sum(B.values())
    # the step as one grouped launch (dec_op_apply_many) is the job's figure;
    # one launch per operator is reported beside it
    ms_per_step = grouped / args.steps
    value = step_bytes / (ms_per_step * 1e-3) / 1e9
    ms_seq = total / args.steps
    ops = {n: {"ms": per[n], "MB": B[n] / 1e6, "GBps": B[n] / (per[n] * 1e-3) / 1e9,
               "frac_hbm_peak": B[n] / (per[n] * 1e-3) / 1e9 / HBM_PEAK_GBPS,
               "cache_resident_GBps": (B[n] / (cache[n] * 1e-3) / 1e9) if n in cache else None} for n in OPSET}
    print(f"# f64 op set: {value:.0f} GB/s grouped launch, {step_bytes / (ms_seq * 1e-3) / 1e9:.0f} GB/s one "
          "launch per op", file=sys.stderr, flush=True)
    # dominant kernel: the grouped kernel is the only one of the timed step
    roofline = {"bound": "hbm", "kernel": "k_apply_group", "op": "operator set (grouped)", "achieved": value,
                "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": value / HBM_PEAK_GBPS, "traffic": None,
                "alg_bytes_per_launch": step_bytes, "avg_launch_ms": ms_per_step}
    if not args.no_traffic:
        try:
            t = traffic("group", "k_apply_group", args.k, args.sets)
            roofline["traffic"] = t["bytes"]
            roofline["traffic_detail"] = t
            print(f"# traffic k_apply_group: {t['bytes'] / 1e6:.2f} MB per launch (algorithmic "
                  f"{step_bytes / 1e6:.2f} MB)", file=sys.stderr, flush=True)
        except (RuntimeError, OSError) as e:
            roofline["traffic_error"] = str(e)
            print(f"# traffic: {e}", file=sys.stderr, flush=True)
    single = {"ms_per_step": ms_seq, "GBps": step_bytes / (ms_seq * 1e-3) / 1e9}

    extra = {}
    extra_names = ["wedge_dd01", "boundary1", "boundary2"]
    if not args.no_extra:
        per_x, _, _, _, _ = time_opset(torch, cs, mesh, extra_names, torch.float64, args.sets,
                                       max(10, args.steps // 2), 2, False)
        extra["other_ops_f64"] = {n: {"ms": per_x[n], "GBps": alg_bytes(n, nv, ne, nt, 8) / (per_x[n] * 1e-3) / 1e9}
                                  for n in extra_names}
        extra["composite_ops_f64"] = musical_ops(torch, cs, mesh)
        print(f"# composite ops: {json.dumps(extra['composite_ops_f64'])}", file=sys.stderr, flush=True)
        extra["interpolation_ops_f64"] = interpolation_ops(torch, cs, mesh)
        print(f"# interpolation ops: {json.dumps(extra['interpolation_ops_f64'])}", file=sys.stderr, flush=True)
    if not args.no_f32:
        per32, total32, grouped32, cache32, lap32 = time_opset(torch, cs, mesh, OPSET, torch.float32, args.sets,
                                                               args.steps, args.warmup, False)
        B32 = {n: alg_bytes(n, nv, ne, nt, 4, lap32) for n in OPSET}
        extra["f32"] = {"value_GBps": sum(B32.values()) / (grouped32 / args.steps * 1e-3) / 1e9,
                        "one_launch_per_op_GBps": sum(B32.values()) / (total32 / args.steps * 1e-3) / 1e9,
                        "ops": {n: {"ms": per32[n], "GBps": B32[n] / (per32[n] * 1e-3) / 1e9} for n in OPSET}}
        print(f"# f32 op set: {extra['f32']['value_GBps']:.0f} GB/s aggregate", file=sys.stderr, flush=True)
    if not args.no_3d:
        extra["ops_3d_f64"] = ops_3d(torch, cs, args.n3d)
        print(f"# 3-D op set: {extra['ops_3d_f64']['grouped_GBps']:.0f} GB/s grouped launch", file=sys.stderr,
              flush=True)
    try:
        extra["stream_GBps"] = stream_gbps(torch)
        print(f"# stream: {json.dumps(extra['stream_GBps'])}", file=sys.stderr, flush=True)
    except (RuntimeError, OSError) as e:
        extra["stream_GBps"] = None
        print(f"# stream skipped: {e}", file=sys.stderr)
    if not args.no_solve:
        extra["poisson"], md, fd = poisson(torch, cs)
        print(f"# poisson: {json.dumps(extra['poisson'])}", file=sys.stderr, flush=True)
        if not args.no_learned:
            extra["learned"] = learned(torch, cs, md, fd, npairs=args.learned_pairs, steps=args.learned_steps)
            print(f"# learned: {json.dumps(extra['learned'])}", file=sys.stderr, flush=True)
        extra["multigrid_reference_calls"] = mg_reference_calls(torch, cs, fd)
        print(f"# multigrid reference calls: {json.dumps(extra['multigrid_reference_calls'])}", file=sys.stderr,
              flush=True)
        del md, fd
    if not args.no_batch:
        extra["batched_heat"] = batched_heat(torch, cs, counters=not args.no_traffic)
        print(f"# batched heat: {json.dumps(extra['batched_heat'])}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()

    line = {
        "metric": METRIC, "value": value, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"icosphere k={args.k} ({nv} V / {ne} E / {nt} T), circumcentric dual, "
                               "one application each of wedge 0-1, 1-1, 0-2, d0, d1, ⋆0, ⋆1, ⋆2 and the "
                               "0-form Laplacian per step in one grouped launch, float64 forms, int32 tables, "
                               f"{args.sets} rotating table+buffer sets (HBM, not cache)",
                   "mesh_vertices": nv, "step_bytes": step_bytes, "rotation_sets": args.sets},
        "roofline": roofline, "cpu_baseline": cpu, "one_launch_per_op": single, "ops": ops,
        "setup_mesh_s": setup_mesh, **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
