"""Batched heat steps: one persistent launch per step over many meshes of
uneven size, against the single-mesh operators and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_setup(cs):
    import torch

    rng = np.random.default_rng(5)
    levels = [int(x) for x in rng.choice([2, 3, 4, 5], size=12, p=[0.4, 0.3, 0.2, 0.1])]
    meshes = {k: cs.dualize(cs.icosphere(k)) for k in sorted(set(levels))}
    inst = [meshes[k] for k in levels]
    b64 = cs.HeatBatch(inst, torch.float64)
    b32 = cs.HeatBatch(inst, torch.float32)
    u0 = torch.empty(b64.total, dtype=torch.float64, device="cuda")
    for i in range(len(inst)):
        cs.fill_uniform(b64.instance(u0, i), 300 + i)
    return levels, meshes, inst, b64, b32, u0


def eager_steps(cs, mesh, dt, u, nsteps):
    """u ← u - dt L u with the single-mesh Laplacian operator."""
    L = cs.fast_laplace_beltrami(mesh, cs.DiagonalHodge, cast_float=u.dtype)
    for _ in range(nsteps):
        u = u - dt * (L @ u)
    return u


def test_layout(batch_setup):
    levels, meshes, inst, b64, b32, u0 = batch_setup
    sizes = np.diff(b64.offsets)
    assert list(sizes) == [10 * 4 ** k + 2 for k in levels]
    assert b64.total == sizes.sum()
    for i, k in enumerate(levels):
        e = meshes[k].export_dual()
        h = e["length"].mean()
        assert b64.dt[i] == pytest.approx(0.1 * h * h, rel=1e-12)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("nsteps", [1, 4])
def test_matches_single_mesh_operator(cs, batch_setup, graph, nsteps):
    levels, meshes, inst, b64, b32, u0 = batch_setup
    u = u0.clone()
    b64.steps(u, nsteps, graph=graph)
    for i, m in enumerate(inst):
        ref = eager_steps(cs, m, b64.dt[i], b64.instance(u0, i).clone(), nsteps)
        assert (b64.instance(u, i) == ref).all()          # same sums, same order: bit-exact


def test_matches_oracle(O, cs, batch_setup):
    levels, meshes, inst, b64, b32, u0 = batch_setup
    u = u0.clone()
    b64.steps(u, 3, graph=False)
    for i, k in enumerate(levels[:4]):
        L = O.Dual(O.icosphere(k)).laplacian0()
        x = b64.instance(u0, i).cpu().numpy()
        for _ in range(3):
            x = x - b64.dt[i] * L.matvec(x)
        y = b64.instance(u, i).cpu().numpy()
        assert np.max(np.abs(y - x)) <= 1e-12 * np.max(np.abs(x))


def test_graph_replays(cs, batch_setup):
    """Replay the captured graph several times with eager work in between and
    check every replay against the eager result."""
    import torch

    levels, meshes, inst, b64, b32, u0 = batch_setup
    ref = u0.clone()
    b64.steps(ref, 5, graph=False)
    u = torch.empty_like(u0)
    for rep in range(3):
        u.copy_(u0)
        b64.steps(u, 5, graph=True)
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 16, dtype=torch.float64, device="cuda").sum()   # eager work between replays
        assert torch.isfinite(scratch)
        assert torch.equal(u, ref), f"replay {rep}"


def test_autotune_and_cache(cs, batch_setup, tmp_path):
    import torch

    levels, meshes, inst, b64, b32, u0 = batch_setup
    u32 = u0.to(torch.float32)
    keep = u32.clone()
    path = str(tmp_path / "tune.txt")
    shape = b32.tune(u32, path)
    assert torch.equal(u32, keep)                        # tuning leaves the state alone
    assert shape["block"] in (64, 128, 256, 512) and shape["rows_per_thread"] in (1, 2, 4)
    assert open(path).read().strip()
    assert b32.tune(u32, path) == shape                  # second call reads the cache
    # tuned shape gives the same result as the default one
    a = u32.clone()
    b32.steps(a, 3, graph=True)
    ref = torch.cat([eager_steps(cs, m, b32.dt[i], b32.instance(u32, i).clone(), 3) for i, m in enumerate(inst)])
    assert torch.equal(a, ref)


def test_f32_against_f64_trajectory(cs, batch_setup):
    import torch

    levels, meshes, inst, b64, b32, u0 = batch_setup
    u64 = u0.clone()
    u32 = u0.to(torch.float32)
    b64.steps(u64, 20)
    b32.steps(u32, 20)
    err = (u32.double() - u64).abs().max().item() / u64.abs().max().item()
    assert err < 1e-5


def test_cache_entries_outside_the_candidates_are_retuned(cs, batch_setup, tmp_path):
    """A cache line whose shape is not one the tuner tries (block 0, a grid
    multiplier of 0) is ignored: the batch tunes again, appends a valid line,
    and its steps still match the single-mesh operator."""
    import torch

    levels, meshes, inst, b64, b32, u0 = batch_setup
    u32 = u0.to(torch.float32)
    path = tmp_path / "tune.txt"
    b32.tune(u32, str(path))
    key = path.read_text().split()[0]
    assert ":gfx" in key                                    # the device architecture is part of the key
    path.write_text(f"{key} 0 1 4\n{key} 256 1 0\n")
    shape = b32.tune(u32, str(path))
    assert shape["block"] in (64, 128, 256, 512) and shape["grid_mult"] in (2, 4, 8)
    assert len(path.read_text().splitlines()) == 3
    a = u32.clone()
    b32.steps(a, 2, graph=False)
    ref = torch.cat([eager_steps(cs, m, b32.dt[i], b32.instance(u32, i).clone(), 2) for i, m in enumerate(inst)])
    assert torch.equal(a, ref)
