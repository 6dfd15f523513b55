"""Batched stepping (fdtd_amd.batch.GridBatch) against the grids stepped one
by one: the same kernels' arithmetic, so the results must be bit-identical,
direct launches and captured-graph replays alike, for every storage type.
"""

import numpy as np
import pytest
import torch

import fdtd_amd as fdtd
from fdtd_amd.batch import GridBatch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def make_grids(seed, n=5):
    """n grids of uneven shapes with PML faces (some missing), objects of
    every kind on some, 1-2 point sources and random initial fields."""
    rs = np.random.RandomState(seed)
    grids = []
    for g in range(n):
        shape = tuple(int(v) for v in rs.randint(10, 25, size=3))
        grid = fdtd.Grid(shape, grid_spacing=50e-9, permittivity=rs.uniform(1.0, 4.0, size=shape + (1,)),
                         courant_number=0.5)
        t = int(rs.randint(2, 5))
        for axis in range(3):
            for high in (False, True):
                if rs.rand() < 0.8:
                    key = [slice(None)] * 3
                    key[axis] = slice(-t, None) if high else slice(0, t)
                    grid[tuple(key)] = fdtd.PML()
        if g % 2 == 1:
            grid[2:6, 3:7, 1:5] = fdtd.Object(permittivity=2.5)
            grid[7:9, 1:4, 4:8] = fdtd.AbsorbingObject(permittivity=1.5, conductivity=1e5)
        if g % 3 == 2:
            grid[1:3, 5:8, 6:9] = fdtd.AnisotropicObject(permittivity=rs.uniform(1.5, 3.0, size=(2, 3, 3)))
        for _ in range(int(rs.randint(1, 3))):
            p = [int(rs.randint(0, s)) for s in shape]
            grid[p[0], p[1], p[2]] = fdtd.PointSource(period=float(rs.uniform(8, 20)), pulse=bool(rs.rand() < 0.5))
        grid.E = torch.from_numpy(rs.uniform(-1, 1, size=shape + (3,)))
        grid.H = torch.from_numpy(rs.uniform(-1, 1, size=shape + (3,)))
        grids.append(grid)
    return grids


def _fields(grids):
    return [(g.full_precision("E").cpu().numpy(), g.full_precision("H").cpu().numpy()) for g in grids]


@pytest.mark.parametrize("dealing", ["contiguous", "cyclic"])
@pytest.mark.parametrize("backend", ["hip.float64", "hip.float32", "hip.float32c"])
def test_batch_matches_grids_stepped_alone(backend, dealing, monkeypatch):
    """Small batches deal their tiles to the XCDs contiguously by default;
    FDTD_HIP_CYCLIC=1 forces the cyclic dealing of large batches (idle
    positions at both ends of every grid's span), which must not change a
    bit."""
    fdtd.set_backend(backend)
    steps = 37
    ref = make_grids(3)
    for g in ref:
        g.run(steps, progress_bar=False)
    want = _fields(ref)
    if dealing == "cyclic":
        monkeypatch.setenv("FDTD_HIP_CYCLIC", "1")
    for graph, shape in ((False, (256, 1, 0)), (True, (512, 2, 64)), (True, (256, 3, 1000))):
        grids = make_grids(3)
        batch = GridBatch(grids, graph_steps=10)
        batch.set_launch(*shape)
        _say(f"[gpu] batch {backend} {dealing} graph={graph} launch={shape}")
        batch.run(steps, graph=graph)
        _say(f"[gpu]   work list positions: {batch._batch.work_items()}")
        got = _fields(grids)
        for i, ((eg, hg), (ew, hw)) in enumerate(zip(got, want)):
            assert np.array_equal(eg, ew) and np.array_equal(hg, hw), f"grid {i} differs ({backend}, {shape})"
        assert all(g.time_steps_passed == steps for g in grids)
    fdtd.set_backend("hip.float64")


def test_batch_graph_replays_interleaved_with_eager_work():
    """Replays of one captured graph, with host work and syncs between the
    runs, against the grids stepped alone after each run."""
    fdtd.set_backend("hip.float64")
    ref = make_grids(7, n=4)
    grids = make_grids(7, n=4)
    batch = GridBatch(grids, graph_steps=8)
    for rep in range(4):
        for g in ref:
            g.run(16, progress_bar=False)
        batch.run(16, graph=True)
        torch.cuda.synchronize()
        _ = [float(g.E.abs().sum()) for g in grids]  # eager device work between replays
        for (eg, hg), (ew, hw) in zip(_fields(grids), _fields(ref)):
            assert np.array_equal(eg, ew) and np.array_equal(hg, hw), f"replay {rep}"


def test_autotune_restores_state_and_keeps_results():
    fdtd.set_backend("hip.float32")
    ref = make_grids(11, n=3)
    for g in ref:
        g.run(12, progress_bar=False)
    grids = make_grids(11, n=3)
    batch = GridBatch(grids, graph_steps=4)
    times = batch.autotune(steps=3, candidates=[(256, 1, 0), (512, 2, 0), (256, 4, 512)])
    assert len(times) == 3 and batch.launch in times
    assert all(g.time_steps_passed == 0 for g in grids)
    batch.run(12)
    for (eg, hg), (ew, hw) in zip(_fields(grids), _fields(ref)):
        assert np.array_equal(eg, ew) and np.array_equal(hg, hw)
    fdtd.set_backend("hip.float64")


def test_batch_rejects_ineligible_grids():
    fdtd.set_backend("hip.float64")
    g1, g2 = make_grids(5, n=2)
    g2[3, :, :] = fdtd.LineDetector(name="det")
    with pytest.raises(ValueError):
        GridBatch([g1, g2])
    g3 = fdtd.Grid((12, 12, 12), grid_spacing=50e-9, courant_number=0.5)
    g3[0, :, :] = fdtd.PeriodicBoundary()
    g3[6, 6, 6] = fdtd.PointSource(period=10.0)
    with pytest.raises(RuntimeError):
        GridBatch([g1, g3]).run(2)


# ------------------------------------------------------------------ C4
# The batching configuration itself (bench.batch_specs / bench.batch_grid):
# 64 grids with edges drawn from 64-192 cells (uneven z rows), smoothed-noise
# permittivity in [1, 12], 10-cell PML on all faces, a pulsed point source.

def _zero_state(batch):
    for t in batch._state():
        t.zero_()
    for g in batch.grids:
        g.time_steps_passed = 0


def test_c4_batch_matches_grids_stepped_alone_and_oracle():
    """float32, 50 steps: the 64-grid batch (graph replays of 10 steps with
    the default launch, direct launches of a persistent launch shape) bit-
    identical to every grid stepped alone; the three smallest grids' float32
    error against the oracle stated and bounded."""
    import bench
    from oracle import fdtd_oracle as orc

    fdtd.set_backend("hip.float32")
    steps = 50
    specs = bench.batch_specs(64)
    grids = [bench.batch_grid(fdtd, s) for s in specs]
    _say(f"[gpu] c4 batch: {len(grids)} grids, {sum(g.Nx * g.Ny * g.Nz for g in grids)} cells, "
         f"nz from {min(g.Nz for g in grids)} to {max(g.Nz for g in grids)}")
    for g in grids:
        g.run(steps, progress_bar=False)
    want = [(g._E.clone(), g._H.clone()) for g in grids]
    batch = GridBatch(grids, graph_steps=10)
    for graph, shape in ((True, (256, 2, 0)), (False, (256, 1, 2048))):
        _zero_state(batch)
        batch.set_launch(*shape)
        _say(f"[gpu] c4 batch float32 graph={graph} launch={shape}")
        batch.run(steps, graph=graph)
        torch.cuda.synchronize()
        for i, (g, (ew, hw)) in enumerate(zip(grids, want)):
            assert torch.equal(g._E, ew) and torch.equal(g._H, hw), f"grid {i} {g.shape} differs ({shape})"
        assert all(g.time_steps_passed == steps for g in grids)
    smallest = sorted(range(len(specs)), key=lambda i: np.prod(specs[i][0]))[:3]
    for i in smallest:
        sim = bench.batch_oracle(orc, specs[i])
        for _ in range(steps):
            sim.step()
        ref = np.concatenate([sim.E.reshape(-1), sim.H.reshape(-1)])
        got = np.concatenate([grids[i].E.double().cpu().numpy().reshape(-1),
                              grids[i].H.double().cpu().numpy().reshape(-1)])
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        m = np.abs(got - ref).max() / np.abs(ref).max()
        _say(f"[gpu] c4 grid {i} {specs[i][0]} float32 batch vs oracle, {steps} steps: rel L2 {e:.2e}, "
             f"max-norm {m:.2e}")
        assert e <= 1e-4 and m <= 1e-4
    fdtd.set_backend("hip.float64")


def test_c4_batch_float64_smallest_grids_bit_exact_against_oracle():
    """float64, 40 steps of the whole 64-grid batch (graph replays): the three
    smallest grids bit for bit against the oracle."""
    import bench
    from oracle import fdtd_oracle as orc

    fdtd.set_backend("hip.float64")
    steps = 40
    specs = bench.batch_specs(64)
    grids = [bench.batch_grid(fdtd, s) for s in specs]
    GridBatch(grids, graph_steps=10).run(steps)
    torch.cuda.synchronize()
    smallest = sorted(range(len(specs)), key=lambda i: np.prod(specs[i][0]))[:3]
    for i in smallest:
        sim = bench.batch_oracle(orc, specs[i])
        for _ in range(steps):
            sim.step()
        eq_E = np.array_equal(grids[i].E.cpu().numpy(), sim.E)
        eq_H = np.array_equal(grids[i].H.cpu().numpy(), sim.H)
        _say(f"[gpu] c4 grid {i} {specs[i][0]} float64 batch vs oracle, {steps} steps: bit-exact E {eq_E}, H {eq_H}, "
             f"max |E| {np.abs(sim.E).max():.3e}")
        assert np.abs(sim.E).max() > 0
        assert eq_E and eq_H
