"""fdtd_batch_run_static -- the batched run whose tables live in the
library's static device array, with no device allocation -- against the
allocating entry points (fdtd_batch_create + fdtd_batch_set_launch +
fdtd_batch_run) on identical copies of the same grids. The arithmetic is the
same kernels', so the fields and CPML state must be bit-identical. Two calls
in a row: the second over fewer and smaller grids, from a later time index
and with another launch shape, so that tables left by the first call would
show. A batch whose tables exceed the array is refused with an error.
"""

import numpy as np
import pytest
import torch

import fdtd_amd as fdtd
from fdtd_amd import _native

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def make_grids(seed, shapes):
    """Grids of the given shapes with PML faces, objects of every kind on
    some, one or two point sources and random initial fields (float64)."""
    rs = np.random.RandomState(seed)
    grids = []
    for g, shape in enumerate(shapes):
        grid = fdtd.Grid(shape, grid_spacing=50e-9, permittivity=rs.uniform(1.0, 4.0, size=shape + (1,)),
                         courant_number=0.5)
        for axis in range(3):
            for high in (False, True):
                key = [slice(None)] * 3
                key[axis] = slice(-3, None) if high else slice(0, 3)
                grid[tuple(key)] = fdtd.PML()
        if g % 2 == 1:
            grid[2:6, 3:7, 1:5] = fdtd.Object(permittivity=2.5)
            grid[7:9, 1:4, 4:8] = fdtd.AbsorbingObject(permittivity=1.5, conductivity=1e5)
        if g % 3 == 2:
            grid[1:3, 5:8, 6:9] = fdtd.AnisotropicObject(permittivity=rs.uniform(1.5, 3.0, size=(2, 3, 3)))
        for _ in range(1 + g % 2):
            p = [int(rs.randint(0, s)) for s in shape]
            grid[p[0], p[1], p[2]] = fdtd.PointSource(period=float(rs.uniform(8, 20)), pulse=bool(g % 2))
        grid.E = torch.from_numpy(rs.uniform(-1, 1, size=shape + (3,)))
        grid.H = torch.from_numpy(rs.uniform(-1, 1, size=shape + (3,)))
        grids.append(grid)
    return grids


def _state(grids):
    out = []
    for g in grids:
        out += [g._E, g._H]
        for b in g.boundaries:
            out += [b._psiE, b._psiH]
    return [t.clone() for t in out]


def _run_allocating(grids, steps, launch):
    q0 = grids[0].time_steps_passed
    plans = [g._prepare(steps, steps, steps) for g in grids]
    batch = _native.Batch(plans)
    batch.set_launch(*launch)
    batch.run(q0, steps, torch.cuda.current_stream().cuda_stream)
    for g in grids:
        g.time_steps_passed += steps
    torch.cuda.synchronize()
    del batch


def _run_static(grids, steps, launch):
    q0 = grids[0].time_steps_passed
    plans = [g._prepare(steps, steps, steps) for g in grids]
    _native.batch_run_static(plans, *launch, q0, steps, torch.cuda.current_stream().cuda_stream)
    for g in grids:
        g.time_steps_passed += steps
    torch.cuda.synchronize()


CALLS = [
    # seed, grid shapes, steps before the call (grid.run), steps of the call, launch shape
    (11, [(22, 18, 20), (16, 24, 13), (19, 14, 26), (25, 17, 15), (14, 21, 18)], 0, 23, (256, 2, 0)),
    (12, [(12, 10, 11), (10, 13, 9)], 3, 17, (256, 1, 5)),
]


def test_static_batch_matches_allocating_batch():
    fdtd.set_backend("hip.float64")
    for seed, shapes, pre, steps, launch in CALLS:
        want = make_grids(seed, shapes)
        got = make_grids(seed, shapes)
        for g in want + got:
            if pre:
                g.run(pre, progress_bar=False)
        _say(f"[gpu] static batch: {len(shapes)} grids {shapes}, q0 {pre}, {steps} steps, launch {launch}")
        before = _state(got)
        _run_allocating(want, steps, launch)
        _run_static(got, steps, launch)
        a, b = _state(want), _state(got)
        assert len(a) == len(b)
        for i, (x, y, z) in enumerate(zip(a, b, before)):
            assert torch.equal(x, y), f"state tensor {i} differs"
        changed = sum(int(not torch.equal(y, z)) for y, z in zip(b, before))
        assert changed >= len(shapes) * 2
        assert all(g.time_steps_passed == pre + steps for g in got)


def test_static_batch_refuses_tables_beyond_the_array():
    fdtd.set_backend("hip.float64")
    # one work item per x plane for grids one cell wide in y and z: 18 grids
    # of 60000 planes list more items (4-byte entries) than the array holds
    n = 60_000
    grids = [fdtd.Grid((n, 1, 1), grid_spacing=50e-9, courant_number=0.5) for _ in range(18)]
    for g in grids:
        g[n // 2, 0, 0] = fdtd.PointSource(period=20.0)
    plans = [g._prepare(2, 2, 2) for g in grids]
    before = [g._E.clone() for g in grids]
    with pytest.raises(RuntimeError, match="static tables"):
        _native.batch_run_static(plans, 256, 1, 0, 0, 2, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert all(torch.equal(g._E, b) for g, b in zip(grids, before))
