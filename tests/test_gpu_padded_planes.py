"""The padded x-plane layout on small grids (FDTD_HIP_PITCH=1: every x plane
whose ny*nz is not a multiple of 32 cells padded to one; by default only
planes of 4096 cells and more are padded). The layout moves data, never
arithmetic, so every result must equal the unpadded run's bit for bit:

  * every golden scenario the reference wrote, in float64, against its
    fixture (fields, CPML state, detector and source traces)
  * the compensated storage (records) and the complex-field path against the
    same runs unpadded
  * batched grids against the same grids stepped alone
  * the gradient of the periodic + anisotropic case against the unpadded one
"""

import numpy as np
import pytest
import torch

import fdtd_amd as fdtd
import scenarios
from fdtd_amd.batch import GridBatch

import adjoint_scenarios as A
from test_gpu_parity import PSI_NAMES, SCENARIOS, _assert_same, _build, _np, _say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fp64():
    fdtd.set_backend("hip.float64")
    yield
    fdtd.set_backend("hip.float64")


def _padded(monkeypatch, on):
    monkeypatch.setenv("FDTD_HIP_PITCH", "1" if on else "0")


@pytest.mark.parametrize("name", SCENARIOS)
def test_padded_planes_match_reference_fixture(golden, name, monkeypatch):
    _padded(monkeypatch, True)
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, comps = _build(sc)
    _say(f"[gpu] padded planes, fixture {name}: shape {grid.shape}, pitch {grid._px} (plane {grid.Ny * grid.Nz})")
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    _assert_same(_np(grid.E), g["E"], "E")
    _assert_same(_np(grid.H), g["H"], "H")
    for i, c in enumerate(comps):
        p = f"out/{i}/"
        cls = str(g[f"reg/{i}/class"])
        if cls.startswith("_PML"):
            for q, nm in enumerate(PSI_NAMES[c.axis]):
                store = c._psiE if nm[0] == "E" else c._psiH
                _assert_same(_np(store[q % 2]), g[p + "psi_" + nm], f"psi_{nm} of face {i}")
        elif cls in ("LineDetector", "BlockDetector"):
            _assert_same(_np(c.trace("E")), g[p + "E"], f"detector {i} E")
            _assert_same(_np(c.trace("H")), g[p + "H"], f"detector {i} H")
        elif cls == "CurrentDetector":
            _assert_same(_np(c.trace("I")), g[p + "I"], f"current detector {i}")


def _run_fixture(name, backend, monkeypatch, on):
    _padded(monkeypatch, on)
    fdtd.set_backend(backend)
    sc = scenarios.get_scenario(name)
    grid, comps = _build(sc)
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    return grid, [_np(grid.full_precision("E")), _np(grid.full_precision("H"))]


@pytest.mark.parametrize("backend,name", [("hip.float32c", "manual"), ("hip.float32c", "periodic_x_first"),
                                          ("hip.float32", "manual"), ("hip.float64", "complex_object")])
def test_padded_planes_are_bit_invisible(backend, name, monkeypatch):
    g0, ref = _run_fixture(name, backend, monkeypatch, False)
    g1, got = _run_fixture(name, backend, monkeypatch, True)
    _say(f"[gpu] padded planes, {backend} {name}: pitch {g0._px} -> {g1._px}")
    assert g1._px != g0._px
    for what, a, b in zip(("E", "H"), got, ref):
        assert np.array_equal(a, b), f"{backend} {name}: {what} differs with padded planes"


def test_padded_planes_batch_matches_grids_stepped_alone(monkeypatch):
    from test_gpu_batch import _fields, make_grids

    _padded(monkeypatch, True)
    fdtd.set_backend("hip.float32c")
    steps = 23
    ref = make_grids(5)
    for g in ref:
        g.run(steps, progress_bar=False)
    want = _fields(ref)
    grids = make_grids(5)
    assert any(g._px != g.Ny * g.Nz for g in grids)
    GridBatch(grids, graph_steps=10).run(steps)
    for i, ((eg, hg), (ew, hw)) in enumerate(zip(_fields(grids), want)):
        assert np.array_equal(eg, ew) and np.array_equal(hg, hw), f"grid {i} differs"


def test_padded_planes_gradient_equals_unpadded(monkeypatch):
    from fdtd_amd.adjoint import IntensityGradient

    sc = A.TINY_PA
    out = []
    for on in (False, True):
        _padded(monkeypatch, on)
        perm = A.design_permittivity(sc)
        grid, design = A.build(fdtd, sc, torch.from_numpy(perm))
        prob = IntensityGradient(grid, design, sc["plane"], sc["steps"], "sqrt")
        J, g = prob.gradient(torch.from_numpy(perm))
        out.append((grid._px, J, g.cpu().numpy()))
    _say(f"[gpu] padded planes, gradient tiny_pa: pitch {out[0][0]} -> {out[1][0]}, J {out[1][1]:.6e}")
    assert out[0][0] != out[1][0]
    assert out[0][1] == out[1][1] and np.array_equal(out[0][2], out[1][2])
