"""The precision configuration on the GPU: 512x512x384, 20-cell PML on all
faces, an absorbing slab (conductivity 1e4 S/m), an anisotropic block with a
seeded SPD permittivity (eigenvalues 1-9) and a continuous 1550 nm point
source, 5000 steps, built exactly as bench.py builds it
(bench.precision_components).

  * a copy of the configuration scaled to 64x64x48 (8-cell PML, objects
    scaled with the grid) against the numpy oracle after 300 steps: float64
    within 1e-12 relative L2 (the anisotropic 3x3 product follows numpy's
    einsum association, which is machine-dependent in the last bit), float32
    and the compensated float32 form with their errors stated and bounded
  * full size, float64: exact linearity in the source amplitude (doubling the
    amplitude doubles every field value bit for bit, except for values in the
    subnormal range, where rounding is not scale-invariant)
  * full size, 5000 steps: hip.float32 and hip.float32c against hip.float64,
    relative L2 error and max-norm error (max |difference| / max |float64
    value|) stated; the compensated form within 1e-5 on both

Errors are computed on the device over all E and H values.
"""

import numpy as np
import pytest
import torch

import bench
import fdtd_amd as fdtd
from oracle import fdtd_oracle as orc

pytestmark = pytest.mark.gpu

SHAPE = (512, 512, 384)
PML = 20
STEPS = 5000


def _say(msg):
    print(msg, flush=True)


def _fields(grid):
    return torch.cat([grid.full_precision("E").reshape(-1), grid.full_precision("H").reshape(-1)])


def errors(got, ref):
    """(relative L2, max-norm relative, pointwise relative with a floor of
    1e-3 max|ref|) of got against ref, as floats."""
    d = (got - ref).abs()
    mx = ref.abs().max()
    rel_l2 = float((got - ref).norm() / ref.norm())
    max_norm = float(d.max() / mx)
    pointwise = float((d / ref.abs().clamp_min(1e-3 * mx)).max())
    return rel_l2, max_norm, pointwise


@pytest.fixture(autouse=True)
def _restore_backend():
    yield
    fdtd.set_backend("hip.float64")


def test_precision_small_copy_matches_oracle():
    shape, pml, steps = (64, 64, 48), 8, 300
    sim = bench.build_precision_oracle(orc, shape, pml)
    for _ in range(steps):
        sim.step()
    ref = torch.from_numpy(np.concatenate([sim.E.reshape(-1), sim.H.reshape(-1)])).cuda()
    bounds = {"float64": 1e-12, "float32": 1e-4, "float32c": 1e-7}
    for name, bound in bounds.items():
        fdtd.set_backend(f"hip.{name}")
        grid = bench.build_precision_grid(fdtd, shape, pml)
        grid.run(steps, progress_bar=False)
        got = _fields(grid)
        e = errors(got, ref)
        exact = float((got == ref).double().mean())
        _say(f"[gpu] precision 64x64x48 {name} vs oracle, {steps} steps: rel L2 {e[0]:.2e}, max-norm {e[1]:.2e}, "
             f"pointwise (floor 1e-3 max) {e[2]:.2e}, bit-identical values {100 * exact:.3f} %")
        assert e[0] <= bound and e[1] <= bound
        del grid


def test_precision_full_size_float64_linearity():
    fdtd.set_backend("hip.float64")
    steps = 200
    out = []
    for amp in (1.0, 2.0):
        grid = bench.build_precision_grid(fdtd, SHAPE, PML, amplitude=amp)
        grid.run(steps, progress_bar=False)
        out.append(_fields(grid))
        del grid
        torch.cuda.empty_cache()
    x, y = out
    tiny = 2.0 ** -1000  # far above the subnormal range
    normal = (2 * x).abs() > tiny
    n_diff = int(((y != 2 * x) & normal).sum())
    worst_tiny = float((y - 2 * x)[~normal].abs().max()) if bool((~normal).any()) else 0.0
    _say(f"[gpu] precision 512x512x384 float64 linearity, {steps} steps: {int(normal.sum())} values above "
         f"2^-1000, {n_diff} of them not exactly doubled; largest deviation below: {worst_tiny:.1e}")
    assert float(x.abs().max()) > 0
    assert n_diff == 0
    assert worst_tiny <= 2.0 ** -990


@pytest.fixture(scope="module")
def f64_fields():
    fdtd.set_backend("hip.float64")
    grid = bench.build_precision_grid(fdtd, SHAPE, PML)
    for k in range(STEPS // 1000):
        grid.run(1000, progress_bar=False)
        torch.cuda.synchronize()
        _say(f"[gpu] precision float64 reference: {1000 * (k + 1)} of {STEPS} steps")
    ref = _fields(grid)
    del grid
    torch.cuda.empty_cache()
    return ref


@pytest.mark.parametrize("name,bound", [("float32c", 1e-5), ("float32", 1e-3)])
def test_precision_full_size_error_against_float64(f64_fields, name, bound):
    fdtd.set_backend(f"hip.{name}")
    grid = bench.build_precision_grid(fdtd, SHAPE, PML)
    for k in range(STEPS // 1000):
        grid.run(1000, progress_bar=False)
        torch.cuda.synchronize()
        _say(f"[gpu] precision {name}: {1000 * (k + 1)} of {STEPS} steps")
    e = errors(_fields(grid), f64_fields)
    _say(f"[gpu] precision 512x512x384 {name} vs float64 after {STEPS} steps: rel L2 {e[0]:.2e}, "
         f"max-norm {e[1]:.2e}, pointwise (floor 1e-3 max) {e[2]:.2e}")
    assert e[0] <= bound and e[1] <= bound
    del grid
    torch.cuda.empty_cache()
