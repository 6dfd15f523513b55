"""The reference's own test cases for the components on the step path
(tests/test_detectors.py, tests/test_fft.py, tests/test_sources.py of
flaport/fdtd), run on the HIP backend with the grid of its fixtures.py:
10x10x10 cells of 100 nm, default Courant number. They assert shapes and
behaviour; the values themselves are covered by the fixture tests."""

import numpy as np
import pytest
import torch

import fdtd_amd as fdtd
from fdtd_amd.sources import SoftArbitraryPointSource
from fdtd_amd.waveforms import normalized_gaussian_pulse
from oracle import fdtd_oracle as orc

pytestmark = pytest.mark.gpu


def _arr(history):
    """A detector/source history (list of device tensors) as one numpy array,
    as the reference's bd.array(...) builds it."""
    return torch.stack(list(history)).cpu().numpy() if len(history) else np.zeros((0,))


@pytest.fixture
def grid():
    fdtd.set_backend("hip.float64")
    return fdtd.Grid(shape=(10, 10, 10), grid_spacing=100e-9, permittivity=1.0, permeability=1.0,
                     courant_number=None)


def test_current_detector_shape(grid):
    """test_detectors.py:91-100: loop current and block detector record one
    sample per step."""
    edetector = fdtd.BlockDetector()
    cdetector = fdtd.CurrentDetector()
    grid[4, 4, 4] = cdetector
    grid[4, 4, 5] = edetector
    grid.run(10)
    I = _arr(cdetector.I)
    E = _arr(edetector.E)
    assert I.shape[0:2] == E.shape[0:2]
    assert I.shape[0] == 10


def test_soft_source_registers_its_current_detector(grid):
    """test_detectors.py:103-107: a SoftArbitraryPointSource registers a
    CurrentDetector; its one-sample waveform is 0 after the first step
    (sources.py:604-609)."""
    source = SoftArbitraryPointSource(np.zeros(1), impedance=50.0)
    grid[4, 4, 4] = source
    grid.run(100)
    assert isinstance(grid.detectors[0], fdtd.CurrentDetector)
    vin = _arr(source.input_voltage)
    assert vin.shape[0] == 100 and not np.any(vin)


def test_absorbing_current_detector_with_host_writes(grid):
    """test_detectors.py:32-88: half steps driven by hand with a host write
    of E every step and the time index advanced by hand; one current and
    one field sample per step."""
    grid[3:5, 3:5, 3:5] = fdtd.AbsorbingObject(permittivity=1.0, conductivity=1e9)
    cd = fdtd.CurrentDetector(name="Dave")
    grid[4, 4, 4] = cd
    grid[4:4, 4:4, 4:4] = fdtd.BlockDetector()
    n = 200
    for _ in range(n):
        grid.update_E()
        grid.E[4, 4, 4, 2] = normalized_gaussian_pulse(grid.time_passed, (n / 10.0) * grid.time_step,
                                                       center=(n / 2) * grid.time_step)
        grid.update_H()
        grid.time_steps_passed += 1
    current = _arr(grid.detectors[0].I).reshape(n)
    efield = _arr(grid.detectors[1].E).reshape(n, 3)[:, 2]
    assert np.all(np.isfinite(current)) and np.all(np.isfinite(efield))
    # the detector samples E after each E half step, before the host write
    # of that step: sample s holds the value written at step s-1, updated
    assert np.abs(efield).max() > 0


def test_fft_of_a_block_detector(grid):
    """test_fft.py:37-50: FFT with no steps taken, then after 100 steps."""
    detector = fdtd.BlockDetector()
    grid[4, 4, 4] = detector
    assert fdtd.FrequencyRoutines(grid, detector).FFT() == ([], [])
    grid[2, 2, 2] = fdtd.PointSource(period=20)
    grid.run(100)
    freqs, spectrum = fdtd.FrequencyRoutines(grid, detector).FFT()
    assert freqs.shape == spectrum.shape and spectrum.is_complex()


def test_soft_source_impedance(grid):
    """test_fft.py:53-58: the impedance routine on a soft source's records."""
    source = SoftArbitraryPointSource(np.zeros(1), impedance=50.0)
    grid[4, 4, 4] = source
    grid.run(100)
    f, z = fdtd.FrequencyRoutines(grid, source).impedance()
    assert f.shape == z.shape


def test_padding(grid):
    """test_fft.py:17-34: padding for a 1000-sample record at 1 GHz."""
    fr = fdtd.FrequencyRoutines(grid, None)
    pad, end_time = fr.compute_padding(torch.zeros(1000), dt=1e-9)
    assert end_time == pytest.approx(1000 * 1e-9)
    assert pad == orc.fr_padding(1000, 1e-9)  # 1001: the reference's formula rounds up past 2n
