"""Sources on the HIP backend (API of fdtd/sources.py).

Injection runs on the device after the boundaries of each half step, in
registration order. The per-step scalar a source multiplies into its cells
is computed on the host with Python ``math`` for every time index, exactly as
the reference computes it inside update_E/update_H, and stored in a float64
table on the device, so no host work or synchronisation happens per step:

 - PointSource: E[x,y,z,2] += A*hanning(...) or A*sin(2 pi q/P + phi)  (sources.py:93-109)
 - LineSource:  E[x_i,y_i,z_i,2] += profile_i * (hanning | sin)       (sources.py:278-297)
 - PlaneSource: E[box, Epol] = A*sin(...); H[box, Hpol] = A*sin(...)  (sources.py:476-486)
"""

from math import pi, sin

import numpy as np
import torch

from . import _native
from .backend import backend as bd
from .typing_ import List, ListOrSlice, Number, Tuple
from .waveforms import hanning


class _SourceBase:
    def _wave_params(self):
        return (self.period, self.amplitude, self.phase_shift, getattr(self, "pulse", False),
                getattr(self, "cycle", None), getattr(self, "hanning_dt", None))

    def _plan_key(self):
        return (type(self).__name__,) + self._wave_params() + (id(getattr(self, "profile", None)),)

    def _pulse_window(self, q):
        """hanning window value at step q, or None after the window (sources.py:97-105)"""
        t1 = int(2 * pi / (self.frequency * self.hanning_dt / self.cycle))
        if q < t1:
            return hanning(self.frequency, q * self.hanning_dt, self.cycle)
        return None

    def update_E(self):
        raise NotImplementedError("source injection runs inside the HIP step")

    def update_H(self):
        raise NotImplementedError("source injection runs inside the HIP step")

    def __repr__(self):
        return (
            f"{self.__class__.__name__}(period={self.period}, "
            f"amplitude={self.amplitude}, phase_shift={self.phase_shift}, "
            f"name={repr(self.name)})"
        )


class PointSource(_SourceBase):
    """A source placed at a single point (grid cell) in the grid"""

    def __init__(self, period: Number = 15, amplitude: float = 1.0, phase_shift: float = 0.0, name: str = None,
                 pulse: bool = False, cycle: int = 5, hanning_dt: float = 10.0):
        self.grid = None
        self.period = period
        self.amplitude = amplitude
        self.phase_shift = phase_shift
        self.name = name
        self.pulse = pulse
        self.cycle = cycle
        self.frequency = 1.0 / period
        self.hanning_dt = hanning_dt if hanning_dt is not None else 0.5 / self.frequency

    def _register_grid(self, grid, x: Number, y: Number, z: Number):
        self.grid = grid
        self.grid.sources.append(self)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        try:
            (x,), (y,), (z,) = x, y, z
        except (TypeError, ValueError):
            raise ValueError("a point source should be placed on a single grid cell.")
        self.x, self.y, self.z = grid._handle_tuple((x, y, z))
        self.period = grid._handle_time(self.period)
        self.frequency = 1.0 / self.period

    def _waveform(self, q):
        if self.pulse:
            h = self._pulse_window(q)
            return self.amplitude * h if h is not None else 0.0
        return self.amplitude * sin(2 * pi * q / self.period + self.phase_shift)

    def _add_to_plan(self, plan):
        cell = self.grid._cell(self.x, self.y, self.z)
        plan.add_source(_native.SRC_POINT, [cell], None, self.amplitude, 2, 0)

    def __str__(self):
        return "    " + repr(self) + "\n" + f"        @ x={self.x}, y={self.y}, z={self.z}\n"


class LineSource(_SourceBase):
    """A source along a line in the FDTD grid"""

    def __init__(self, period: Number = 15, amplitude: float = 1.0, phase_shift: float = 0.0, name: str = None,
                 pulse: bool = False, cycle: int = 5, hanning_dt: float = 10.0):
        self.grid = None
        self.period = period
        self.amplitude = amplitude
        self.phase_shift = phase_shift
        self.name = name
        self.pulse = pulse
        self.cycle = cycle
        self.frequency = 1.0 / period
        self.hanning_dt = hanning_dt if hanning_dt is not None else 0.5 / self.frequency

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        self.grid = grid
        self.grid.sources.append(self)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        self.x, self.y, self.z = self._handle_slices(x, y, z)
        self.period = grid._handle_time(self.period)
        self.frequency = 1.0 / self.period
        self.profile = line_profile(self.x, self.y, self.z, self.amplitude, bd.numpy_float)

    def _handle_slices(self, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice) -> Tuple[List, List, List]:
        return line_indices(self.grid, x, y, z, convert=self.grid._handle_distance, min_points=2)

    def _waveform(self, q):
        if self.pulse:
            h = self._pulse_window(q)
            return h if h is not None else 0.0
        return sin(2 * pi * q / self.period + self.phase_shift)

    def _add_to_plan(self, plan):
        g = self.grid
        cells = [g._cell(a, b, c) for a, b, c in zip(self.x, self.y, self.z)]
        prof = self.profile
        prof = prof.detach().cpu().numpy() if torch.is_tensor(prof) else np.asarray(prof)
        plan.add_source(_native.SRC_LINE, cells, prof.astype(np.float64), self.amplitude, 2, 0)

    def __str__(self):
        s = "    " + repr(self) + "\n"
        s += f"        @ x=[{self.x[0]}, ... , {self.x[-1]}], y=[{self.y[0]}, ... , {self.y[-1]}], " \
             f"z=[{self.z[0]}, ... , {self.z[-1]}]\n"
        return s


def line_indices(grid, x, y, z, convert, min_points=None):
    """Index lists of a line along the diagonal of the box spanned by the key
    (sources.py:209-276 with convert=grid._handle_distance, min_points=2;
    detectors.py:63-112 with convert=identity): linspace(endpoint=False)
    truncated to int; list keys are taken as the points themselves."""
    if isinstance(x, list) and isinstance(y, list) and isinstance(z, list):
        if len(x) != len(y) or len(y) != len(z) or len(z) != len(x):
            raise IndexError("sources require grid to be indexed with slices or equal length list-indices")
        return [convert(v) for v in x], [convert(v) for v in y], [convert(v) for v in z]
    ends = []
    for s, n in ((x, grid.Nx), (y, grid.Ny), (z, grid.Nz)):
        if isinstance(s, list):
            s = slice(convert(s[0]), convert(s[-1]), None)
        ends.append((convert(s.start if s.start is not None else 0), convert(s.stop if s.stop is not None else n)))
    m = max(abs(b - a) for a, b in ends)
    if min_points is not None and m < min_points:
        raise ValueError("a LineSource should consist of at least two gridpoints")
    return tuple([int(v) for v in np.array(np.linspace(a, b, m, endpoint=False), np.int64)] for a, b in ends)


def line_profile(xs, ys, zs, amplitude, dtype=np.float64):
    """exp(-d^4 / (2 (max d^2 / 2)^2)) over the squared distance d^2 to the
    middle point, normalised to sum 1 and scaled by the amplitude
    (sources.py:197-207)."""
    L = len(xs)
    vect = np.array((np.array(xs) - xs[L // 2]) ** 2 + (np.array(ys) - ys[L // 2]) ** 2
                    + (np.array(zs) - zs[L // 2]) ** 2, dtype)
    profile = np.exp(-(vect ** 2) / (2 * (0.5 * vect.max()) ** 2))
    profile /= profile.sum()
    profile *= amplitude
    return profile


class PlaneSource(_SourceBase):
    """A source along a plane in the FDTD grid (a hard source)"""

    def __init__(self, period: Number = 15, amplitude: float = 1.0, phase_shift: float = 0.0, name: str = None,
                 polarization: str = "z"):
        self.grid = None
        self.period = period
        self.amplitude = amplitude
        self.phase_shift = phase_shift
        self.name = name
        self.polarization = polarization

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        self.grid = grid
        self.grid.sources.append(self)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        self.x, self.y, self.z = self._handle_slices(x, y, z)
        self.period = grid._handle_time(self.period)
        self.frequency = 1.0 / self.period
        shape = (self.x.stop - self.x.start, self.y.stop - self.y.start, self.z.stop - self.z.start)
        self.profile = self.amplitude * torch.ones(shape, dtype=grid.dtype, device=grid._E.device)
        self._profile_amplitude = self.amplitude  # the profile is fixed at registration

    def _handle_slices(self, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        """ascending slices, flatness checks, E/H polarisation (sources.py:390-474)"""
        g = self.grid
        out = []
        for s, n in ((x, g.Nx), (y, g.Ny), (z, g.Nz)):
            if not isinstance(s, slice):
                if isinstance(s, list):
                    (s,) = s
                s = slice(g._handle_distance(s), g._handle_distance(s) + 1, None)
            a = g._handle_distance(s.start if s.start is not None else 0)
            b = g._handle_distance(s.stop if s.stop is not None else n)
            out.append(slice(a, b) if a < b else (slice(b, a) if a > b else slice(a, a + 1)))
        x, y, z = out
        lens = [x.stop - x.start, y.stop - y.start, z.stop - z.start]
        if lens.count(0) > 0:
            raise ValueError("Given location for PlaneSource results in slices of length 0!")
        if lens.count(1) == 0:
            raise ValueError("Given location for PlaneSource is not a 2D plane!")
        if lens.count(1) > 1:
            raise ValueError(
                "Given location for PlaneSource should have no more than one dimension in which it's flat.\n"
                "Use a LineSource for lower dimensional sources.")
        self._Epol = "xyz".index(self.polarization)
        if (x.stop - x.start == 1 and self.polarization == "x") or \
           (y.stop - y.start == 1 and self.polarization == "y") or \
           (z.stop - z.start == 1 and self.polarization == "z"):
            raise ValueError("PlaneSource cannot be polarized perpendicular to the orientation of the plane.")
        probe, first, second = [(z, 1, 2), (z, 0, 2), (y, 0, 1)][self._Epol]
        self._Hpol = first if probe.stop - probe.start == 1 else second
        return x, y, z

    def _waveform(self, q):
        return sin(2 * pi * q / self.period + self.phase_shift)

    def _add_to_plan(self, plan):
        g = self.grid
        cells = [g._cell(a, b, c) for a in range(self.x.start, self.x.stop)
                 for b in range(self.y.start, self.y.stop) for c in range(self.z.start, self.z.stop)]
        plan.add_source(_native.SRC_PLANE, cells, None, self._profile_amplitude, self._Epol, self._Hpol)

    def __repr__(self):
        return (
            f"{self.__class__.__name__}(period={self.period}, "
            f"amplitude={self.amplitude}, phase_shift={self.phase_shift}, "
            f"name={repr(self.name)}, polarization={repr(self.polarization)})"
        )

    def __str__(self):
        s = "    " + repr(self) + "\n"
        s += f"        @ x=[{self.x.start}, ... , {self.x.stop}], y=[{self.y.start}, ... , {self.y.stop}], " \
             f"z=[{self.z.start}, ... , {self.z.stop}]\n"
        return s


class SoftArbitraryPointSource:
    """A resistive point source driven by a waveform array (sources.py:504-637).

    In every E half step, on the device: v = waveform_array[q] (0 past its
    end), I = the current its CurrentDetector measured after the last H half
    step (0 at q == 0), u = v + impedance * I when impedance > 0, else v;
    E[x, y, z, 2] += u / grid_spacing. ``input_voltage`` / ``source_voltage``
    are the reference's per-step lists of v and u. Unlike the reference,
    naming the source works (its detector is named ``<name>_I``)."""

    def __init__(self, waveform_array, name: str = None, impedance: float = 0.0):
        self.grid = None
        self.name = name
        self.current_detector = None
        if torch.is_tensor(waveform_array):
            waveform_array = waveform_array.detach().cpu().numpy()
        self.waveform_array = np.asarray(waveform_array)
        self.impedance = impedance
        self._vin = None
        self._vout = None
        self._count = 0

    def _register_grid(self, grid, x: Number, y: Number, z: Number):
        from .detectors import CurrentDetector

        self.grid = grid
        self.grid.sources.append(self)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        try:
            (x,), (y,), (z,) = x, y, z
        except (TypeError, ValueError):
            raise ValueError("a point source should be placed on a single grid cell.")
        self.x, self.y, self.z = grid._handle_tuple((x, y, z))
        self._cell = grid._cell(self.x, self.y, self.z)
        self.current_detector = CurrentDetector(name=None if self.name is None else self.name + "_I")
        grid[x, y, z] = self.current_detector
        self._reserve(0)

    def _waveform(self, q):
        return float(self.waveform_array[q]) if q < self.waveform_array.shape[0] else 0.0

    def _reserve(self, extra):
        need = self._count + int(extra)
        cap = 0 if self._vin is None else self._vin.shape[0]
        if self._vin is not None and need <= cap:
            return False
        g = self.grid
        new_cap = max(need, 2 * cap, 64)
        vin = torch.zeros(new_cap, dtype=g.dtype, device=g._E.device)
        vout = torch.zeros(new_cap, dtype=g.dtype, device=g._E.device)
        if self._vin is not None:
            vin[: self._count] = self._vin[: self._count]
            vout[: self._count] = self._vout[: self._count]
        self._vin, self._vout = vin, vout
        return True

    def _plan_key(self):
        return (type(self).__name__, self.impedance, id(self.waveform_array), id(self._vin))

    def _add_to_plan(self, plan):
        curs = [d for d in self.grid.detectors if type(d).__name__ == "CurrentDetector"]
        det = next(i for i, d in enumerate(curs) if d is self.current_detector)
        plan.add_soft_source(self._cell, float(self.impedance), self.grid.grid_spacing, det, self._vin, self._vout,
                             self._count)

    def trace(self, field="input_voltage"):
        """All recorded voltages as one device tensor (steps, 1, 1, 1)."""
        buf = {"input_voltage": self._vin, "source_voltage": self._vout}[field]
        return buf[: self._count].reshape(self._count, 1, 1, 1)

    @property
    def input_voltage(self) -> List:
        t = self.trace("input_voltage")
        return [t[i] for i in range(t.shape[0])]

    @property
    def source_voltage(self) -> List:
        t = self.trace("source_voltage")
        return [t[i] for i in range(t.shape[0])]

    def update_E(self):
        raise NotImplementedError("source injection runs inside the HIP step")

    def update_H(self):
        pass

    def __repr__(self):
        return f"{self.__class__.__name__}()"

    def __str__(self):
        return "    " + repr(self) + "\n" + f"        @ x={self.x}, y={self.y}, z={self.z}\n"
