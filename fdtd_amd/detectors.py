"""Detectors on the HIP backend (API of fdtd/detectors.py).

Sampling runs on the device at the end of every half step, after the
sources: the three field components at the detector's cells are copied into
a device trace buffer (capacity, n, 3) that grows geometrically before runs
that need room. ``det.E`` / ``det.H`` present the reference's per-step list
(detectors.py:114-124) as views into that buffer; ``det.trace("E")`` gives
the whole (steps, n, 3) tensor without building the list.
"""

import numpy as np
import torch

from .sources import line_indices
from .typing_ import List, ListOrSlice, Tuple


class _DetectorBase:
    def __init__(self, name=None):
        self.grid = None
        self.name = name
        self._bufE = None
        self._bufH = None
        self._count_E = 0
        self._count_H = 0

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        self.grid = grid
        self.grid.detectors.append(self)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        self.x, self.y, self.z = self._handle_slices(x, y, z)
        self._cells = [grid._cell(a, b, c) for a, b, c in self._points()]
        self._reserve(0)

    # -- device trace buffers
    def _reserve(self, extra):
        """Make room for `extra` more samples; True if the buffers moved."""
        need = max(self._count_E, self._count_H) + int(extra)
        cap = 0 if self._bufE is None else self._bufE.shape[0]
        if self._bufE is not None and need <= cap:
            return False
        new_cap = max(need, 2 * cap, 64)
        n = len(self._cells)
        g = self.grid
        bufE = torch.zeros((new_cap, n, 3), dtype=g.field_dtype, device=g._E.device)
        bufH = torch.zeros((new_cap, n, 3), dtype=g.field_dtype, device=g._E.device)
        if self._bufE is not None:
            bufE[: self._count_E] = self._bufE[: self._count_E]
            bufH[: self._count_H] = self._bufH[: self._count_H]
        self._bufE, self._bufH = bufE, bufH
        return True

    def _promote(self, cdt):
        """Complex traces once the fields are complex (samples so far kept)."""
        if self._bufE is not None:
            self._bufE, self._bufH = self._bufE.to(cdt), self._bufH.to(cdt)

    def _plan_key(self):
        return (type(self).__name__, id(self._bufE))

    def _add_to_plan(self, plan):
        plan.add_detector(self._cells, self._bufE, self._bufH, self._count_E, self._count_H)

    def _sample_shape(self):
        return (len(self._cells), 3)

    def trace(self, field="E"):
        """All samples so far as one device tensor (steps, *sample_shape)."""
        buf, n = (self._bufE, self._count_E) if field == "E" else (self._bufH, self._count_H)
        return buf[:n].reshape((n,) + self._sample_shape())

    @property
    def E(self) -> List:
        t = self.trace("E")
        return [t[i] for i in range(t.shape[0])]

    @property
    def H(self) -> List:
        t = self.trace("H")
        return [t[i] for i in range(t.shape[0])]

    def detect_E(self):
        raise NotImplementedError("detector sampling runs inside the HIP step")

    def detect_H(self):
        raise NotImplementedError("detector sampling runs inside the HIP step")

    def detector_values(self):
        """outputs what detector detects"""
        return {"E": self.E, "H": self.H}

    def __repr__(self):
        return f"{self.__class__.__name__}(name={repr(self.name)})"

    def __str__(self):
        s = "    " + repr(self) + "\n"
        s += f"        @ x=[{self.x[0]}, ... , {self.x[-1]}], y=[{self.y[0]}, ... , {self.y[-1]}], " \
             f"z=[{self.z[0]}, ... , {self.z[-1]}]\n"
        return s


class LineDetector(_DetectorBase):
    """A detector along a line in the FDTD grid"""

    def _handle_slices(self, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice) -> Tuple[List, List, List]:
        return line_indices(self.grid, x, y, z, convert=lambda v: v)

    def _points(self):
        return zip(self.x, self.y, self.z)


class BlockDetector(_DetectorBase):
    """A detector over a block in the FDTD grid (inclusive ranges)"""

    def _handle_slices(self, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice) -> Tuple[List, List, List]:
        if isinstance(x, list) and isinstance(y, list) and isinstance(z, list):
            if len(x) != len(y) or len(y) != len(z) or len(z) != len(x):
                raise IndexError("sources require grid to be indexed with slices or equal length list-indices")
            return x, y, z
        out = []
        for s, n in ((x, self.grid.Nx), (y, self.grid.Ny), (z, self.grid.Nz)):
            if isinstance(s, list):
                s = slice(s[0], s[-1], None)
            a = s.start if s.start is not None else 0
            b = s.stop if s.stop is not None else n
            out.append([int(v) for v in np.arange(a, b + 1)])  # inclusive (detectors.py:236-238)
        return tuple(out)

    def _points(self):
        return ((a, b, c) for a in self.x for b in self.y for c in self.z)

    def _sample_shape(self):
        return (len(self.x), len(self.y), len(self.z), 3)


class CurrentDetector:
    """Loop current around z at each point of a block (detectors.py:284-476),
    sampled on the device after every H half step; ``I`` is the reference's
    per-step list, each entry an (nx, ny, nz) view indexable as I[-1][i][j][k]."""

    def __init__(self, name=None):
        self.grid = None
        self.orientation = None
        self.name = name
        self._buf = None
        self._count_E = 0  # unused: detect_E records nothing (detectors.py:385-386)
        self._count_H = 0

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        self.grid = grid
        self.grid.detectors.append(self)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        self.x, self.y, self.z = self._handle_slices(x, y, z)
        for v, n in ((self.x, grid.Nx), (self.y, grid.Ny), (self.z, grid.Nz)):
            if any(p < -n or p >= n for p in v):
                raise IndexError("current detector point outside the grid")
        self._pts = [[p % n for p in v] for v, n in ((self.x, grid.Nx), (self.y, grid.Ny), (self.z, grid.Nz))]
        self._reserve(0)

    def _handle_slices(self, x, y, z):
        """Point lists as given, otherwise inclusive ranges (detectors.py:330-372)."""
        if isinstance(x, list) and isinstance(y, list) and isinstance(z, list):
            if len(x) != len(y) or len(y) != len(z) or len(z) != len(x):
                raise IndexError("sources require grid to be indexed with slices or equal length list-indices")
            return x, y, z
        out = []
        for s, n in ((x, self.grid.Nx), (y, self.grid.Ny), (z, self.grid.Nz)):
            if isinstance(s, list):
                s = slice(s[0], s[-1], None)
            a = s.start if s.start is not None else 0
            b = s.stop if s.stop is not None else n
            out.append([int(v) for v in np.arange(a, b + 1)])
        return tuple(out)

    def _points(self):
        px, py, pz = self._pts
        return [(a, b, c) for a in px for b in py for c in pz]

    def _promote(self, cdt):
        raise NotImplementedError("CurrentDetector is not supported on the complex-field path")

    def _reserve(self, extra):
        need = self._count_H + int(extra)
        cap = 0 if self._buf is None else self._buf.shape[0]
        if self._buf is not None and need <= cap:
            return False
        g = self.grid
        buf = torch.zeros((max(need, 2 * cap, 64), len(self._points())), dtype=g.dtype, device=g._E.device)
        if self._buf is not None:
            buf[: self._count_H] = self._buf[: self._count_H]
        self._buf = buf
        return True

    def _plan_key(self):
        return (type(self).__name__, id(self._buf))

    def _add_to_plan(self, plan):
        pts = self._points()
        plan.add_current_detector([p[0] for p in pts], [p[1] for p in pts], [p[2] for p in pts],
                                  self.grid.grid_spacing, self._buf, self._count_H)

    def trace(self, field="I"):
        """All current samples as one device tensor (steps, nx, ny, nz)."""
        if field != "I":
            raise KeyError("a current detector records I only")
        n = self._count_H
        return self._buf[:n].reshape(n, len(self.x), len(self.y), len(self.z))

    @property
    def I(self) -> List:
        t = self.trace("I")
        return [t[i] for i in range(t.shape[0])]

    def single_point_current(self, px, py, pz):
        """The loop current at one point now (detectors.py:388-460), from the
        current H."""
        H = self.grid.H
        dx = self.grid.grid_spacing
        v1 = (H[px, py - 1, pz, 0] - H[px, py, pz, 0]) * dx
        v2 = (H[px, py, pz, 1] - H[px - 1, py, pz, 1]) * dx
        c1 = v1 + v2
        v1 = (H[px, py - 1, pz - 1, 0] - H[px, py, pz - 1, 0]) * dx
        v2 = v2 + (H[px, py, pz - 1, 1] - H[px - 1, py, pz - 1, 1]) * dx
        c2 = v1 + v2
        return (c1 + c2) / 2.0

    def detect_E(self):
        raise NotImplementedError("detector sampling runs inside the HIP step")

    def detect_H(self):
        raise NotImplementedError("detector sampling runs inside the HIP step")

    def detector_values(self):
        return {"I": self.I}

    def __repr__(self):
        return f"{self.__class__.__name__}(name={repr(self.name)})"

    def __str__(self):
        s = "    " + repr(self) + "\n"
        s += f"        @ x=[{self.x[0]}, ... , {self.x[-1]}], y=[{self.y[0]}, ... , {self.y[-1]}], " \
             f"z=[{self.z[0]}, ... , {self.z[-1]}]\n"
        return s
