"""Boundaries on the HIP backend (API of fdtd/boundaries.py).

 - PeriodicBoundary: plane copy E[0] <- E[-1], H[-1] <- H[0] along one axis.
 - PML: one CPML face per registration (Schneider ch. 11).

Registration swaps the class to a face subclass exactly as the reference
does (boundaries.py:153-180, 295-358). A PML face only keeps what is not
identically zero in the reference's state: along the normal axis the
conductivity profile, so only that component of the b/c coefficient arrays
differs from (exp(-a*S), 0), and only the normal-axis entry of each psi array
ever becomes non-zero. The face therefore stores 1-D b/c profiles (length t)
and two psi scalars per slab cell for each of the E and H sides. The full
reference-shaped arrays (``psi_Ex``, ``bE``, ``phi_E`` ...) are rebuilt on
attribute access for inspection.
"""

import numpy as np
import torch

from . import _native
from .backend import backend as bd
from .typing_ import IntOrSlice, ListOrSlice


class Boundary:
    """an FDTD Boundary [base class]"""

    def __init__(self, name: str = None):
        self.grid = None
        self.name = name

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        # the reference appends before validating the location (boundaries.py:47-60)
        self.grid = grid
        self.grid.boundaries.append(self)
        self.x = self._handle_slice(x)
        self.y = self._handle_slice(y)
        self.z = self._handle_slice(z)
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")

    def _handle_slice(self, s: ListOrSlice) -> IntOrSlice:
        if isinstance(s, list):
            if len(s) > 1:
                raise ValueError("Use slices or single numbers to index the grid for a boundary")
            return s[0]
        if isinstance(s, slice):
            if s.start is not None and s.stop is not None and (s.start == s.stop or abs(s.start - s.stop) == 1):
                return s.start
            return s
        raise ValueError("Invalid grid indexing used for boundary")

    def promote_dtypes_to_complex(self):
        # boundaries.py:106-107: the base class converts CPML state, which only
        # a PML has; anything else fails on its first attribute
        raise AttributeError(f"'{type(self).__name__}' object has no attribute 'phi_E'")

    # plan interface
    def _plan_key(self):
        return (type(self).__name__,)

    def _add_to_plan(self, plan):
        """An unplaced boundary does nothing, as in the reference."""

    def __repr__(self):
        return f"{self.__class__.__name__}(name={repr(self.name)})"

    def __str__(self):
        s = "    " + repr(self) + "\n"

        def _h(v):
            return str(v).replace("slice(", "").replace(")", "").replace(", ", ":").replace("None", "")

        s += f"        @ x={_h(self.x)}, y={_h(self.y)}, z={_h(self.z)}".replace(":,", ",")
        if s[-1] == ":":
            s = s[:-1]
        return s + "\n"


# ------------------------------------------------------------------ periodic
class PeriodicBoundary(Boundary):
    """A periodic boundary; becomes _PeriodicBoundaryX/Y/Z on registration."""

    axis = None

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        super()._register_grid(grid=grid, x=x, y=y, z=z)
        for axis, (pos, cls) in enumerate(((self.x, _PeriodicBoundaryX), (self.y, _PeriodicBoundaryY),
                                           (self.z, _PeriodicBoundaryZ))):
            if pos == 0 or pos == -1:
                self.__class__ = cls
                n = "xyz"[axis]
                if hasattr(grid, f"_{n}low_boundary") or hasattr(grid, f"_{n}high_boundary"):
                    raise AttributeError(f"grid already has an {n}low/{n}high boundary!")
                setattr(grid, f"_{n}low_boundary", self)
                setattr(grid, f"_{n}high_boundary", self)
                return
        raise IndexError(
            "A periodic boundary should be placed at the boundary of the grid using a single index (either 0 or -1)")

    def _add_to_plan(self, plan):
        if self.axis is not None:
            plan.add_periodic(self.axis)


class _PeriodicBoundaryX(PeriodicBoundary):
    axis = 0


class _PeriodicBoundaryY(PeriodicBoundary):
    axis = 1


class _PeriodicBoundaryZ(PeriodicBoundary):
    axis = 2


# ----------------------------------------------------------------------- PML
def pml_profiles(thickness, high, a, k, courant_number, dtype=np.float64):
    """1-D normal-axis CPML coefficients (bE, cE, bH, cH), each of length t.

    sigma(v) = 40 v^3 / (t+1)^4 sampled at t-1/2 ... 1/2 (E) and t-1 ... 1
    then 0 (H) on a low face, mirrored on a high face (boundaries.py:291-293,
    502-625); b = exp(-(sigma/k + a) S), c = (b - 1) sigma / (sigma k + a k^2)
    (boundaries.py:395-407). Same numpy operations, element by element, as the
    reference applies to its full arrays, in the field dtype."""
    t = int(thickness)
    f = np.dtype(dtype).type

    def sig(v):
        return 40 * v ** 3 / (t + 1) ** 4

    if high:
        sE = sig(np.arange(0.5, t + 0.5, 1.0).astype(dtype))
        sH_part = sig(np.arange(1.0, t, 1.0).astype(dtype))
    else:
        sE = sig(np.arange(t - 0.5, -0.5, -1.0).astype(dtype))
        sH_part = sig(np.arange(t - 1.0, 0, -1.0).astype(dtype))
    sE = np.asarray(sE, dtype=dtype)
    sH = np.zeros(t, dtype=dtype)
    sH[:-1] = sH_part
    bE = np.exp(-(sE / k + a) * courant_number)
    cE = (bE - f(1.0)) * sE / (sE * k + a * k ** 2)
    bH = np.exp(-(sH / k + a) * courant_number)
    cH = (bH - f(1.0)) * sH / (sH * k + a * k ** 2)
    return (np.asarray(bE, dtype), np.asarray(cE, dtype), np.asarray(bH, dtype), np.asarray(cH, dtype),
            np.asarray(sE, dtype), sH)


class PML(Boundary):
    """A perfectly matched layer; becomes _PMLXlow ... _PMLZhigh on registration."""

    axis = None
    high = None

    def __init__(self, a: float = 1e-8, name: str = None):
        super().__init__(name=name)
        self.k = 1.0
        self.thickness = 0
        self.a = a

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        super()._register_grid(grid=grid, x=x, y=y, z=z)
        faces = ((self.x, _PMLXlow, _PMLXhigh, "x"), (self.y, _PMLYlow, _PMLYhigh, "y"),
                 (self.z, _PMLZlow, _PMLZhigh, "z"))
        for s, low_cls, high_cls, n in faces:
            if (s.start is None or s.start == 0) and (s.stop is not None) and (s.stop > 0):
                self.__class__ = low_cls
                if hasattr(grid, f"_{n}low_boundary"):
                    raise AttributeError(f"grid already has an {n}low boundary!")
                setattr(grid, f"_{n}low_boundary", self)
                self._calculate_parameters(thickness=s.stop)
                return
            if (s.start is not None) and (s.stop is None) and (s.start < 0):
                self.__class__ = high_cls
                if hasattr(grid, f"_{n}high_boundary"):
                    raise AttributeError(f"grid already has an {n}high boundary!")
                setattr(grid, f"_{n}high_boundary", self)
                self._calculate_parameters(thickness=-s.start)
                return
        raise IndexError("not a valid slice for a PML. Make sure the slice is at the border of the PML")

    def _handle_slice(self, s: ListOrSlice) -> slice:
        if isinstance(s, list):
            raise ValueError("One can only use slices to index the grid for a PML")
        if isinstance(s, slice):
            return s
        raise ValueError("Invalid grid indexing used for boundary")

    def _calculate_parameters(self, thickness: int = 10):
        self.thickness = thickness
        g = self.grid
        n = [g.Nx, g.Ny, g.Nz]
        n[self.axis] = thickness
        self.shape = tuple(n)
        prof = pml_profiles(thickness, self.high, self.a, self.k, g.courant_number, bd.numpy_float)
        self._bE, self._cE, self._bH, self._cH, self._sigE, self._sigH = prof
        dev = g._E.device
        self._psiE = torch.zeros((2,) + self.shape, dtype=g.mat_dtype, device=dev)
        self._psiH = torch.zeros((2,) + self.shape, dtype=g.mat_dtype, device=dev)
        sl = [slice(None)] * 3
        sl[self.axis] = slice(-thickness, None) if self.high else slice(None, thickness)
        self.loc = tuple(sl) + (slice(None),)

    # plan interface
    def _plan_key(self):
        return (type(self).__name__, self.thickness, self.a, id(self._psiE))

    def _add_to_plan(self, plan):
        if self.axis is None:
            raise RuntimeError(f"{self!r} is registered in the grid but not placed on a face")
        plan.add_pml(self.axis, self.high, self.thickness, self._bE, self._cE, self._bH, self._cH,
                     self._psiE, self._psiH)

    # reference-shaped views of the compact state, for inspection
    def _full(self, profile, other):
        g = self.grid
        arr = np.full(self.shape + (3,), other, dtype=bd.numpy_float)
        shp = [1, 1, 1]
        shp[self.axis] = -1
        arr[..., self.axis] = profile.reshape(shp)
        return torch.from_numpy(arr).to(g._E.device)

    @property
    def bE(self):
        return self._full(self._bE, np.exp(-(0.0 / self.k + self.a) * self.grid.courant_number))

    @property
    def bH(self):
        return self._full(self._bH, np.exp(-(0.0 / self.k + self.a) * self.grid.courant_number))

    @property
    def cE(self):
        return self._full(self._cE, 0.0)

    @property
    def cH(self):
        return self._full(self._cH, 0.0)

    @property
    def sigmaE(self):
        return self._full(self._sigE, 0.0)

    @property
    def sigmaH(self):
        return self._full(self._sigH, 0.0)

    def _psi_full(self, store, comp):
        out = torch.zeros(self.shape + (3,), dtype=store.dtype, device=store.device)
        comps = [c for c in range(3) if c != self.axis]
        if comp in comps:
            out[..., self.axis] = store[comps.index(comp)]
        return out

    def _phi(self, store):
        out = torch.zeros(self.shape + (3,), dtype=store.dtype, device=store.device)
        for q, c in enumerate(c for c in range(3) if c != self.axis):
            out[..., c] = store[q] if self.axis == (c + 1) % 3 else -store[q]
        return out

    psi_Ex = property(lambda self: self._psi_full(self._psiE, 0))
    psi_Ey = property(lambda self: self._psi_full(self._psiE, 1))
    psi_Ez = property(lambda self: self._psi_full(self._psiE, 2))
    psi_Hx = property(lambda self: self._psi_full(self._psiH, 0))
    psi_Hy = property(lambda self: self._psi_full(self._psiH, 1))
    psi_Hz = property(lambda self: self._psi_full(self._psiH, 2))
    phi_E = property(lambda self: self._phi(self._psiE))
    phi_H = property(lambda self: self._phi(self._psiH))

    def promote_dtypes_to_complex(self):
        """The CPML state becomes complex (boundaries.py:106-116)."""
        cdt = self.grid.field_dtype
        self._psiE = self._psiE.to(cdt)
        self._psiH = self._psiH.to(cdt)

    def __repr__(self):
        return f"PML(name={repr(self.name)})"


class _PMLXlow(PML):
    axis, high = 0, False


class _PMLXhigh(PML):
    axis, high = 0, True


class _PMLYlow(PML):
    axis, high = 1, False


class _PMLYhigh(PML):
    axis, high = 1, True


class _PMLZlow(PML):
    axis, high = 2, False


class _PMLZhigh(PML):
    axis, high = 2, True


def DomainBorderPML(grid, border_cells=5):
    """PML on all six faces (boundaries.py:628-656). The tangential extents of
    the keys are ignored by the PML, so the faces overlap at edges and corners
    exactly as in the reference."""
    if grid.Nx < border_cells * 2 or grid.Ny < border_cells * 2 or grid.Nz < border_cells * 2:
        raise IndexError("PML border_cells larger than domain!")
    grid[:, :, 0:border_cells] = PML()
    grid[:, :, -border_cells:] = PML()
    grid[0:border_cells, :, border_cells:-border_cells] = PML()
    grid[-border_cells:, :, border_cells:-border_cells] = PML()
    grid[border_cells:-border_cells, 0:border_cells, border_cells:-border_cells] = PML()
    grid[border_cells:-border_cells, -border_cells:, border_cells:-border_cells] = PML()
