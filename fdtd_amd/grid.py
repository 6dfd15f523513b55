"""The FDTD grid on the HIP backend (API of fdtd/grid.py:80-514).

State lives in HBM as component-major torch tensors, E and H of shape
(3, Nx, Ny, Nz) (z fastest), and ``grid.E`` / ``grid.H`` /
``grid.inverse_permittivity`` / ``grid.inverse_permeability`` are
(Nx, Ny, Nz, 3) views of them, so user indexing such as
``grid.E[x, y, z, 2] = v`` writes device memory directly and assignment
``grid.E = array`` copies into it.

``update_E`` / ``update_H`` / ``step`` / ``run`` do not walk the component
hooks the way the reference does (grid.py:275-325); they hand the whole half
step to libfdtd_hip, which applies the same updates in the same order inside
fused kernels. The native plan is (re)built lazily whenever the registered
components or the arrays they own change.
"""

import os
from datetime import datetime
from os import makedirs, path

import numpy as np
import torch

from . import _native
from . import constants as const
from .backend import backend as bd
from .typing_ import Number, Tuple


def _soa_view(t):
    return t.permute(1, 2, 3, 0)


def plane_pitch(ny, nz):
    """Cells from one x plane of a device array to the next: ny*nz rounded up
    to a multiple of 32, so that every x plane of every per-cell array starts
    on a 128-byte line (a wavefront's 64 consecutive cells then span two lines
    instead of three). Planes below 4096 cells, where the padding would cost
    more than 1 % of the memory and alignment does not pay, stay unpadded.
    FDTD_HIP_PITCH: 0 never pads, 1 pads every plane that is not a multiple of
    32 cells (as a test of the padded layout on small grids)."""
    p = int(ny) * int(nz)
    mode = os.environ.get("FDTD_HIP_PITCH", "auto")
    if mode == "0" or p % 32 == 0 or (mode != "1" and p < 4096):
        return p
    return (p + 31) // 32 * 32


def curl_E(E):
    """Forward-difference curl of an E-type (Nx,Ny,Nz,3) array (fdtd/grid.py:29-51).
    Host-side helper for inspection; the time step never calls it."""
    curl = torch.zeros_like(E) if torch.is_tensor(E) else np.zeros_like(E)
    curl[:, :-1, :, 0] += E[:, 1:, :, 2] - E[:, :-1, :, 2]
    curl[:, :, :-1, 0] -= E[:, :, 1:, 1] - E[:, :, :-1, 1]
    curl[:, :, :-1, 1] += E[:, :, 1:, 0] - E[:, :, :-1, 0]
    curl[:-1, :, :, 1] -= E[1:, :, :, 2] - E[:-1, :, :, 2]
    curl[:-1, :, :, 2] += E[1:, :, :, 1] - E[:-1, :, :, 1]
    curl[:, :-1, :, 2] -= E[:, 1:, :, 0] - E[:, :-1, :, 0]
    return curl


def curl_H(H):
    """Backward-difference curl of an H-type (Nx,Ny,Nz,3) array (fdtd/grid.py:54-76).
    Host-side helper for inspection; the time step never calls it."""
    curl = torch.zeros_like(H) if torch.is_tensor(H) else np.zeros_like(H)
    curl[:, 1:, :, 0] += H[:, 1:, :, 2] - H[:, :-1, :, 2]
    curl[:, :, 1:, 0] -= H[:, :, 1:, 1] - H[:, :, :-1, 1]
    curl[:, :, 1:, 1] += H[:, :, 1:, 0] - H[:, :, :-1, 0]
    curl[1:, :, :, 1] -= H[1:, :, :, 2] - H[:-1, :, :, 2]
    curl[1:, :, :, 2] += H[1:, :, :, 1] - H[:-1, :, :, 1]
    curl[:, 1:, :, 2] -= H[:, 1:, :, 0] - H[:, :-1, :, 0]
    return curl


class Grid:
    """The FDTD grid: fields, materials, registered components and the loop."""

    def __init__(
        self,
        shape: Tuple[Number, Number, Number],
        grid_spacing: float = 155e-9,
        permittivity: float = 1.0,
        permeability: float = 1.0,
        courant_number: float = None,
    ):
        self.grid_spacing = float(grid_spacing)
        self.Nx, self.Ny, self.Nz = self._handle_tuple(shape)
        self.D = int(self.Nx > 1) + int(self.Ny > 1) + int(self.Nz > 1)
        max_courant_number = float(self.D) ** (-0.5)
        if courant_number is None:
            self.courant_number = 0.99 * max_courant_number
        elif courant_number > max_courant_number:
            raise ValueError(f"courant_number {courant_number} too high for a {self.D}D simulation")
        else:
            self.courant_number = float(courant_number)
        self.time_step = self.courant_number * self.grid_spacing / const.c

        self.dtype = bd.float                  # fields
        self.mat_dtype = bd.material_float     # materials, object arrays, CPML state
        self.compensated = bool(bd.compensated)
        self.complex = False  # promote_dtypes_to_complex (grid.py:353-356)
        dev = bd.device
        # Device arrays keep x planes `_px` cells apart (plane_pitch); E, H and
        # the materials are (3, Nx, Ny, Nz) views of (3, Nx, _px) storage.
        self._px = plane_pitch(self.Ny, self.Nz)
        self._E = self._alloc_soa(self.dtype)
        self._H = self._alloc_soa(self.dtype)
        self._inv_eps = self._alloc_soa(self.mat_dtype)
        self._inv_mu = self._alloc_soa(self.mat_dtype)
        self._fill_inverse(self._inv_eps, permittivity)
        self._fill_inverse(self._inv_mu, permeability)
        if self.compensated:
            # 8 compensation bits per float32 field value. A host write into
            # grid.E/H[...] changes the float part only: the byte left there
            # lies below the new float's last bit, so the written value is
            # kept to float32 precision and every other cell keeps 32 bits.
            # One 16-byte record per cell, (Nx, Ny, Nz, 4) float32: x, y, z and
            # a 32-bit word whose byte c holds component c's bits. _E/_H are
            # the (3, Nx, Ny, Nz) float views of the records, _E_lo/_H_lo the
            # words, so that the kernels read a cell with one 16-byte access.
            self._E_rec = self._alloc_cells((4,), torch.float32)
            self._H_rec = self._alloc_cells((4,), torch.float32)
            self._E = self._E_rec[..., :3].permute(3, 0, 1, 2)
            self._H = self._H_rec[..., :3].permute(3, 0, 1, 2)
            self._E_lo = self._E_rec[..., 3].view(torch.int32)
            self._H_lo = self._H_rec[..., 3].view(torch.int32)
        # materials read through a <=256-entry float64 palette: "auto" uses one
        # in the compensated mode when the grid has few distinct materials
        self.material_palette = "auto"
        self._palette = None
        self._palette_key = None

        self.time_steps_passed = 0
        self.sources = []
        self.boundaries = []
        self.detectors = []
        self.objects = []
        self.folder = None

        self._plan = None
        self._plan_key = None
        self._wave = None
        self._wave_key = None

    # ------------------------------------------------------------ storage
    def _alloc_soa(self, dtype):
        """Zeroed (3, Nx, Ny, Nz) device array in the pitched layout."""
        P = self.Ny * self.Nz
        store = torch.zeros((3, self.Nx, self._px), dtype=dtype, device=bd.device)
        return store[:, :, :P].unflatten(2, (self.Ny, self.Nz))

    def _alloc_cells(self, trail, dtype):
        """Zeroed (Nx, Ny, Nz, *trail) device array in the pitched layout."""
        P = self.Ny * self.Nz
        store = torch.zeros((self.Nx, self._px) + tuple(trail), dtype=dtype, device=bd.device)
        return store[:, :P].unflatten(1, (self.Ny, self.Nz))

    # ------------------------------------------------------------ materials
    def _fill_inverse(self, dst, value):
        """dst (3,Nx,Ny,Nz) <- 1 / value broadcast to (Nx,Ny,Nz,3), computed on
        the host in the field dtype (grid.py:141-151)."""
        npdt = bd.numpy_float
        if torch.is_tensor(value):
            value = value.detach().cpu().numpy()
        v = np.asarray(value)
        if v.ndim == 3:
            v = v[..., None]
        v = v.astype(npdt, copy=False)
        if v.ndim == 0:
            dst.fill_(float(np.ones((), npdt) / v))
            return
        if v.ndim == 4 and v.shape[3] == 1 and v.shape[:3] == (self.Nx, self.Ny, self.Nz):
            inv = np.ones(v.shape[:3], npdt) / v[..., 0]
            t = torch.from_numpy(np.ascontiguousarray(inv)).to(dst.device)
            for c in range(3):
                dst[c].copy_(t)
            return
        inv = np.ones((self.Nx, self.Ny, self.Nz, 3), npdt) / v
        dst.copy_(torch.from_numpy(np.ascontiguousarray(np.moveaxis(inv, 3, 0))).to(dst.device))

    def _set_soa(self, dst, value):
        view = _soa_view(dst)
        if torch.is_tensor(value) and value.data_ptr() == view.data_ptr() and value.stride() == view.stride():
            return
        if not torch.is_tensor(value):
            value = torch.as_tensor(np.asarray(value))
        if self.compensated and (dst is self._E or dst is self._H):
            # whole-field assignment keeps 32 significant bits: the float part
            # and the next 8 mantissa bits (rounded at bit 21 of the double);
            # the material arrays are plain float64 and take the copy below
            v = value.to(device=dst.device, dtype=torch.float64).expand_as(view)
            bits = v.contiguous().view(torch.int64) + (1 << 20)
            q = (bits >> 21) & 0xFF  # (Nx, Ny, Nz, 3)
            word = q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16)
            (self._E_lo if dst is self._E else self._H_lo).copy_(word.to(torch.int32))
            view.copy_((bits & ~((1 << 29) - 1)).view(torch.float64).to(torch.float32))
            return
        view.copy_(value.to(device=dst.device, dtype=dst.dtype).expand_as(view))

    @property
    def E(self):
        return _soa_view(self._E)

    @E.setter
    def E(self, value):
        self._set_soa(self._E, value)

    @property
    def H(self):
        return _soa_view(self._H)

    def full_precision(self, field="E"):
        """The field ("E" or "H") as a float64 (Nx, Ny, Nz, 3) tensor. On
        hip.float32c this decodes the 40-bit values (the float plus its 8
        compensation bits, inserted below the float's mantissa) that E/H --
        the float part alone -- round away; elsewhere it is a cast."""
        soa = self._E if field == "E" else self._H
        if self.complex:
            return _soa_view(soa.to(torch.complex128))
        if not self.compensated:
            return _soa_view(soa.to(torch.float64))
        word = (self._E_lo if field == "E" else self._H_lo).to(torch.int64)
        q = torch.stack([(word >> (8 * c)) & 0xFF for c in range(3)])  # (3, Nx, Ny, Nz)
        bits = soa.to(torch.float64).view(torch.int64) | (q << 21)
        return _soa_view(bits.view(torch.float64))

    @H.setter
    def H(self, value):
        self._set_soa(self._H, value)

    @property
    def inverse_permittivity(self):
        return _soa_view(self._inv_eps)

    @inverse_permittivity.setter
    def inverse_permittivity(self, value):
        self._set_soa(self._inv_eps, value)

    @property
    def inverse_permeability(self):
        return _soa_view(self._inv_mu)

    @inverse_permeability.setter
    def inverse_permeability(self, value):
        self._set_soa(self._inv_mu, value)

    # ------------------------------------------------------- unit handling
    def _handle_distance(self, distance: Number) -> int:
        """transform a distance to an integer number of gridpoints (grid.py:171-175)"""
        if not isinstance(distance, int):
            return int(float(distance) / self.grid_spacing + 0.5)
        return distance

    def _handle_time(self, time: Number) -> int:
        """transform a time value to an integer number of timesteps (grid.py:177-181)"""
        if not isinstance(time, int):
            return int(float(time) / self.time_step + 0.5)
        return time

    def _handle_tuple(self, shape):
        if len(shape) != 3:
            raise ValueError(
                f"invalid grid shape {shape}\ngrid shape should be a 3D tuple containing floats or ints")
        x, y, z = shape
        return self._handle_distance(x), self._handle_distance(y), self._handle_distance(z)

    def _handle_slice(self, s: slice) -> slice:
        f = lambda v: v if not isinstance(v, float) else self._handle_distance(v)
        return slice(f(s.start), f(s.stop), f(s.step))

    def _handle_single_key(self, key):
        try:
            len(key)
            return [self._handle_distance(k) for k in key]
        except TypeError:
            if isinstance(key, slice):
                return self._handle_slice(key)
            return [self._handle_distance(key)]

    @property
    def x(self) -> float:
        return self.Nx * self.grid_spacing

    @property
    def y(self) -> float:
        return self.Ny * self.grid_spacing

    @property
    def z(self) -> float:
        return self.Nz * self.grid_spacing

    @property
    def shape(self) -> Tuple[int, int, int]:
        return (self.Nx, self.Ny, self.Nz)

    @property
    def time_passed(self) -> float:
        return self.time_steps_passed * self.time_step

    # ------------------------------------------------------------- plan
    def _cell(self, x, y, z):
        """Linear cell index with numpy's negative-index rule; IndexError if outside."""
        out = []
        for v, n in ((x, self.Nx), (y, self.Ny), (z, self.Nz)):
            v = int(v)
            if v < -n or v >= n:
                raise IndexError(f"index {v} is out of bounds for a grid axis of size {n}")
            out.append(v + n if v < 0 else v)
        return out[0] * self._px + out[1] * self.Nz + out[2]

    def _components(self):
        return list(self.boundaries) + list(self.objects) + list(self.sources) + list(self.detectors)

    def _palette_wanted(self):
        if self.complex:
            return False
        mode = self.material_palette
        return bool(self.compensated) if mode == "auto" else bool(mode)

    # ------------------------------------------------------ complex fields
    @property
    def field_dtype(self):
        """dtype of E, H, the CPML state and detector traces."""
        return self._E.dtype

    @property
    def obj_dtype(self):
        """dtype of the object arrays: the material dtype, complex on the
        complex-field path."""
        if not self.complex:
            return self.mat_dtype
        return torch.complex128 if self.mat_dtype == torch.float64 else torch.complex64

    def promote_dtypes_to_complex(self):
        """E and H become complex, and so does the CPML state of every boundary
        registered so far (grid.py:353-356); called when an Object with a
        complex permittivity is registered. A PeriodicBoundary has no such
        state and fails as in the reference. Complex fields are stepped by the
        complex128 (hip.float64) or complex64 (hip.float32) kernels, bit for
        bit against numpy's complex arithmetic in float64."""
        if self.compensated:
            raise NotImplementedError("complex fields are not supported by hip.float32c")
        if self.complex:
            return
        cdt = torch.complex128 if self.dtype == torch.float64 else torch.complex64
        E, H = self._alloc_soa(cdt), self._alloc_soa(cdt)
        E.copy_(self._E)
        H.copy_(self._H)
        self._E, self._H = E, H
        self.complex = True
        for det in self.detectors:
            det._promote(cdt)
        for boundary in self.boundaries:
            boundary.promote_dtypes_to_complex()

    def _material_palette(self):
        """(eps_id, eps_table, mu_id, mu_table) or None when a grid material has
        more than 256 distinct (x, y, z) triples. Rebuilt when the material
        arrays change (their torch version counters)."""
        key = (self._inv_eps._version, self._inv_mu._version, self._inv_eps.data_ptr(), self._inv_mu.data_ptr())
        if key == self._palette_key:
            return self._palette
        out = []
        for arr in (self._inv_eps, self._inv_mu):
            rows = arr.reshape(3, -1).transpose(0, 1)
            uniq, idx = torch.unique(rows, dim=0, return_inverse=True)
            if uniq.shape[0] > 256:
                out = None
                break
            tab = torch.zeros((256, 3), dtype=torch.float64, device=arr.device)
            tab[: uniq.shape[0]] = uniq.to(torch.float64)
            ids = self._alloc_cells((), torch.uint8)
            ids.copy_(idx.to(torch.uint8).view(self.Nx, self.Ny, self.Nz))
            out += [ids, tab]
        self._palette = tuple(out) if out else None
        self._palette_key = key
        return self._palette

    def _current_plan_key(self):
        key = [len(self.boundaries), len(self.objects), len(self.sources), len(self.detectors), self.courant_number,
               self._palette_wanted(), self._E.data_ptr(), self._H.data_ptr()]
        if self._palette_wanted():
            key += [self._inv_eps._version, self._inv_mu._version]
        for comp in self._components():
            key.append(id(comp))
            key.extend(comp._plan_key())
        return tuple(key)

    def _ensure_plan(self):
        if bd.device.type != "cuda":
            raise RuntimeError("the HIP backend needs a HIP device to step (arrays are in host memory)")
        key = self._current_plan_key()
        if self._plan is not None and key == self._plan_key:
            return self._plan
        if self.complex:
            for b in self.boundaries:
                if hasattr(b, "_psiE") and not b._psiE.is_complex():
                    # the reference's in-place CPML update fails the same way
                    raise TypeError(f"Cannot cast ufunc 'add' output from dtype('{self.field_dtype}') to the real "
                                    f"CPML state of {b!r}, registered after the fields became complex")
            dtype_code = _native.C128 if self.dtype == torch.float64 else _native.C64
        elif self.compensated:
            dtype_code = _native.F32C
        else:
            dtype_code = _native.F64 if self.dtype == torch.float64 else _native.F32
        plan = _native.Plan(dtype_code, (self.Nx, self.Ny, self.Nz), self.courant_number,
                            self._E, self._H, self._inv_eps, self._inv_mu, pitch=self._px)
        if self.compensated:
            plan.set_compensation(self._E_lo, self._H_lo)
        pal = self._material_palette() if self._palette_wanted() else None
        if pal is not None:
            plan.set_material_palette(*pal)
        for comp in self._components():
            comp._add_to_plan(plan)
        self._plan = plan
        self._plan_key = self._current_plan_key()
        self._wave_key = None
        self._ensure_waveforms(self.time_steps_passed + 1)
        plan.finalize()
        return plan

    def _ensure_waveforms(self, q_end):
        """Device table of per-step source values for q < q_end, computed on the
        host with Python math (the values the reference computes each step)."""
        if not self.sources:
            self._plan.set_waveforms(None)
            return
        cap = 0 if self._wave is None else self._wave.shape[1]
        key = tuple(id(s) for s in self.sources)
        if self._wave is not None and key == self._wave_key and cap >= q_end:
            return
        new_cap = max(int(q_end), 2 * cap if key == self._wave_key else 0, 256)
        table = np.empty((len(self.sources), new_cap), dtype=np.float64)
        for r, src in enumerate(self.sources):
            table[r] = [src._waveform(q) for q in range(new_cap)]
        self._wave = torch.from_numpy(table).to(self._E.device)
        self._wave_key = key
        self._plan.set_waveforms(self._wave)

    def _prepare(self, nsteps, e_steps, h_steps):
        plan = self._ensure_plan()
        grown = False
        for det in self.detectors:
            grown |= det._reserve(max(e_steps, h_steps))
        for src in self.sources:
            if hasattr(src, "_reserve"):  # sources that record (soft sources)
                grown |= src._reserve(e_steps)
        if grown:
            plan = self._ensure_plan()
        self._ensure_waveforms(self.time_steps_passed + nsteps)
        return plan

    @staticmethod
    def _stream():
        return torch.cuda.current_stream().cuda_stream

    # ------------------------------------------------------------- the loop
    def run(self, total_time: Number, progress_bar: bool = True):
        """run an FDTD simulation (grid.py:250-265).

        The whole run is queued on the current HIP stream without host
        synchronisation; with progress_bar the queueing is split into chunks
        so the bar advances as work is submitted."""
        if isinstance(total_time, float):
            total_time /= self.time_step
        n = int(total_time)
        if n <= 0:
            return
        if type(self).step is not Grid.step:
            # a subclass redefined step(): honour it step by step
            it = range(n)
            if progress_bar:
                from tqdm import tqdm
                it = tqdm(it)
            for _ in it:
                self.step()
            return
        chunks = [n]
        bar = None
        if progress_bar:
            from tqdm import tqdm
            bar = tqdm(total=n)
            per = max(1, n // 100)
            chunks = [per] * (n // per) + ([n % per] if n % per else [])
        for m in chunks:
            plan = self._prepare(m, m, m)
            plan.run(self.time_steps_passed, m, self._stream())
            self._count_halves(m, m)
            self.time_steps_passed += m
            if bar is not None:
                bar.update(m)
        if bar is not None:
            bar.close()

    def has_nonfinite(self) -> bool:
        """True when a value of E or H is a NaN or an infinity: one check on
        the device and one 4-byte read-back (no counterpart in the reference,
        whose only guard against divergence is the Courant limit at
        construction, grid.py:123-127 -- a courant_number raised afterwards,
        or a NaN written into the fields, goes unnoticed there)."""
        plan = self._ensure_plan()
        flag = torch.zeros(1, dtype=torch.int32, device=self._E.device)
        plan.check_finite(flag, self._stream())
        return bool(flag.item())

    def step(self):
        """E half step, H half step, then advance the time index (grid.py:267-273)."""
        self.update_E()
        self.update_H()
        self.time_steps_passed += 1

    def update_E(self):
        """Everything the reference's update_E does, in one fused launch
        sequence (grid.py:275-299)."""
        plan = self._prepare(1, 1, 0)
        plan.half_step_E(self.time_steps_passed, self._stream())
        self._count_halves(1, 0)

    def update_H(self):
        """Everything the reference's update_H does (grid.py:301-325)."""
        plan = self._prepare(1, 0, 1)
        plan.half_step_H(self.time_steps_passed, self._stream())
        self._count_halves(0, 1)

    def _count_halves(self, e, h):
        """Samples the device recorded during e E and h H half steps."""
        for det in self.detectors:
            det._count_E += e
            det._count_H += h
        for src in self.sources:
            if hasattr(src, "_count"):
                src._count += e

    def reset(self):
        """zero E, H and the time index; PML state and detector traces are kept
        (grid.py:327-331)"""
        self.H *= 0.0
        self.E *= 0.0
        if self.compensated:
            self._E_lo.zero_()
            self._H_lo.zero_()
        self.time_steps_passed *= 0

    def add_source(self, name, source):
        """the reference's broken helper (grid.py:333-336) is not carried over"""
        raise NotImplementedError("register components with grid[x, y, z] = component")

    add_boundary = add_detector = add_object = add_source

    def __setitem__(self, key, attr):
        if not isinstance(key, tuple):
            x, y, z = key, slice(None), slice(None)
        elif len(key) == 1:
            x, y, z = key[0], slice(None), slice(None)
        elif len(key) == 2:
            x, y, z = key[0], key[1], slice(None)
        elif len(key) == 3:
            x, y, z = key
        else:
            raise KeyError("maximum number of indices for the grid is 3")
        attr._register_grid(
            grid=self,
            x=self._handle_single_key(x),
            y=self._handle_single_key(y),
            z=self._handle_single_key(z),
        )

    def __repr__(self):
        return (
            f"{self.__class__.__name__}(shape=({self.Nx},{self.Ny},{self.Nz}), "
            f"grid_spacing={self.grid_spacing:.2e}, courant_number={self.courant_number:.2f})"
        )

    def __str__(self):
        s = repr(self) + "\n"
        for title, comps in (("sources", self.sources), ("detectors", self.detectors),
                             ("boundaries", self.boundaries), ("objects", self.objects)):
            if comps:
                s = s + f"\n{title}:\n"
                for c in comps:
                    s += str(c)
        return s

    # --------------------------------------------------------------- output
    def save_simulation(self, sim_name=None):
        """Create fdtd_output/fdtd_output_<timestamp> (grid.py:407-439)."""
        makedirs("fdtd_output", exist_ok=True)
        now = datetime.now()
        full_sim_name = "-".join(str(v) for v in (now.year, now.month, now.day, now.hour, now.minute, now.second))
        if sim_name is not None:
            full_sim_name = full_sim_name + " (" + sim_name + ")"
        self.folder = os.path.abspath(path.join("fdtd_output", "fdtd_output_" + full_sim_name))
        self.full_sim_name = full_sim_name
        makedirs(self.folder, exist_ok=True)
        return self.folder

    def save_data(self):
        """Detector traces to <folder>/detector_readings.npz with keys
        "<name> (E)" / "<name> (H)" (grid.py:490-514); one device-to-host
        copy per trace."""
        if self.folder is None:
            raise Exception(
                "Save location not initialized. Please read about 'fdtd.Grid.saveSimulation()' or try running "
                "'grid.saveSimulation()'.")
        dic = {}
        for detector in self.detectors:
            dic[str(detector.name) + " (E)"] = detector.trace("E").cpu().numpy()
            dic[str(detector.name) + " (H)"] = detector.trace("H").cpu().numpy()
        np.savez(path.join(self.folder, "detector_readings"), **dic)

    def generate_video(self, delete_frames=False):
        raise NotImplementedError("video generation (matplotlib + ffmpeg) is not part of the HIP backend")

    def grid_energy(self, x=None, y=None, z=None):
        """The optical energy plane that the reference's visualize() draws
        (visualization.py:86-146): sum over the components of E**2 + H**2,
        at one x, y or z index, computed on the device for that plane only.
        The y plane is returned transposed, (z, x), as it is drawn there; the
        argument checks raise the same errors."""
        checks = (("x", x, (y, z), "an `x`", "a `y` or a `z`"), ("y", y, (z, x), "a `y`", "a `z` or a `x`"),
                  ("z", z, (x, y), "a `z`", "a `x` or a `y`"))
        for name, v, others, this, rest in checks:
            if v is None:
                continue
            if not isinstance(v, int):
                raise ValueError(f"the `{name}`-location supplied should be a single integer")
            if any(o is not None for o in others):
                raise ValueError(f"if {this}-location is supplied, one should not supply {rest}-location!")
            break
        else:
            raise ValueError("at least one projection plane (x, y or z) should be supplied to visualize the grid!")
        E, H = self.E, self.H
        if x is not None:
            assert self.Ny > 1 and self.Nz > 1
            e, h = E[x], H[x]
        elif y is not None:
            assert self.Nx > 1 and self.Nz > 1
            e, h = E[:, y], H[:, y]
        else:
            assert self.Nx > 1 and self.Ny > 1
            e, h = E[:, :, z], H[:, :, z]
        sq = e ** 2 + h ** 2
        energy = (sq[..., 0] + sq[..., 1]) + sq[..., 2]  # numpy's order for a 3-term sum
        return energy.T if y is not None else energy

    def visualize(self, *args, **kwargs):
        raise NotImplementedError("plotting is not part of the HIP backend: draw grid.grid_energy(x=..|y=..|z=..) "
                                  "(the energy plane visualize() shows, computed on the device) with matplotlib")
