"""Objects on the HIP backend (API of fdtd/objects.py).

An object is a box with its own E update, applied inside the fused E kernel
after the grid update and before the boundaries, in registration order:

 - Object:            E += (S * inv) * curl                      (objects.py:118-129)
 - AbsorbingObject:   E *= (1-f)/(1+f); E += (S*inv*curl)/(1+f)  (objects.py:207-221)
 - AnisotropicObject: E += S * (inv @ curl), inv 3x3 per cell    (objects.py:254-269)

Registration does what the reference does (objects.py:37-92): the object's
inverse permittivity takes the grid's value on its last plane per axis from
the grid's LAST plane, and the grid's inverse permittivity is zeroed inside
the box (which also silences CPML corrections there). The object's arrays are
device tensors that the kernels read every step, so in-place edits take
effect; rebinding ``obj.inverse_permittivity`` re-plans on the next step.
"""

import numpy as np
import torch

from . import _native
from . import constants as const
from .backend import backend as bd
from .typing_ import ListOrSlice, Tensorlike


def _is_complex(a):
    if torch.is_tensor(a):
        return a.is_complex()
    return np.iscomplexobj(np.asarray(a))


def _host(a, dtype=None):
    if torch.is_tensor(a):
        a = a.detach().cpu().numpy()
    a = np.asarray(a)
    return a.astype(dtype, copy=False) if dtype is not None else a


class Object:
    """An object to place in the grid"""

    _kind = _native.OBJ_PLAIN

    def __init__(self, permittivity: Tensorlike, name: str = None):
        self.grid = None
        self.name = name
        self.permittivity = permittivity

    def _register_grid(self, grid, x: ListOrSlice, y: ListOrSlice, z: ListOrSlice):
        self.grid = grid
        self.grid.objects.append(self)
        complex_perm = _is_complex(self.permittivity)
        # the reference keeps a real object's arrays real on complex fields
        self._real_arrays = not complex_perm
        if complex_perm:
            # objects.py:51-52: E, H and the CPML state registered so far
            # become complex before the object computes its inverse
            self.grid.promote_dtypes_to_complex()
        if self.name is not None:
            if not hasattr(grid, self.name):
                setattr(grid, self.name, self)
            else:
                raise ValueError(f"The grid already has an attribute with name {self.name}")
        self.x = self._handle_slice(x, max_index=self.grid.Nx)
        self.y = self._handle_slice(y, max_index=self.grid.Ny)
        self.z = self._handle_slice(z, max_index=self.grid.Nz)
        self.Nx = abs(self.x.stop - self.x.start)
        self.Ny = abs(self.y.stop - self.y.start)
        self.Nz = abs(self.z.stop - self.z.start)

        npdt = bd.numpy_float
        if complex_perm:  # numpy's complex division, in the fields' complex dtype
            npdt = np.complex128 if npdt == np.float64 else np.complex64
        perm = _host(self.permittivity, npdt)
        if perm.ndim == 3:
            perm = perm[..., None]
        inv = np.ones((self.Nx, self.Ny, self.Nz, 3), dtype=npdt) / perm
        # border trick (objects.py:79-90): values from the grid's last plane
        g = grid.inverse_permittivity
        if self.Nx > 1:
            inv[-1, :, :, 0] = g[-1, self.y, self.z, 0].cpu().numpy()
        if self.Ny > 1:
            inv[:, -1, :, 1] = g[self.x, -1, self.z, 1].cpu().numpy()
        if self.Nz > 1:
            inv[:, :, -1, 2] = g[self.x, self.y, -1, 2].cpu().numpy()
        g[self.x, self.y, self.z] = 0
        self._inv_host = inv
        self.inverse_permittivity = torch.from_numpy(np.ascontiguousarray(inv)).to(grid._E.device)

    def _handle_slice(self, s: ListOrSlice, max_index: int = None) -> slice:
        if isinstance(s, list):
            if len(s) == 1:
                return slice(s[0], s[0] + 1, None)
            raise IndexError("One can only use slices or single indices to index the grid for an Object")
        if isinstance(s, slice):
            start, stop, step = s.start, s.stop, s.step
            if step is not None and step != 1:
                raise IndexError("Can only use slices with unit step to index the grid for an Object")
            if start is None:
                start = 0
            if start < 0:
                start = max_index + start
            if stop is None:
                stop = max_index
            if stop < 0:
                stop = max_index + stop
            return slice(start, stop, None)
        raise ValueError("Invalid grid indexing used for object")

    # plan interface
    def _device_array(self, attr, shape):
        """Bind `attr` as a contiguous device tensor of the field dtype."""
        v = getattr(self, attr)
        g = self.grid
        ok = (torch.is_tensor(v) and v.device == g._E.device and v.dtype == g.obj_dtype and v.is_contiguous()
              and tuple(v.shape) == shape)
        if not ok:
            t = torch.as_tensor(_host(v) if not torch.is_tensor(v) else v).to(device=g._E.device, dtype=g.obj_dtype)
            t = t.expand(shape).contiguous()
            setattr(self, attr, t)
            v = t
        return v

    def _arrays(self):
        return (self._device_array("inverse_permittivity", (self.Nx, self.Ny, self.Nz, 3)), None)

    def _plan_key(self):
        key = [type(self).__name__]
        for name in ("inverse_permittivity", "absorption_factor"):
            v = getattr(self, name, None)
            key.append(id(v))
        return tuple(key)

    def _add_to_plan(self, plan):
        inv, absorb = self._arrays()
        box = ((self.x.start, self.x.stop), (self.y.start, self.y.stop), (self.z.start, self.z.stop))
        kind = self._kind | (_native.OBJ_REAL_ARRAYS if self._real_arrays else 0)
        plan.add_object(kind, box, inv, absorb)

    def update_E(self, curl_H):
        raise NotImplementedError("object updates run inside the fused HIP step")

    def update_H(self, curl_E):
        """objects leave H unchanged"""

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


class AbsorbingObject(Object):
    """An absorbing object takes conductivity into account"""

    _kind = _native.OBJ_ABSORBING

    def __init__(self, permittivity: Tensorlike, conductivity: Tensorlike, name: str = None):
        super().__init__(permittivity, name)
        self.conductivity = conductivity

    def _register_grid(self, grid, x=None, y=None, z=None):
        super()._register_grid(grid=grid, x=x, y=y, z=z)
        npdt = bd.numpy_float
        cond = _host(self.conductivity, npdt)
        while cond.ndim < 4:
            cond = cond[..., None]
        cond = np.broadcast_to(cond, self._inv_host.shape)
        self.conductivity = torch.from_numpy(np.ascontiguousarray(cond)).to(grid._E.device)
        # objects.py:198-205, evaluated left to right in the dtype of the
        # object's inverse permittivity (complex for a complex permittivity)
        f = 0.5 * grid.courant_number * self._inv_host * cond * grid.grid_spacing * const.eta0
        f = f.astype(self._inv_host.dtype)
        self._real_arrays = not np.iscomplexobj(f)
        self.absorption_factor = torch.from_numpy(np.ascontiguousarray(f)).to(grid._E.device)

    def _arrays(self):
        shape = (self.Nx, self.Ny, self.Nz, 3)
        return (self._device_array("inverse_permittivity", shape), self._device_array("absorption_factor", shape))


class AnisotropicObject(Object):
    """An object with anisotropic permittivity tensor"""

    _kind = _native.OBJ_ANISOTROPIC

    def _register_grid(self, grid, x=None, y=None, z=None):
        super()._register_grid(grid=grid, x=x, y=y, z=z)
        n = self.Nx * self.Ny * self.Nz
        eye = np.zeros((n, 3, 3), dtype=bd.numpy_float)
        eye[:, range(3), range(3)] = 1.0
        full = (self._inv_host.reshape(-1, 1, 3) * eye).reshape(self.Nx, self.Ny, self.Nz, 3, 3)
        self.inverse_permittivity = torch.from_numpy(np.ascontiguousarray(full)).to(grid._E.device)

    def _arrays(self):
        return (self._device_array("inverse_permittivity", (self.Nx, self.Ny, self.Nz, 3, 3)), None)
