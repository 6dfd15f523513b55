"""Backend selection for fdtd_amd (mirrors fdtd/backend.py:367-439).

The package provides the HIP backends only:

    - ``hip``          float64 fields
    - ``hip.float64``
    - ``hip.float32``  float32 fields (all arithmetic in float32)
    - ``hip.float32c`` float32 fields with an 8-bit compensation term per
                       value (the 8 mantissa bits the float drops: 32
                       significant bits in 5 bytes), float64 arithmetic,
                       materials and CPML state

Arrays are torch tensors on the HIP device (torch names it "cuda" on ROCm);
the time stepping runs in hand-written HIP kernels (libfdtd_hip.so). The
numpy and torch backends of the reference are not re-implemented here: the
reference package serves them.

Like the reference, ``backend`` is a module-level singleton whose class and
``float`` attribute ``set_backend`` swaps, so ``from fdtd_amd.backend import
backend as bd`` stays valid across switches. Unlike the reference, importing
this module changes no torch global state (default dtype, autograd).
"""

from functools import wraps  # noqa: F401

import numpy
import torch

backend_names = [
    dict(backends="hip.float64"),
    dict(backends="hip.float32"),
    dict(backends="hip.float32c"),
]

_DTYPES = {"float64": torch.float64, "float32": torch.float32, "float32c": torch.float32}
_NP = {torch.float64: numpy.float64, torch.float32: numpy.float32}

# Registration-only mode for hosts without a HIP device (the unit tests of the
# host logic): arrays are kept in host memory and any time step raises.
_HOST_ONLY = False


def host_only(flag=True):
    """Keep grid arrays in host memory so that registration logic can be
    exercised without a GPU; stepping then raises RuntimeError."""
    global _HOST_ONLY
    _HOST_ONLY = bool(flag)


def backend_names_available():
    """Backend names set_backend accepts."""
    return ["hip"] + ["hip." + d for d in _DTYPES]


def device_available():
    return torch.cuda.is_available()


class Backend:
    """Backend Base Class"""

    pi = numpy.pi

    def __repr__(self):
        return self.__class__.__name__


class HipBackend(Backend):
    """Torch-ROCm tensors on the HIP device; stepping through libfdtd_hip."""

    int = torch.int64
    float = torch.float64
    compensated = False

    @property
    def complex(self):
        """complex type of the fields after promotion (backend.py:216-220)"""
        return torch.complex64 if self.float == torch.float32 else torch.complex128

    @property
    def material_float(self):
        """dtype of materials, object arrays and CPML state"""
        return torch.float64 if self.compensated else self.float

    @property
    def device(self):
        if _HOST_ONLY:
            return torch.device("cpu")
        if not torch.cuda.is_available():
            raise RuntimeError(
                "The HIP backend needs a HIP device and none is visible "
                "(torch.cuda.is_available() is False)")
        return torch.device("cuda", torch.cuda.current_device())

    @property
    def numpy_float(self):
        """numpy dtype of host-side material / profile computations"""
        return _NP[self.material_float]

    # -- constructors
    def _dtype(self, dtype):
        if dtype is None or dtype is float:
            return self.float
        if isinstance(dtype, torch.dtype):
            return dtype
        return torch.from_numpy(numpy.zeros(0, dtype=dtype)).dtype

    def array(self, arr, dtype=None, **kwargs):
        """create an array from an array-like sequence (lists of arrays are stacked)"""
        dt = self._dtype(dtype)
        if torch.is_tensor(arr):
            return arr.detach().clone().to(device=self.device, dtype=dt)
        if isinstance(arr, (list, tuple)) and arr and any(torch.is_tensor(a) for a in arr):
            return torch.stack([self.array(a, dtype=dt) for a in arr])
        return torch.as_tensor(numpy.asarray(arr), device=self.device).to(dt)

    def asarray(self, arr, dtype=None):
        if torch.is_tensor(arr):
            dt = arr.dtype if dtype is None else self._dtype(dtype)
            return arr.to(device=self.device, dtype=dt)
        a = numpy.asarray(arr)
        dt = self._dtype(dtype) if (dtype is not None or a.dtype.kind == "f") else None
        t = torch.as_tensor(a, device=self.device)
        return t.to(dt) if dt is not None else t

    def ones(self, shape, dtype=None, **kwargs):
        return torch.ones(shape, dtype=self._dtype(dtype), device=self.device)

    def zeros(self, shape, dtype=None, **kwargs):
        return torch.zeros(shape, dtype=self._dtype(dtype), device=self.device)

    zeros_like = staticmethod(torch.zeros_like)

    def linspace(self, start, stop, num=50, endpoint=True):
        return torch.as_tensor(numpy.linspace(start, stop, num, endpoint=endpoint), device=self.device).to(self.float)

    def arange(self, start, stop=None, step=1, **kwargs):
        a = numpy.arange(start, stop, step) if stop is not None else numpy.arange(start)
        t = torch.as_tensor(a, device=self.device)
        return t.to(self.float) if a.dtype.kind == "f" else t

    # -- elementwise and shape ops
    exp = staticmethod(torch.exp)
    sin = staticmethod(torch.sin)
    cos = staticmethod(torch.cos)
    sum = staticmethod(torch.sum)
    max = staticmethod(torch.max)
    stack = staticmethod(torch.stack)
    squeeze = staticmethod(torch.squeeze)
    broadcast_arrays = staticmethod(torch.broadcast_tensors)
    broadcast_to = staticmethod(torch.broadcast_to)
    reshape = staticmethod(torch.reshape)
    divide = staticmethod(torch.div)
    pad = staticmethod(torch.nn.functional.pad)
    fftfreq = staticmethod(numpy.fft.fftfreq)

    @staticmethod
    def fft(arr, *args, **kwargs):
        return torch.fft.fft(arr, *args, **kwargs)

    @staticmethod
    def bmm(arr1, arr2):
        """batch matrix multiply two arrays"""
        return torch.bmm(arr1, arr2.contiguous())

    @staticmethod
    def transpose(arr, axes=None):
        if axes is None:
            axes = tuple(range(len(arr.shape) - 1, -1, -1))
        return arr.permute(*axes)

    @staticmethod
    def is_array(arr):
        return isinstance(arr, numpy.ndarray) or torch.is_tensor(arr)

    @staticmethod
    def is_complex(x):
        return isinstance(x, complex) or (torch.is_tensor(x) and torch.is_complex(x)) or (
            isinstance(x, numpy.ndarray) and numpy.iscomplexobj(x))

    def numpy(self, arr):
        """convert to a host numpy array (lists of arrays are stacked)"""
        if torch.is_tensor(arr):
            return arr.detach().cpu().numpy()
        if isinstance(arr, (list, tuple)) and arr and any(torch.is_tensor(a) for a in arr):
            return numpy.stack([self.numpy(a) for a in arr])
        return numpy.asarray(arr)


backend = HipBackend()


def set_backend(name: str):
    """Select ``hip`` / ``hip.float64`` / ``hip.float32`` / ``hip.float32c``
    (fdtd/backend.py:367-439)."""
    if name.count(".") == 0:
        base, dtype = name, "float64"
    elif name.count(".") == 1:
        base, dtype = name.split(".")
    else:
        raise ValueError(f"Unknown backend '{name}'")
    if base != "hip":
        raise ValueError(
            f"Unknown backend '{name}'. fdtd_amd provides 'hip', 'hip.float64' and 'hip.float32'; "
            "the numpy/torch backends belong to the reference fdtd package")
    if dtype not in _DTYPES:
        raise ValueError(f"Unknown dtype '{dtype}' for the HIP backend (float64, float32 or float32c)")
    if not _HOST_ONLY and not torch.cuda.is_available():
        raise RuntimeError(
            "HIP backend is not available.\nDo you have an AMD GPU visible to PyTorch-ROCm?")
    backend.__class__ = HipBackend
    backend.float = _DTYPES[dtype]
    backend.compensated = dtype == "float32c"
