"""fdtd_amd -- an MI355X-native FDTD engine with the Python API of flaport/fdtd.

    import fdtd_amd as fdtd
    fdtd.set_backend("hip.float64")          # or "hip.float32"
    grid = fdtd.Grid((256, 256, 256), grid_spacing=50e-9, courant_number=3 ** -0.5)
    grid[0:16, :, :] = fdtd.PML()
    grid[128, 128, 128] = fdtd.PointSource(period=1550e-9 / 3e8, pulse=True)
    grid.run(1000, progress_bar=False)

Each half step runs as fused HIP kernels (libfdtd_hip.so, C ABI in
include/fdtd_hip.h); the package fails loudly if the library is missing.
"""

__version__ = "0.3.6"

from .grid import Grid
from .sources import PointSource, LineSource, PlaneSource
from .detectors import LineDetector, BlockDetector, CurrentDetector
from .objects import Object, AbsorbingObject, AnisotropicObject
from .boundaries import PeriodicBoundary, PML, DomainBorderPML
from .backend import backend
from .backend import set_backend
from .backend import backend_names_available
from .fourier import FrequencyRoutines
