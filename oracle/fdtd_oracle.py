"""CPU oracle for the FDTD time-step path -- TEST INFRASTRUCTURE ONLY.

This module restates, in numpy float64, the algorithm of the reference
package (flaport/fdtd 0.3.6) for the hot path: grid set-up, the Yee leap-frog
update, CPML faces, periodic boundaries, the three object kinds, the point /
line / plane sources and the line / block detectors. It is imported only by
tests/, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
bench.py, always as the checker, never as the product: the product path
(``fdtd_amd`` on a ``hip.*`` backend) runs HIP kernels and fails loudly when
they are missing.

Structure is the oracle's own (one ``Sim`` object, components as small
records, one generic PML face instead of six subclasses), but every array
statement performs the same IEEE operations, in the same order, as the
reference's statement it cites, so that fp64 results can be compared
bit-for-bit. Whole-array temporaries are kept on purpose: the CPU baseline
timed from this module then costs what the reference's numpy backend costs.

Pinning (checked by tests/test_oracle.py):
  * curl_E / curl_H            - the reference's known-answer vectors
                                 (tests/test_grid.py:47-164 of the reference)
  * every Sim feature          - golden fixtures written by the reference
                                 itself (tests/golden/*.npz, bit-exact)
  * closed forms               - courant / time step, 1-D magic time step,
                                 uniform absorbing decay, PML sigma profile
"""

import math

import numpy as np

# fdtd/constants.py:5-23
SPEED_OF_LIGHT = 299792458.0
MU0 = 4e-7 * math.pi
ETA0 = MU0 * SPEED_OF_LIGHT


# ---------------------------------------------------------------- curls
def curl_E(E):
    """Forward-difference curl of an E-type field (fdtd/grid.py:29-51).
    A derivative whose forward neighbour is outside the grid is absent."""
    out = np.zeros(E.shape, dtype=E.dtype)
    # (component written, component differenced, axis of the difference, sign)
    for comp, src, axis, sign in ((0, 2, 1, +1), (0, 1, 2, -1),
                                  (1, 0, 2, +1), (1, 2, 0, -1),
                                  (2, 1, 0, +1), (2, 0, 1, -1)):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(None, -1)
        hi[axis] = slice(1, None)
        d = E[tuple(hi) + (src,)] - E[tuple(lo) + (src,)]
        if sign > 0:
            out[tuple(lo) + (comp,)] += d
        else:
            out[tuple(lo) + (comp,)] -= d
    return out


def curl_H(H):
    """Backward-difference curl of an H-type field (fdtd/grid.py:54-76).
    A derivative whose backward neighbour is outside the grid is absent."""
    out = np.zeros(H.shape, dtype=H.dtype)
    for comp, src, axis, sign in ((0, 2, 1, +1), (0, 1, 2, -1),
                                  (1, 0, 2, +1), (1, 2, 0, -1),
                                  (2, 1, 0, +1), (2, 0, 1, -1)):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(None, -1)
        hi[axis] = slice(1, None)
        d = H[tuple(hi) + (src,)] - H[tuple(lo) + (src,)]
        if sign > 0:
            out[tuple(hi) + (comp,)] += d
        else:
            out[tuple(hi) + (comp,)] -= d
    return out


# ------------------------------------------------------------- waveforms
def hanning(f, t, n):
    """Hanning-windowed sine, fdtd/waveforms.py:8-9 (no 2*pi factor)."""
    return (1 / 2) * (1 - math.cos(f * t / n)) * (math.sin(f * t))


# ------------------------------------------------------------ components
class _Pml:
    """One CPML face (fdtd/boundaries.py:225-487, faces :490-625).

    Only the face's normal axis and side matter: the tangential extents of the
    registration key are ignored, so the slab always spans full planes."""

    def __init__(self, sim, axis, high, thickness, a=1e-8):
        self.sim, self.axis, self.high, self.t = sim, axis, high, thickness
        self.k, self.a = 1.0, a
        n = [sim.Nx, sim.Ny, sim.Nz]
        n[axis] = thickness
        self.shape = tuple(n)
        sl = [slice(None)] * 3
        sl[axis] = slice(-thickness, None) if high else slice(None, thickness)
        self.loc = tuple(sl)
        t = thickness
        # conductivity profiles along the normal axis, boundaries.py:291-293,502-625
        prof = lambda v: 40 * v ** 3 / (t + 1) ** 4
        if high:
            sig_e = prof(np.arange(0.5, t + 0.5, 1.0))
            sig_h = prof(np.arange(1.0, t, 1.0))
        else:
            sig_e = prof(np.arange(t - 0.5, -0.5, -1.0))
            sig_h = prof(np.arange(t - 1.0, 0, -1.0))
        bshape = [1, 1, 1]
        bshape[axis] = -1
        self.sigmaE = np.zeros(self.shape + (3,))
        self.sigmaE[..., axis] = sig_e.reshape(bshape)
        self.sigmaH = np.zeros(self.shape + (3,))
        hsl = [slice(None)] * 3
        hsl[axis] = slice(None, -1)
        self.sigmaH[tuple(hsl) + (axis,)] = sig_h.reshape(bshape)
        sc = sim.courant_number
        # CPML recursion coefficients, boundaries.py:395-407
        self.bE = np.exp(-(self.sigmaE / self.k + self.a) * sc)
        self.cE = (self.bE - 1.0) * self.sigmaE / (self.sigmaE * self.k + self.a * self.k ** 2)
        self.bH = np.exp(-(self.sigmaH / self.k + self.a) * sc)
        self.cH = (self.bH - 1.0) * self.sigmaH / (self.sigmaH * self.k + self.a * self.k ** 2)
        self.psi_E = [np.zeros(self.shape + (3,)) for _ in range(3)]  # psi_Ex, psi_Ey, psi_Ez
        self.psi_H = [np.zeros(self.shape + (3,)) for _ in range(3)]
        self.phi_E = np.zeros(self.shape + (3,))
        self.phi_H = np.zeros(self.shape + (3,))

    @staticmethod
    def _ax(axis, s):
        out = [slice(None)] * 3
        out[axis] = s
        return tuple(out)

    def _update_psi(self, psi, field, c, forward):
        """psi_<i>[..., j] += Delta_j field_k * c[..., j] with k = 3-i-j,
        boundaries.py:439-455 (E side, stored at slab index >= 1) and
        :467-483 (H side, stored at slab index <= n-2)."""
        for p in psi:
            p *= (self.bE if not forward else self.bH)
        slab = field[self.loc]
        for i in range(3):
            for j in ((i + 1) % 3, (i + 2) % 3):
                k = 3 - i - j
                f = slab[..., k]
                d = f[self._ax(j, slice(1, None))] - f[self._ax(j, slice(None, -1))]
                at = self._ax(j, slice(None, -1) if forward else slice(1, None))
                psi[i][at + (j,)] += d * c[at + (j,)]

    def update_phi_E(self):
        self._update_psi(self.psi_E, self.sim.H, self.cE, forward=False)
        for i in range(3):
            self.phi_E[..., i] = self.psi_E[i][..., (i + 1) % 3] - self.psi_E[i][..., (i + 2) % 3]

    def update_phi_H(self):
        self._update_psi(self.psi_H, self.sim.E, self.cH, forward=True)
        for i in range(3):
            self.phi_H[..., i] = self.psi_H[i][..., (i + 1) % 3] - self.psi_H[i][..., (i + 2) % 3]

    def promote_dtypes_to_complex(self):  # boundaries.py:106-116
        self.phi_E = self.phi_E.astype(np.complex128)
        self.phi_H = self.phi_H.astype(np.complex128)
        self.psi_E = [p.astype(np.complex128) for p in self.psi_E]
        self.psi_H = [p.astype(np.complex128) for p in self.psi_H]

    def update_E(self):
        s = self.sim
        s.E[self.loc] += s.courant_number * s.inverse_permittivity[self.loc] * self.phi_E

    def update_H(self):
        s = self.sim
        s.H[self.loc] -= s.courant_number * s.inverse_permeability[self.loc] * self.phi_H


class _Periodic:
    """Plane copy E[0] <- E[-1], H[-1] <- H[0] along one axis
    (fdtd/boundaries.py:184-219)."""

    def __init__(self, sim, axis):
        self.sim, self.axis = sim, axis

    def _plane(self, i):
        sl = [slice(None)] * 3
        sl[self.axis] = i
        return tuple(sl)

    def update_phi_E(self):
        pass

    def update_phi_H(self):
        pass

    def update_E(self):
        self.sim.E[self._plane(0)] = self.sim.E[self._plane(-1)]

    def update_H(self):
        self.sim.H[self._plane(-1)] = self.sim.H[self._plane(0)]


class _Object:
    """Box with its own E update (fdtd/objects.py:24-277); kind is
    'plain', 'absorbing' or 'anisotropic'."""

    def __init__(self, sim, kind, x, y, z, permittivity, conductivity=None):
        self.sim, self.kind = sim, kind
        self.x, self.y, self.z = x, y, z
        self.Nx, self.Ny, self.Nz = x.stop - x.start, y.stop - y.start, z.stop - z.start
        perm = np.array(permittivity)
        if np.iscomplexobj(perm):
            # objects.py:51-52: a complex permittivity promotes the grid fields
            # and the boundaries registered so far before anything else
            perm = perm.astype(np.complex128)
            sim.promote_dtypes_to_complex()
        else:
            perm = perm.astype(np.float64)
        if perm.ndim == 3:
            perm = perm[..., None]
        inv = np.ones((self.Nx, self.Ny, self.Nz, 3), dtype=perm.dtype) / perm
        g = sim.inverse_permittivity
        # "border trick": the object's last plane per axis takes the grid's
        # value from the grid's LAST plane (objects.py:79-90)
        if self.Nx > 1:
            inv[-1, :, :, 0] = g[-1, y, z, 0]
        if self.Ny > 1:
            inv[:, -1, :, 1] = g[x, -1, z, 1]
        if self.Nz > 1:
            inv[:, :, -1, 2] = g[x, y, -1, 2]
        g[x, y, z] = 0  # objects.py:92
        self.inverse_permittivity = inv
        if kind == "absorbing":
            cond = np.asarray(conductivity, dtype=np.float64)
            while cond.ndim < inv.ndim:
                cond = cond[..., None]
            cond = np.broadcast_to(cond, inv.shape)
            self.absorption_factor = (0.5 * sim.courant_number * inv * cond
                                      * sim.grid_spacing * ETA0)  # objects.py:198-205
        if kind == "anisotropic":
            n = self.Nx * self.Ny * self.Nz
            eye = np.zeros((n, 3, 3))
            eye[:, range(3), range(3)] = 1.0
            self.inverse_permittivity = (inv.reshape(-1, 1, 3) * eye).reshape(
                self.Nx, self.Ny, self.Nz, 3, 3)  # objects.py:247-252

    def update_E(self, curl):
        s = self.sim
        loc = (self.x, self.y, self.z)
        sc = s.courant_number
        if self.kind == "plain":  # objects.py:118-129
            s.E[loc] = s.E[loc] + sc * self.inverse_permittivity * curl[loc]
        elif self.kind == "absorbing":  # objects.py:207-221
            f = self.absorption_factor
            s.E[loc] *= (1 - f) / (1 + f)
            s.E[loc] += sc * self.inverse_permittivity * curl[loc] / (1 + f)
        else:  # objects.py:254-269, a collocated 3x3 mat-vec
            m = self.inverse_permittivity.reshape(-1, 3, 3)
            v = curl[loc].reshape(-1, 3, 1)
            s.E[loc] += (sc * np.einsum("ijk,ikl->ijl", m, v)).reshape(self.Nx, self.Ny, self.Nz, 3)


class _Source:
    """Point / line / plane source (fdtd/sources.py:25-501)."""

    def __init__(self, sim, kind, period, amplitude, phase_shift, pulse=False,
                 cycle=5, hanning_dt=10.0, polarization="z"):
        self.sim, self.kind = sim, kind
        self.amplitude, self.phase_shift = amplitude, phase_shift
        self.pulse, self.cycle = pulse, cycle
        # default window step from the constructor-time frequency, sources.py:58-59
        self.hanning_dt = hanning_dt if hanning_dt is not None else 0.5 / (1.0 / period)
        self.period = sim.steps_of(period)  # sources.py:90,194,374
        self.frequency = 1.0 / self.period
        self.polarization = polarization

    def _sine(self, q):
        return math.sin(2 * math.pi * q / self.period + self.phase_shift)

    def _pulse_value(self, q):
        """hanning window value or 0 after t1 (sources.py:97-105)."""
        t1 = int(2 * math.pi / (self.frequency * self.hanning_dt / self.cycle))
        if q < t1:
            return hanning(self.frequency, q * self.hanning_dt, self.cycle)
        return None

    def update_E(self):
        s = self.sim
        q = s.time_steps_passed
        if self.kind == "point":  # sources.py:93-109
            if self.pulse:
                h = self._pulse_value(q)
                src = self.amplitude * h if h is not None else 0
            else:
                src = self.amplitude * self._sine(q)
            s.E[self.x, self.y, self.z, 2] += src
        elif self.kind == "line":  # sources.py:278-297
            if self.pulse:
                h = self._pulse_value(q)
                vect = self.profile * h if h is not None else self.profile * 0
            else:
                vect = self.profile * self._sine(q)
            for x, y, z, v in zip(self.x, self.y, self.z, vect):
                s.E[x, y, z, 2] += v
        else:  # plane, a hard source: sources.py:476-480
            s.E[self.x, self.y, self.z, self.Epol] = self.profile * self._sine(q)

    def update_H(self):
        if self.kind == "plane":  # sources.py:482-486
            s = self.sim
            s.H[self.x, self.y, self.z, self.Hpol] = self.profile * self._sine(s.time_steps_passed)


class _Detector:
    """Line / block detector (fdtd/detectors.py:20-280): one (copied) sample
    of E after every E half step and of H after every H half step."""

    def __init__(self, sim, kind, x, y, z):
        self.sim, self.kind = sim, kind
        self.x, self.y, self.z = x, y, z
        self.E, self.H = [], []

    def _sample(self, F):
        if self.kind == "line":
            return F[self.x, self.y, self.z]
        return np.array([[[F[r, c, [p]][0] for p in self.z] for c in self.y] for r in self.x])

    def detect_E(self):
        self.E.append(self._sample(self.sim.E))

    def detect_H(self):
        self.H.append(self._sample(self.sim.H))


class _CurrentDetector:
    """Loop current around z at each point of an (inclusive) block, sampled
    after every H half step (fdtd/detectors.py:284-476). The second loop's
    y-term is accumulated into the first loop's (the `+=` at
    detectors.py:456), which the restatement keeps."""

    def __init__(self, sim, xs, ys, zs):
        self.sim, self.x, self.y, self.z = sim, xs, ys, zs
        self.I = []

    def _current(self, px, py, pz):
        H, dx = self.sim.H, self.sim.grid_spacing
        v1 = (H[px, py - 1, pz, 0] - H[px, py, pz, 0]) * dx
        v2 = (H[px, py, pz, 1] - H[px - 1, py, pz, 1]) * dx
        c1 = v1 + v2
        v1 = (H[px, py - 1, pz - 1, 0] - H[px, py, pz - 1, 0]) * dx
        v2 += (H[px, py, pz - 1, 1] - H[px - 1, py, pz - 1, 1]) * dx
        c2 = v1 + v2
        return (c1 + c2) / 2.0

    def detect_E(self):
        pass

    def detect_H(self):
        self.I.append([[[self._current(r, c, p) for p in self.z] for c in self.y] for r in self.x])


class _SoftSource:
    """Resistive point source driven by a waveform array, reading the current
    its own current detector measured after the previous H half step
    (fdtd/sources.py:504-637): E_z += (v(q) + R I) / dx."""

    def __init__(self, sim, x, y, z, waveform, impedance):
        self.sim, self.x, self.y, self.z = sim, x, y, z
        self.waveform, self.impedance = waveform, impedance
        self.input_voltage, self.source_voltage = [], []
        self.current_detector = None

    def update_E(self):
        s = self.sim
        q = s.time_steps_passed
        v_in = self.waveform[q] if q < self.waveform.shape[0] else 0.0
        current = self.current_detector.I[-1][0][0][0] if q > 0 else 0.0
        v_out = v_in + self.impedance * current if self.impedance > 0 else v_in
        s.E[self.x, self.y, self.z, 2] += v_out / s.grid_spacing
        self.input_voltage.append([[[v_in]]])
        self.source_voltage.append([[[v_out]]])

    def update_H(self):
        pass


# ---------------------------------------------------------------- the grid
class Sim:
    """Grid state plus the leap-frog loop (fdtd/grid.py:90-331)."""

    def __init__(self, shape, grid_spacing=155e-9, permittivity=1.0, permeability=1.0,
                 courant_number=None):
        self.grid_spacing = float(grid_spacing)
        if len(shape) != 3:
            raise ValueError(f"invalid grid shape {shape}")
        self.Nx, self.Ny, self.Nz = (self.cells_of(v) for v in shape)
        self.D = int(self.Nx > 1) + int(self.Ny > 1) + int(self.Nz > 1)
        cmax = float(self.D) ** (-0.5)
        if courant_number is None:
            self.courant_number = 0.99 * cmax
        elif courant_number > cmax:
            raise ValueError(f"courant_number {courant_number} too high for a {self.D}D simulation")
        else:
            self.courant_number = float(courant_number)
        self.time_step = self.courant_number * self.grid_spacing / SPEED_OF_LIGHT
        shp = (self.Nx, self.Ny, self.Nz, 3)
        self.E = np.zeros(shp)
        self.H = np.zeros(shp)
        eps = np.asarray(permittivity, dtype=np.float64)
        mu = np.asarray(permeability, dtype=np.float64)
        if eps.ndim == 3:
            eps = eps[..., None]
        if mu.ndim == 3:
            mu = mu[..., None]
        self.inverse_permittivity = np.ones(shp) / eps
        self.inverse_permeability = np.ones(shp) / mu
        self.time_steps_passed = 0
        self.boundaries, self.objects, self.sources, self.detectors = [], [], [], []
        self._faces_taken = set()

    def promote_dtypes_to_complex(self):
        """grid.py:353-356: E and H become complex, and so does the CPML state
        of every boundary registered so far. A periodic boundary has no such
        state: the reference's Boundary.promote_dtypes_to_complex then fails
        with AttributeError (boundaries.py:106-107), and so does this."""
        self.E = self.E.astype(np.complex128)
        self.H = self.H.astype(np.complex128)
        for b in self.boundaries:
            if not hasattr(b, "promote_dtypes_to_complex"):
                raise AttributeError(f"'{type(b).__name__}' object has no attribute 'phi_E'")
            b.promote_dtypes_to_complex()

    # -- unit handling, grid.py:171-223
    def cells_of(self, d):
        return d if isinstance(d, int) else int(float(d) / self.grid_spacing + 0.5)

    def steps_of(self, t):
        return t if isinstance(t, int) else int(float(t) / self.time_step + 0.5)

    def _key(self, k):
        if isinstance(k, slice):
            f = lambda v: self.cells_of(v) if isinstance(v, float) else v
            return slice(f(k.start), f(k.stop), f(k.step))
        try:
            return [self.cells_of(v) for v in k]
        except TypeError:
            return [self.cells_of(k)]

    def split_key(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        if len(key) == 0 or len(key) > 3:
            raise KeyError("maximum number of indices for the grid is 3")
        key = tuple(key) + (slice(None),) * (3 - len(key))
        return tuple(self._key(k) for k in key)

    # -- registration (grid.__setitem__ -> component._register_grid)
    def add(self, kind, key, **p):
        x, y, z = self.split_key(key)
        p.pop("name", None)
        if kind == "PML":
            return self._add_pml(x, y, z, a=p.get("a", 1e-8))
        if kind == "PeriodicBoundary":
            return self._add_periodic(x, y, z)
        if kind in ("Object", "AbsorbingObject", "AnisotropicObject"):
            return self._add_object(kind, x, y, z, **p)
        if kind in ("PointSource", "LineSource", "PlaneSource"):
            return self._add_source(kind, x, y, z, **p)
        if kind in ("LineDetector", "BlockDetector"):
            return self._add_detector(kind, x, y, z)
        if kind == "CurrentDetector":
            return self._add_current_detector(x, y, z)
        if kind == "SoftArbitraryPointSource":
            try:
                (x,), (y,), (z,) = x, y, z
            except (TypeError, ValueError):
                raise ValueError("a point source should be placed on a single grid cell.")
            src = _SoftSource(self, self.cells_of(x), self.cells_of(y), self.cells_of(z),
                              np.asarray(p["waveform_array"]), p.get("impedance", 0.0))
            self.sources.append(src)
            src.current_detector = self._add_current_detector([src.x], [src.y], [src.z])
            return src
        raise ValueError(kind)

    def _add_current_detector(self, x, y, z):
        """Point lists stay as given; otherwise inclusive ranges
        (detectors.py:330-372)."""
        if isinstance(x, list) and isinstance(y, list) and isinstance(z, list):
            if not (len(x) == len(y) == len(z)):
                raise IndexError("sources require grid to be indexed with slices or equal length list-indices")
            xs, ys, zs = x, y, z
        else:
            r = []
            for sl, n in ((x, self.Nx), (y, self.Ny), (z, self.Nz)):
                if isinstance(sl, list):
                    sl = slice(sl[0], sl[-1])
                a = sl.start if sl.start is not None else 0
                b = sl.stop if sl.stop is not None else n
                r.append(list(range(a, b + 1)))
            xs, ys, zs = r
        d = _CurrentDetector(self, xs, ys, zs)
        self.detectors.append(d)
        return d

    def _claim(self, *faces):
        for f in faces:
            if f in self._faces_taken:
                raise AttributeError(f"grid already has an {f} boundary!")
        self._faces_taken.update(faces)

    def _add_periodic(self, x, y, z):
        def single(s):  # boundaries.py:62-77
            if isinstance(s, list):
                if len(s) > 1:
                    raise ValueError("Use slices or single numbers to index the grid for a boundary")
                return s[0]
            if s.start is not None and s.stop is not None and (s.start == s.stop or abs(s.start - s.stop) == 1):
                return s.start
            return s
        pos = [single(s) for s in (x, y, z)]
        for axis in range(3):  # boundaries.py:158-180
            if pos[axis] == 0 or pos[axis] == -1:
                n = "xyz"[axis]
                self._claim(n + "low", n + "high")
                b = _Periodic(self, axis)
                self.boundaries.append(b)
                return b
        raise IndexError("A periodic boundary should be placed at the boundary of the grid")

    def _add_pml(self, x, y, z, a):
        for s in (x, y, z):
            if not isinstance(s, slice):
                raise ValueError("One can only use slices to index the grid for a PML")
        for axis, s in enumerate((x, y, z)):  # boundaries.py:301-358
            n = "xyz"[axis]
            if (s.start is None or s.start == 0) and s.stop is not None and s.stop > 0:
                self._claim(n + "low")
                f = _Pml(self, axis, False, s.stop, a)
                self.boundaries.append(f)
                return f
            if s.start is not None and s.stop is None and s.start < 0:
                self._claim(n + "high")
                f = _Pml(self, axis, True, -s.start, a)
                self.boundaries.append(f)
                return f
        raise IndexError("not a valid slice for a PML")

    def _add_object(self, kind, x, y, z, permittivity, conductivity=None):
        def box(s, n):  # objects.py:94-116
            if isinstance(s, list):
                if len(s) == 1:
                    return slice(s[0], s[0] + 1)
                raise IndexError("One can only use slices or single indices to index the grid for an Object")
            if s.step is not None and s.step != 1:
                raise IndexError("Can only use slices with unit step to index the grid for an Object")
            start = 0 if s.start is None else s.start
            start = n + start if start < 0 else start
            stop = n if s.stop is None else s.stop
            stop = n + stop if stop < 0 else stop
            return slice(start, stop)
        kinds = {"Object": "plain", "AbsorbingObject": "absorbing", "AnisotropicObject": "anisotropic"}
        o = _Object(self, kinds[kind], box(x, self.Nx), box(y, self.Ny), box(z, self.Nz),
                    permittivity, conductivity)
        self.objects.append(o)
        return o

    def _line_indices(self, x, y, z, convert):
        """Diagonal index lists (sources.py:209-276, detectors.py:63-112)."""
        if isinstance(x, list) and isinstance(y, list) and isinstance(z, list):
            if not (len(x) == len(y) == len(z)):
                raise IndexError("sources require grid to be indexed with slices or equal length list-indices")
            return ([convert(v) for v in x], [convert(v) for v in y], [convert(v) for v in z]), None
        ends = []
        for s, n in ((x, self.Nx), (y, self.Ny), (z, self.Nz)):
            if isinstance(s, list):
                s = slice(convert(s[0]), convert(s[-1]))
            ends.append((convert(s.start if s.start is not None else 0),
                         convert(s.stop if s.stop is not None else n)))
        m = max(abs(b - a) for a, b in ends)
        idx = tuple([int(v) for v in np.array(np.linspace(a, b, m, endpoint=False), np.int64)]
                    for a, b in ends)
        return idx, m

    def _add_source(self, kind, x, y, z, period=15, amplitude=1.0, phase_shift=0.0,
                    pulse=False, cycle=5, hanning_dt=10.0, polarization="z"):
        if kind == "PointSource":
            try:
                (x,), (y,), (z,) = x, y, z
            except (TypeError, ValueError):
                raise ValueError("a point source should be placed on a single grid cell.")
            s = _Source(self, "point", period, amplitude, phase_shift, pulse, cycle, hanning_dt)
            s.x, s.y, s.z = (self.cells_of(v) for v in (x, y, z))
        elif kind == "LineSource":
            s = _Source(self, "line", period, amplitude, phase_shift, pulse, cycle, hanning_dt)
            (s.x, s.y, s.z), m = self._line_indices(x, y, z, self.cells_of)
            if m is not None and m < 2:
                raise ValueError("a LineSource should consist of at least two gridpoints")
            L = len(s.x)
            # Gaussian-like profile over the squared distance, sources.py:197-207
            vect = np.array((np.array(s.x) - s.x[L // 2]) ** 2 + (np.array(s.y) - s.y[L // 2]) ** 2
                            + (np.array(s.z) - s.z[L // 2]) ** 2, np.float64)
            prof = np.exp(-(vect ** 2) / (2 * (0.5 * vect.max()) ** 2))
            prof /= prof.sum()
            prof *= amplitude
            s.profile = prof
        else:
            s = _Source(self, "plane", period, amplitude, phase_shift, polarization=polarization)
            self._plane_box(s, x, y, z)
            shape = (s.x.stop - s.x.start, s.y.stop - s.y.start, s.z.stop - s.z.start)
            s.profile = amplitude * np.ones(shape)
        self.sources.append(s)
        return s

    def _plane_box(self, s, x, y, z):
        """sources.py:390-474: ascending slices, flatness checks, E/H polarisation."""
        sl = []
        for v, n in ((x, self.Nx), (y, self.Ny), (z, self.Nz)):
            if not isinstance(v, slice):
                if isinstance(v, list):
                    (v,) = v
                v = slice(self.cells_of(v), self.cells_of(v) + 1)
            a = self.cells_of(v.start if v.start is not None else 0)
            b = self.cells_of(v.stop if v.stop is not None else n)
            sl.append(slice(a, b) if a < b else (slice(b, a) if a > b else slice(a, a + 1)))
        lens = [q.stop - q.start for q in sl]
        if lens.count(0) > 0:
            raise ValueError("Given location for PlaneSource results in slices of length 0!")
        if lens.count(1) == 0:
            raise ValueError("Given location for PlaneSource is not a 2D plane!")
        if lens.count(1) > 1:
            raise ValueError("Given location for PlaneSource should have no more than one dimension in which it's flat.")
        ep = "xyz".index(s.polarization)
        if lens[ep] == 1:
            raise ValueError("PlaneSource cannot be polarized perpendicular to the orientation of the plane.")
        # H polarisation: the component orthogonal to E in the plane
        probe, first, second = [(sl[2], 1, 2), (sl[2], 0, 2), (sl[1], 0, 1)][ep]
        s.Hpol = first if probe.stop - probe.start == 1 else second
        s.Epol = ep
        s.x, s.y, s.z = sl

    def _add_detector(self, kind, x, y, z):
        if kind == "LineDetector":
            (xs, ys, zs), _ = self._line_indices(x, y, z, lambda v: v)
        else:
            if isinstance(x, list) and isinstance(y, list) and isinstance(z, list):
                if not (len(x) == len(y) == len(z)):
                    raise IndexError("sources require grid to be indexed with slices or equal length list-indices")
                xs, ys, zs = x, y, z
            else:
                r = []
                for s, n in ((x, self.Nx), (y, self.Ny), (z, self.Nz)):
                    if isinstance(s, list):
                        s = slice(s[0], s[-1])
                    a = s.start if s.start is not None else 0
                    b = s.stop if s.stop is not None else n
                    r.append(list(range(a, b + 1)))  # inclusive, detectors.py:236-238
                xs, ys, zs = r
        d = _Detector(self, "line" if kind == "LineDetector" else "block", xs, ys, zs)
        self.detectors.append(d)
        return d

    # -- the loop, grid.py:250-331
    @property
    def time_passed(self):
        return self.time_steps_passed * self.time_step

    def update_E(self):
        for b in self.boundaries:
            b.update_phi_E()
        curl = curl_H(self.H)
        self.E += self.courant_number * self.inverse_permittivity * curl
        for o in self.objects:
            o.update_E(curl)
        for b in self.boundaries:
            b.update_E()
        for s in self.sources:
            s.update_E()
        for d in self.detectors:
            d.detect_E()

    def update_H(self):
        for b in self.boundaries:
            b.update_phi_H()
        curl = curl_E(self.E)
        self.H -= self.courant_number * self.inverse_permeability * curl
        for b in self.boundaries:
            b.update_H()
        for s in self.sources:
            s.update_H()
        for d in self.detectors:
            d.detect_H()

    def step(self):
        self.update_E()
        self.update_H()
        self.time_steps_passed += 1

    def run(self, total_time, progress_bar=False):
        if isinstance(total_time, float):
            total_time /= self.time_step
        for _ in range(int(total_time)):
            self.step()

    def reset(self):
        """Zero E, H and the step counter; PML state is kept (grid.py:327-331)."""
        self.H *= 0.0
        self.E *= 0.0
        self.time_steps_passed *= 0


# ------------------------------------------------------ frequency routines
# fdtd/fourier.py:52-266 with the numpy backend and default arguments: the
# history padded with its last value ('edge') to the bin count the default
# window asks for, numpy's FFT, bins [0, -1).
def fr_padding(n, dt):
    end_freq = (n / 2.0) / (dt * n)
    return math.ceil(n / ((end_freq - 0) * dt)) - n


def fr_fft(history, dt):
    x = np.asarray(history, dtype=np.float64).reshape(-1)
    x = np.pad(x, (0, fr_padding(x.shape[0], dt)), "edge")
    return np.fft.fftfreq(x.shape[0], d=dt)[0:-1], np.fft.fft(x)[0:-1]


def fr_impedance(voltage, current, dt, steps):
    v = np.asarray(voltage, dtype=np.float64)
    pad = fr_padding(v.shape[0], dt)
    v = np.pad(np.reshape(v, steps), (0, pad), "edge")
    i = np.pad(np.reshape(np.asarray(current, dtype=np.float64), steps), (0, pad), "edge")
    vs, cs = np.fft.fft(v), np.fft.fft(i)
    z = np.divide(vs, cs, out=np.zeros_like(vs), where=cs != 0)
    return np.fft.fftfreq(v.shape[0], d=dt)[0:-1], z[0:-1]
