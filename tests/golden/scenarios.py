"""Declarative FDTD scenarios shared by the fixture generator, the oracle tests
and the HIP parity tests.

A scenario is a plain dict:

    shape, grid_spacing, courant_number, permittivity, permeability
    init        : optional {"E": array, "H": array} written before stepping
    components  : list of (kind, params, key, overrides) in registration order
                  kind      - class name in the fdtd API ("PML", "Object", ...)
                  params    - constructor keyword arguments
                  key       - the index passed to ``grid[key] = component``
                  overrides - attributes assigned on the component right after
                              registration (e.g. a full 3x3 inverse
                              permittivity for an AnisotropicObject)
    program     : list of ops: ("run", T) | ("step", T) | ("reset",) |
                  ("manual_pulse", T, (x, y, z, comp))

``build_fdtd(mod, sc)`` realises a scenario with any module exposing the fdtd
API (the reference package ``fdtd`` when generating fixtures, ``fdtd_amd`` in
the parity tests). ``run_program`` executes the op list. All random inputs come
from ``numpy.random.RandomState`` (a stream numpy keeps stable across
releases); fixtures also store a checksum of the inputs to detect drift.
"""

import math

import numpy as np

C0 = 299792458.0  # speed of light (m/s), fdtd/constants.py:5


def _spd_inverse_field(rs, shape, lo=1.0, hi=9.0):
    """Per-cell inverse of a random SPD 3x3 permittivity with eigenvalues in [lo, hi]."""
    n = int(np.prod(shape))
    out = np.empty((n, 3, 3))
    for i in range(n):
        q, _ = np.linalg.qr(rs.randn(3, 3))
        lam = rs.uniform(lo, hi, size=3)
        eps = q.T @ np.diag(lam) @ q
        out[i] = np.linalg.inv(eps)
    return out.reshape(tuple(shape) + (3, 3))


def scenario_c1():
    """The 2D plain-method configuration (161x97x1, PML on x/y, periodic z,
    continuous 1550 nm LineSource, LineDetector at x=140, eps=1.7^2 block)."""
    s = slice
    return dict(
        name="c1",
        shape=(161, 97, 1),
        grid_spacing=155e-9,
        courant_number=2 ** -0.5,
        permittivity=1.0,
        permeability=1.0,
        components=[
            ("PML", dict(name="pml_xlow"), (s(0, 10), s(None), s(None)), {}),
            ("PML", dict(name="pml_xhigh"), (s(-10, None), s(None), s(None)), {}),
            ("PML", dict(name="pml_ylow"), (s(None), s(0, 10), s(None)), {}),
            ("PML", dict(name="pml_yhigh"), (s(None), s(-10, None), s(None)), {}),
            ("PeriodicBoundary", dict(name="zbounds"), (s(None), s(None), 0), {}),
            ("LineSource", dict(period=1550e-9 / C0, name="source"), (20, s(20, 77), 0), {}),
            ("LineDetector", dict(name="detector"), (140, s(None), 0), {}),
            ("Object", dict(permittivity=1.7 ** 2, name="object"), (s(11, 32), s(30, 84), 0), {}),
        ],
        program=[("run", 300)],
    )


def scenario_g2():
    """Small 3D run touching every hot-path feature: six PML faces of unequal
    thickness, per-voxel eps (3D) and per-component mu (4D), overlapping Object
    and AbsorbingObject, an AnisotropicObject whose inverse permittivity is
    overwritten by a full SPD inverse, pulsed/continuous point and line sources
    (one with a float period in seconds, one diagonal), a hard PlaneSource, two
    LineDetectors (one diagonal) and a BlockDetector, random initial fields."""
    rs = np.random.RandomState(1234)
    shape = (24, 20, 16)
    eps = rs.uniform(1.0, 4.0, size=shape)
    mu = rs.uniform(1.0, 2.0, size=shape + (3,))
    e0 = 0.05 * rs.randn(*shape, 3)
    h0 = 0.05 * rs.randn(*shape, 3)
    aniso_eps = rs.uniform(1.5, 3.0, size=(6, 5, 4, 3))
    aniso_inv = _spd_inverse_field(rs, (6, 5, 4))
    s = slice
    return dict(
        name="g2",
        shape=shape,
        grid_spacing=100e-9,
        courant_number=0.5,
        permittivity=eps,
        permeability=mu,
        init=dict(E=e0, H=h0),
        components=[
            ("PML", dict(), (s(0, 4), s(None), s(None)), {}),
            ("PML", dict(a=2e-8), (s(None), s(-5, None), s(None)), {}),
            ("PML", dict(), (s(-4, None), s(None), s(None)), {}),
            ("PML", dict(), (s(None), s(0, 3), s(None)), {}),
            ("PML", dict(), (s(None), s(None), s(0, 4)), {}),
            ("PML", dict(), (s(None), s(None), s(-4, None)), {}),
            ("Object", dict(permittivity=2.5, name="obj"), (s(6, 12), s(5, 11), s(4, 9)), {}),
            ("AbsorbingObject", dict(permittivity=1.5, conductivity=2e4, name="absorber"),
             (s(10, 16), s(8, 14), s(6, 12)), {}),
            ("AnisotropicObject", dict(permittivity=aniso_eps, name="aniso"),
             (s(14, 20), s(3, 8), s(9, 13)), {"inverse_permittivity": aniso_inv}),
            ("PointSource", dict(period=10, pulse=True, name="psrc"), (12, 10, 8), {}),
            ("LineSource", dict(period=1.3e-14, phase_shift=0.3, amplitude=0.7, name="lsrc"),
             (s(4, 18), 15, 10), {}),
            ("LineSource", dict(period=12, pulse=True, cycle=3, hanning_dt=4.0, name="dsrc"),
             (s(5, 15), s(2, 9), 13), {}),
            ("PlaneSource", dict(period=15, amplitude=0.5, polarization="y", name="plane"),
             (7, s(4, 16), s(3, 13)), {}),
            ("LineDetector", dict(name="ld"), (s(2, 22), 10, 8), {}),
            ("LineDetector", dict(name="ld_diag"), (s(3, 19), s(4, 16), s(2, 14)), {}),
            ("BlockDetector", dict(name="bdet"), (s(10, 12), s(5, 7), s(3, 4)), {}),
        ],
        program=[("run", 60)],
    )


def _periodic_scenario(name, axis, periodic_first, steps=30):
    """Periodic boundary on `axis` combined with PML faces on the two other
    axes, registered either before or after the periodic boundary; the order
    decides which plane-0 values the PML corrections see."""
    rs = np.random.RandomState(77 + axis + (10 if periodic_first else 0))
    shape = (12, 10, 8)
    s = slice
    full = [s(None)] * 3
    per_key = list(full)
    per_key[axis] = 0 if periodic_first else -1
    per = ("PeriodicBoundary", dict(), tuple(per_key), {})
    pmls = []
    for other in range(3):
        if other == axis:
            continue
        lo = list(full)
        lo[other] = s(0, 3)
        hi = list(full)
        hi[other] = s(-2, None)
        pmls.append(("PML", dict(), tuple(lo), {}))
        pmls.append(("PML", dict(), tuple(hi), {}))
    comps = [per] + pmls if periodic_first else pmls[:2] + [per] + pmls[2:]
    src_key = [5, 4, 3]
    comps.append(("PointSource", dict(period=9), tuple(src_key), {}))
    det_key = [s(None)] * 3
    det_key[(axis + 1) % 3] = 2
    det_key[(axis + 2) % 3] = 1
    comps.append(("LineDetector", dict(), tuple(det_key), {}))
    return dict(
        name=name,
        shape=shape,
        grid_spacing=80e-9,
        courant_number=0.55,
        permittivity=rs.uniform(1.0, 3.0, size=shape),
        permeability=1.0,
        init=dict(E=0.1 * rs.randn(*shape, 3), H=0.1 * rs.randn(*shape, 3)),
        components=comps,
        program=[("run", steps)],
    )


def scenario_manual():
    """The manual half-stepping pattern of the reference detector test: the
    host writes E between update_E and update_H and advances the counter by
    hand; then reset() (which keeps the PML state) and a plain run."""
    s = slice
    return dict(
        name="manual",
        shape=(10, 10, 10),
        grid_spacing=100e-9,
        courant_number=None,
        permittivity=1.0,
        permeability=1.0,
        components=[
            ("PML", dict(), (s(0, 3), s(None), s(None)), {}),
            ("PML", dict(), (s(None), s(None), s(-3, None)), {}),
            ("AbsorbingObject", dict(permittivity=1.0, conductivity=1e9), (s(3, 5), s(3, 5), s(3, 5)), {}),
            ("BlockDetector", dict(), (s(4, 4), s(4, 4), s(4, 4)), {}),
            ("LineDetector", dict(), (s(1, 9), 6, 6), {}),
        ],
        program=[("manual_pulse", 40, (4, 4, 4, 2)), ("reset",), ("step", 12)],
    )


def scenario_float_keys():
    """Registration with positions in metres (floats) and a float period:
    exercises the distance/time rounding rules of the grid."""
    s = slice
    dx = 50e-9
    return dict(
        name="float_keys",
        shape=(1.0e-6, 0.8e-6, 12),
        grid_spacing=dx,
        courant_number=0.5,
        permittivity=1.0,
        permeability=1.0,
        components=[
            ("PML", dict(), (s(0, 3.1e-7), s(None), s(None)), {}),
            ("Object", dict(permittivity=3.0), (s(4.0e-7, 6.2e-7), s(2.1e-7, 5.0e-7), s(2, 9)), {}),
            ("PointSource", dict(period=2.0e-15, phase_shift=1.1), (3.0e-7, 4.0e-7, 6), {}),
            ("LineSource", dict(period=11), (s(2.0e-7, 8.0e-7), 5.0e-7, 5), {}),
            ("LineDetector", dict(), (s(None), 7, 6), {}),
        ],
        program=[("run", 25)],
    )


def scenario_1d():
    """1D grid (N,1,1): a hard-sourced pulse travels along x."""
    return dict(
        name="oned",
        shape=(64, 1, 1),
        grid_spacing=100e-9,
        courant_number=1.0,
        permittivity=1.0,
        permeability=1.0,
        components=[
            ("PML", dict(), (slice(-8, None), slice(None), slice(None)), {}),
            ("PointSource", dict(period=16, pulse=True, cycle=2, hanning_dt=8.0), (10, 0, 0), {}),
            ("LineDetector", dict(), (slice(None), 0, 0), {}),
        ],
        program=[("run", 50)],
    )


def scenario_soft_source():
    """A probe-fed patch in miniature (after the reference's antenna test,
    tests/test_antenna_impedance.py:18-95): copper planes as absorbing
    objects, a dielectric substrate, a resistive soft source (Gaussian pulse
    waveform) that reads its own loop current every step, and a stand-alone
    current detector over a block. The impedance is 0.2 in the grid's
    normalised units: the source feeds R*I/dx back into E every step, and
    the loop gain ~2.3 R must stay below 1 for the run to stay bounded."""
    s = slice
    steps = 60
    dx = 0.4e-3
    t = np.arange(steps) * (0.99 * 3 ** -0.5 * dx / 299792458.0)
    wave = _gaussian(t, 12 * t[1], 20 * t[1])
    copper = dict(permittivity=1.0, conductivity=1e8)
    return dict(
        name="soft_source",
        shape=(20, 16, 14),
        grid_spacing=dx,
        courant_number=None,
        permittivity=1.0,
        permeability=1.0,
        components=[
            ("PML", dict(), (s(0, 3), s(None), s(None)), {}),
            ("PML", dict(), (s(-3, None), s(None), s(None)), {}),
            ("PML", dict(), (s(None), s(0, 3), s(None)), {}),
            ("PML", dict(), (s(None), s(-3, None), s(None)), {}),
            ("PML", dict(), (s(None), s(None), s(0, 3)), {}),
            ("PML", dict(), (s(None), s(None), s(-3, None)), {}),
            ("AbsorbingObject", copper, (s(3, -3), s(3, -3), s(3, 4)), {}),
            ("AbsorbingObject", copper, (s(3, -3), s(3, -3), s(7, 8)), {}),
            ("SoftArbitraryPointSource", dict(waveform_array=wave, impedance=0.2), (9, 8, 4), {}),
            ("AbsorbingObject", copper, (s(9, 10), s(8, 9), s(5, 8)), {}),
            ("Object", dict(permittivity=2.42), (s(3, 9), s(3, -3), s(4, 7)), {}),
            ("Object", dict(permittivity=2.42), (s(10, -3), s(3, -3), s(4, 7)), {}),
            ("CurrentDetector", dict(name="loop"), (s(5, 7), s(5, 6), s(5, 5)), {}),
        ],
        program=[("step", steps)],
    )


def scenario_complex_object():
    """Complex fields (grid.py:353-356, objects.py:51-52, boundaries.py:106-116):
    an Object with a complex permittivity promotes E, H and the CPML state
    registered before it to complex128; a second Object carries a per-cell
    complex permittivity array, a third a real one. PML on five faces
    (registered first), a pulsed point source, a line source, a line detector
    and random real initial fields."""
    rs = np.random.RandomState(71)
    s = slice
    shape = (22, 18, 14)
    eps_arr = rs.uniform(1.5, 4.0, size=(5, 4, 3)) + 1j * rs.uniform(0.005, 0.05, size=(5, 4, 3))
    return dict(
        name="complex_object",
        shape=shape,
        grid_spacing=50e-9,
        courant_number=None,
        permittivity=1.0,
        permeability=1.0,
        init={"E": rs.uniform(-1, 1, size=shape + (3,)), "H": rs.uniform(-1, 1, size=shape + (3,))},
        components=[
            ("PML", dict(), (s(0, 3), s(None), s(None)), {}),
            ("PML", dict(), (s(-3, None), s(None), s(None)), {}),
            ("PML", dict(), (s(None), s(0, 3), s(None)), {}),
            ("PML", dict(), (s(None), s(-3, None), s(None)), {}),
            ("PML", dict(), (s(None), s(None), s(0, 3)), {}),
            ("Object", dict(permittivity=2.5 + 0.05j, name="lossy"), (s(6, 12), s(5, 11), s(4, 9)), {}),
            ("Object", dict(permittivity=eps_arr), (s(13, 18), s(4, 8), s(6, 9)), {}),
            ("Object", dict(permittivity=1.8), (s(4, 8), s(11, 15), s(8, 12)), {}),
            ("PointSource", dict(period=10, pulse=True), (10, 9, 7), {}),
            ("LineSource", dict(period=14), (s(5, 17), 13, 10), {}),
            ("LineDetector", dict(name="det"), (s(3, 19), 9, 10), {}),
        ],
        program=[("step", 50)],
    )


def scenario_complex_absorbing():
    """Absorbing and anisotropic objects on the complex-field path
    (objects.py:163-277 after grid.py:353-356): a plain Object with a complex
    permittivity promotes the fields; then an AbsorbingObject with a complex
    permittivity (complex absorption factor: numpy's complex division each
    step), an AbsorbingObject with a real permittivity (real arrays on complex
    fields: a real division, then mixed products), an AnisotropicObject
    whose inverse is overwritten by a per-cell complex 3x3 matrix (complex
    einsum), and one with a real permittivity (real 3x3 on complex fields).
    PML on four faces (registered first), a pulsed point source, a line
    source, a line detector and random real initial fields."""
    rs = np.random.RandomState(83)
    s = slice
    shape = (20, 16, 14)
    aniso_shape = (4, 3, 4)
    full = _spd_inverse_field(rs, aniso_shape, 1.5, 4.0).astype(np.complex128)
    full += 1j * rs.uniform(-0.01, 0.01, size=full.shape)
    # per cell, so that the real division of the real object's factor is
    # seen on many values (a quotient and a product by the reciprocal agree
    # on about half of them)
    real_perm = rs.uniform(1.5, 3.0, size=(4, 4, 4))
    return dict(
        name="complex_absorbing",
        shape=shape,
        grid_spacing=50e-9,
        courant_number=None,
        permittivity=1.0,
        permeability=1.0,
        init={"E": rs.uniform(-1, 1, size=shape + (3,)), "H": rs.uniform(-1, 1, size=shape + (3,))},
        components=[
            ("PML", dict(), (s(0, 3), s(None), s(None)), {}),
            ("PML", dict(), (s(-3, None), s(None), s(None)), {}),
            ("PML", dict(), (s(None), s(0, 3), s(None)), {}),
            ("PML", dict(), (s(None), s(None), s(-3, None)), {}),
            ("Object", dict(permittivity=2.2 + 0.03j), (s(4, 8), s(4, 9), s(3, 7)), {}),
            ("AbsorbingObject", dict(permittivity=3.0 + 0.04j, conductivity=2e4), (s(9, 14), s(3, 7), s(4, 9)), {}),
            ("AbsorbingObject", dict(permittivity=real_perm, conductivity=5e4), (s(12, 16), s(9, 13), s(2, 6)), {}),
            ("AnisotropicObject", dict(permittivity=2.0), (s(5, 9), s(10, 13), s(7, 11)), {"inverse_permittivity": full}),
            ("AnisotropicObject", dict(permittivity=1.6), (s(14, 17), s(4, 8), s(9, 12)), {}),
            ("PointSource", dict(period=10, pulse=True), (10, 8, 7), {}),
            ("LineSource", dict(period=14), (s(4, 16), 12, 9), {}),
            ("LineDetector", dict(name="det"), (s(3, 17), 8, 9), {}),
        ],
        program=[("step", 40)],
    )


def all_scenarios():
    out = [scenario_c1(), scenario_g2(), scenario_manual(), scenario_float_keys(), scenario_1d(),
           scenario_soft_source(), scenario_complex_object(), scenario_complex_absorbing()]
    for axis in range(3):
        for first in (True, False):
            out.append(_periodic_scenario(f"periodic_{'xyz'[axis]}_{'first' if first else 'last'}", axis, first))
    return out


def get_scenario(name):
    for sc in all_scenarios():
        if sc["name"] == name:
            return sc
    raise KeyError(name)


def input_checksum(sc):
    """A float checksum over every array input of a scenario."""
    acc = 0.0
    arrays = []
    for k in ("permittivity", "permeability"):
        if isinstance(sc[k], np.ndarray):
            arrays.append(sc[k])
    for v in sc.get("init", {}).values():
        arrays.append(v)
    for _, params, _, overrides in sc["components"]:
        for v in list(params.values()) + list(overrides.values()):
            if isinstance(v, np.ndarray):
                arrays.append(v)
    for i, a in enumerate(arrays):
        a = np.asarray(a)
        if np.iscomplexobj(a):  # real and imaginary parts both count
            acc += float(np.sum(a.imag.astype(np.float64) * (i + 1)))
            a = a.real
        acc += float(np.sum(a.astype(np.float64) * (i + 1)))
    return acc


# ---------------------------------------------------------------------------
# realisation through an fdtd-compatible module


def build_fdtd(mod, sc, to_array=None):
    """Build a grid with the fdtd API of module `mod`. `to_array` converts a
    numpy input to the module's array type (identity for numpy backends)."""
    conv = to_array or (lambda a: a)
    kw = dict(shape=sc["shape"], grid_spacing=sc["grid_spacing"])
    if sc["courant_number"] is not None:
        kw["courant_number"] = sc["courant_number"]
    for k in ("permittivity", "permeability"):
        v = sc[k]
        kw[k] = conv(v) if isinstance(v, np.ndarray) else v
    grid = mod.Grid(**kw)
    comps = []
    for kind, params, key, overrides in sc["components"]:
        p = {k: (conv(v) if isinstance(v, np.ndarray) else v) for k, v in params.items()}
        cls = getattr(mod, kind, None) or getattr(mod.sources, kind)  # some live in submodules only
        comp = cls(**p)
        grid[key] = comp
        for attr, val in overrides.items():
            setattr(comp, attr, conv(val) if isinstance(val, np.ndarray) else val)
        comps.append(comp)
    init = sc.get("init")
    if init:
        grid.E[...] = conv(init["E"])
        grid.H[...] = conv(init["H"])
    return grid, comps


def _gaussian(x, fwhm, center):
    # normalized_gaussian_pulse, fdtd/waveforms.py:34-46 (peak 1 at `center`)
    sigma = fwhm / (2.0 * math.sqrt(2.0 * math.log(2)))
    return np.exp(-(((x - center) ** 2.0) / (2.0 * (sigma ** 2.0))))


def run_program(grid, sc, write=None):
    """Execute the scenario's op list on an fdtd-API grid. `write(grid, idx,
    value)` performs a host write into grid.E (default: item assignment)."""
    for op in sc["program"]:
        if op[0] == "run":
            grid.run(op[1], progress_bar=False)
        elif op[0] == "step":
            for _ in range(op[1]):
                grid.step()
        elif op[0] == "reset":
            grid.reset()
        elif op[0] == "manual_pulse":
            n, idx = op[1], op[2]
            for _ in range(n):
                grid.update_E()
                val = float(_gaussian(grid.time_passed, (n / 10.0) * grid.time_step, (n / 2) * grid.time_step))
                if write is None:
                    grid.E[idx] = val
                else:
                    write(grid, idx, val)
                grid.update_H()
                grid.time_steps_passed += 1
        else:
            raise ValueError(op)


def build_oracle(oracle_mod, sc):
    """Build the same scenario with the oracle's own registration API."""
    sim = oracle_mod.Sim(sc["shape"], sc["grid_spacing"], sc["permittivity"], sc["permeability"],
                         sc["courant_number"])
    comps = []
    for kind, params, key, overrides in sc["components"]:
        comp = sim.add(kind, key, **params)
        for attr, val in overrides.items():
            setattr(comp, attr, val)
        comps.append(comp)
    init = sc.get("init")
    if init:
        sim.E[...] = init["E"]
        sim.H[...] = init["H"]
    return sim, comps


def random_complex_scenario(seed, steps=15):
    """random_scenario(seed) moved onto the complex-field path: periodic
    boundaries dropped (the reference cannot promote them), the PML faces
    registered first (a PML registered after the promotion fails there),
    the first object with a complex permittivity (scalar or per cell,
    imaginary part <= 0.05: the complex leap-frog grows otherwise; an
    overwritten anisotropic inverse gets imaginary parts <= 0.01), a plain
    Object added first when the scenario has no object; the other objects,
    of every kind, keep real arrays on the complex fields."""
    sc = random_scenario(seed, steps)
    rs = np.random.RandomState(5000 + seed)
    comps = [c for c in sc["components"] if c[0] != "PeriodicBoundary"]
    pmls = [c for c in comps if c[0] == "PML"]
    objs = [c for c in comps if c[0].endswith("Object")]
    rest = [c for c in comps if c[0] != "PML" and not c[0].endswith("Object")]
    if not objs:
        shape = sc["shape"]
        key = tuple(slice(n // 3, max(n // 3 + 1, 2 * n // 3)) for n in shape)
        objs.append(("Object", dict(permittivity=2.0), key, {}))
    kind, params, key, over = objs[0]
    params, over = dict(params), dict(over)
    perm = params["permittivity"]
    if np.ndim(perm) == 0:
        params["permittivity"] = complex(perm, float(rs.uniform(0.005, 0.05)))
    else:
        params["permittivity"] = perm + 1j * rs.uniform(0.005, 0.05, size=np.shape(perm))
    if "inverse_permittivity" in over:
        inv = over["inverse_permittivity"]
        over["inverse_permittivity"] = inv + 1j * rs.uniform(-0.01, 0.01, size=inv.shape)
    objs[0] = (kind, params, key, over)
    sc["components"] = pmls + objs + rest
    sc["name"] = f"random_complex_{seed}"
    return sc


def random_scenario(seed, steps=15):
    """A random but valid configuration: random grid (sometimes 2D), random
    subset and registration order of PML faces, an optional periodic boundary
    at a random position among them, random objects of every kind, point /
    line / plane sources, line / block detectors, random initial fields."""
    rs = np.random.RandomState(1000 + seed)
    two_d = rs.rand() < 0.25
    shape = [int(rs.randint(6, 17)), int(rs.randint(6, 15)), 1 if two_d else int(rs.randint(6, 15))]
    D = sum(n > 1 for n in shape)
    s = slice
    full = [s(None)] * 3
    bnds = []
    per_axis = int(rs.randint(0, 3)) if rs.rand() < 0.5 else None
    for axis in range(3):
        if shape[axis] < 6 or axis == per_axis:
            continue
        for high in (False, True):
            if rs.rand() < 0.7:
                t = int(rs.randint(2, max(3, shape[axis] // 3 + 1)))
                k = list(full)
                k[axis] = s(-t, None) if high else s(0, t)
                bnds.append(("PML", dict(a=float(rs.choice([1e-8, 3e-8]))), tuple(k), {}))
    rs.shuffle(bnds)
    if per_axis is not None:
        k = list(full)
        k[per_axis] = int(rs.choice([0, -1]))
        bnds.insert(int(rs.randint(0, len(bnds) + 1)), ("PeriodicBoundary", dict(), tuple(k), {}))

    def box():
        lo = [int(rs.randint(0, n)) for n in shape]
        hi = [min(n, l + int(rs.randint(1, 6))) for l, n in zip(lo, shape)]
        return lo, hi, tuple(s(a, b) for a, b in zip(lo, hi))

    objs = []
    for _ in range(int(rs.randint(0, 4))):
        lo, hi, key = box()
        kind = ["Object", "AbsorbingObject", "AnisotropicObject"][int(rs.randint(0, 3))]
        dims = tuple(b - a for a, b in zip(lo, hi))
        perm = float(rs.uniform(1.2, 4.0)) if rs.rand() < 0.5 else rs.uniform(1.2, 4.0, size=dims)
        p = dict(permittivity=perm)
        over = {}
        if kind == "AbsorbingObject":
            p["conductivity"] = float(rs.choice([1e3, 5e4, 1e6]))
        if kind == "AnisotropicObject" and rs.rand() < 0.6:
            over["inverse_permittivity"] = _spd_inverse_field(rs, dims, 1.0, 5.0)
        objs.append((kind, p, key, over))

    srcs = []
    pt = tuple(int(rs.randint(0, n)) for n in shape)
    srcs.append(("PointSource", dict(period=int(rs.randint(6, 20)), pulse=bool(rs.rand() < 0.5),
                                     phase_shift=float(rs.uniform(0, 3))), pt, {}))
    if rs.rand() < 0.8:
        lo = [int(rs.randint(0, n)) for n in shape]
        hi = [int(rs.randint(0, n)) for n in shape]
        ax = int(np.argmax(shape))
        lo[ax], hi[ax] = 0, shape[ax]
        key = tuple(s(a, b) if a != b else a for a, b in zip(lo, hi))
        srcs.append(("LineSource", dict(period=int(rs.randint(6, 20)), pulse=bool(rs.rand() < 0.5),
                                        amplitude=float(rs.uniform(0.5, 2))), key, {}))
    flat = [a for a in range(3) if shape[a] > 1]
    if len(flat) == 3 and rs.rand() < 0.7:
        n_ax = int(rs.choice(flat))
        pol = [a for a in range(3) if a != n_ax][int(rs.randint(0, 2))]
        k = [s(1, n - 1) for n in shape]
        k[n_ax] = int(rs.randint(0, shape[n_ax]))
        srcs.append(("PlaneSource", dict(period=int(rs.randint(6, 20)), amplitude=float(rs.uniform(0.1, 1)),
                                         polarization="xyz"[pol]), tuple(k), {}))
    dets = []
    ax = int(np.argmax(shape))
    k = [int(rs.randint(0, n)) for n in shape]
    k[ax] = s(None)
    dets.append(("LineDetector", dict(), tuple(k), {}))
    lo = [int(rs.randint(0, n - 1)) if n > 1 else 0 for n in shape]
    k = tuple(s(a, min(a + 1, n - 1)) for a, n in zip(lo, shape))
    dets.append(("BlockDetector", dict(), k, {}))

    comps = []
    groups = [bnds, objs, srcs, dets]
    while any(groups):
        g = groups[int(rs.randint(0, 4))]
        if g:
            comps.append(g.pop(0))
    eps = rs.uniform(1.0, 4.0, size=tuple(shape)) if rs.rand() < 0.7 else 2.0
    mu = rs.uniform(1.0, 2.0, size=tuple(shape) + (3,)) if rs.rand() < 0.4 else 1.0
    return dict(
        name=f"random_{seed}",
        shape=tuple(shape),
        grid_spacing=float(rs.choice([50e-9, 100e-9])),
        courant_number=0.95 * D ** -0.5,
        permittivity=eps,
        permeability=mu,
        init=dict(E=0.1 * rs.randn(*shape, 3), H=0.1 * rs.randn(*shape, 3)),
        components=comps,
        program=[("run", steps)],
    )
