"""Gradient scenarios: a grid with CPML on all faces, a Hanning-pulse line
source (the reference's own pulse, fdtd/sources.py:278-297), a plain Object
as the design region and an objective plane. The same builder drives the
reference package (torch backend, for the autograd fixture), the oracle
tests and the HIP tests, since the HIP package has the reference's API.

J = sum over the steps of sum over the plane cells of |E|^2, E after each step.
"""

import numpy as np

C0 = 299792458.0  # fdtd/constants.py

# 48x48x24: the size at which the reference's torch autograd gives the
# fixture gradient
G5 = dict(
    name="g5",
    shape=(48, 48, 24),
    spacing=50e-9,
    courant=3 ** -0.5,
    pml=6,
    steps=200,
    source=(10, slice(14, 34), 12),
    design=(slice(18, 30), slice(16, 32), slice(8, 16)),
    plane=(36, slice(20, 28), slice(8, 16)),
    seed=5,
)

# 128x128x64 with a 64x64x16 design region and a 16x16 plane, 2000 steps
C5 = dict(
    name="c5",
    shape=(128, 128, 64),
    spacing=50e-9,
    courant=3 ** -0.5,
    pml=12,
    steps=2000,
    source=(16, slice(32, 96), 32),
    design=(slice(32, 96), slice(32, 96), slice(24, 40)),
    plane=(110, slice(56, 72), slice(24, 40)),
    seed=7,
)

# a small case for quick checks
TINY = dict(
    name="tiny",
    shape=(20, 18, 14),
    spacing=50e-9,
    courant=3 ** -0.5,
    pml=4,
    steps=60,
    source=(6, slice(5, 13), 7),
    design=(slice(8, 13), slice(6, 12), slice(4, 10)),
    plane=(15, slice(6, 12), slice(5, 9)),
    seed=3,
)


# TINY plus an absorbing slab behind the design region (covers the E-update
# scaling of absorbing objects in the transposed steps)
TINY_ABSORBER = dict(TINY, name="tiny_absorber",
                     absorber=((slice(14, 17), slice(4, 14), slice(3, 11)), 2.0, 1e4))


# TINY plus a hard PlaneSource (x = 10) that cuts through the design region:
# its cells' E (polarisation y) and H components are assigned every half step
TINY_PLANE = dict(TINY, name="tiny_plane",
                  plane_source=((10, slice(4, 14), slice(2, 12)), dict(period=12, amplitude=0.3, polarization="y")))

# Periodic in z (no z faces: the PML faces on x and y are registered first,
# then the PeriodicBoundary) with an AnisotropicObject between the design
# region and the objective plane whose inverse permittivity is overwritten by
# a seeded per-cell SPD inverse (eigenvalues 1.5-4); the reference's torch
# autograd gives the fixture gradient at this size
PA = dict(
    name="pa",
    shape=(36, 32, 12),
    spacing=50e-9,
    courant=3 ** -0.5,
    pml=6,
    steps=150,
    periodic_z=True,
    source=(9, slice(10, 22), 6),
    design=(slice(13, 21), slice(10, 22), slice(2, 10)),
    plane=(28, slice(12, 20), slice(3, 9)),
    aniso=((slice(22, 25), slice(8, 24), slice(0, 12)), 13),
    seed=9,
)

# the same features on TINY (central differences over every design voxel)
TINY_PA = dict(TINY, name="tiny_pa", periodic_z=True, aniso=((slice(13, 15), slice(4, 14), slice(0, 14)), 17))

# tiles that cover the (y, z) plane exactly (nz = 32: 32-cell z rows, ny = 48:
# 8 rows of 256-thread tiles), for the transposed H kernel's guard-free form
FULL_TILES = dict(G5, name="full_tiles", shape=(40, 48, 32), steps=120, source=(10, slice(14, 34), 16),
                  plane=(30, slice(20, 28), slice(8, 16)))


def aniso_inverse(sc):
    """The overwritten inverse permittivity of the anisotropic object:
    (nx, ny, nz, 3, 3), per cell the inverse of a seeded SPD matrix."""
    from scenarios import _spd_inverse_field

    key, seed = sc["aniso"]
    shape = tuple(k.stop - k.start for k in key)
    return _spd_inverse_field(np.random.RandomState(seed), shape, 1.5, 4.0)


def pml_keys(sc):
    """The PML faces, in registration order: x and y, and z unless z is
    periodic."""
    t = sc["pml"]
    s = slice
    keys = [(s(0, t), s(None), s(None)), (s(-t, None), s(None), s(None)),
            (s(None), s(0, t), s(None)), (s(None), s(-t, None), s(None))]
    if not sc.get("periodic_z"):
        keys += [(s(None), s(None), s(0, t)), (s(None), s(None), s(-t, None))]
    return keys


def design_shape(sc):
    return tuple(s.stop - s.start for s in sc["design"])


def design_permittivity(sc):
    """Seeded eps ~ U[1, 12] over the design region, (Nx, Ny, Nz) float64."""
    rng = np.random.default_rng(sc["seed"])
    return rng.uniform(1.0, 12.0, size=design_shape(sc))


def build(mod, sc, permittivity):
    """Grid + registered components in a fixed order; returns (grid, design)."""
    grid = mod.Grid(sc["shape"], grid_spacing=sc["spacing"], permittivity=1.0, permeability=1.0,
                    courant_number=sc["courant"])
    names = ["pml_xlow", "pml_xhigh", "pml_ylow", "pml_yhigh", "pml_zlow", "pml_zhigh"]
    for key, name in zip(pml_keys(sc), names):
        grid[key] = mod.PML(name=name)
    if sc.get("periodic_z"):
        grid[:, :, 0] = mod.PeriodicBoundary(name="zbounds")
    grid[sc["source"]] = mod.LineSource(period=1550e-9 / C0, pulse=True, name="source")
    design = mod.Object(permittivity, name="design")
    grid[sc["design"]] = design
    if sc.get("absorber"):
        key, eps, sigma = sc["absorber"]
        grid[key] = mod.AbsorbingObject(permittivity=eps, conductivity=sigma, name="absorber")
    if sc.get("plane_source"):
        key, kw = sc["plane_source"]
        grid[key] = mod.PlaneSource(name="plane_source", **kw)
    if sc.get("aniso"):
        import torch

        obj = mod.AnisotropicObject(permittivity=2.0, name="aniso")
        grid[sc["aniso"][0]] = obj
        obj.inverse_permittivity = torch.from_numpy(aniso_inverse(sc))
    return grid, design


def sample_voxels(sc, n, seed=11):
    """n distinct design voxels (i, j, k), seeded."""
    rng = np.random.default_rng(seed)
    shape = design_shape(sc)
    flat = rng.choice(int(np.prod(shape)), size=n, replace=False)
    return [tuple(int(v) for v in np.unravel_index(f, shape)) for f in flat]


def build_oracle(oracle_mod, sc, permittivity):
    """The same scenario through the oracle's registration API; returns sim."""
    sim = oracle_mod.Sim(sc["shape"], sc["spacing"], 1.0, 1.0, sc["courant"])
    for key in pml_keys(sc):
        sim.add("PML", key)
    if sc.get("periodic_z"):
        sim.add("PeriodicBoundary", (slice(None), slice(None), 0))
    sim.add("LineSource", sc["source"], period=1550e-9 / C0, pulse=True)
    sim.add("Object", sc["design"], permittivity=permittivity)
    if sc.get("absorber"):
        key, eps, sigma = sc["absorber"]
        sim.add("AbsorbingObject", key, permittivity=eps, conductivity=sigma)
    if sc.get("plane_source"):
        key, kw = sc["plane_source"]
        sim.add("PlaneSource", key, **kw)
    if sc.get("aniso"):
        obj = sim.add("AnisotropicObject", sc["aniso"][0], permittivity=2.0)
        obj.inverse_permittivity = aniso_inverse(sc)
    return sim


def oracle_objective(oracle_mod, sc, permittivity, records=False):
    """J of the scenario stepped by the oracle; with records, also the
    per-step curl H inside the design box and E on the plane, (T, 3, cells)
    each, z fastest (what the forward pass of the gradient records)."""
    sim = build_oracle(oracle_mod, sc, permittivity)
    J = 0.0
    curls, planes = [], []
    for _ in range(sc["steps"]):
        if records:
            c = oracle_mod.curl_H(sim.H)[sc["design"]]
            curls.append(np.moveaxis(c, -1, 0).reshape(3, -1))
        sim.update_E()
        e = sim.E[sc["plane"]]
        J += float((e ** 2).sum())
        if records:
            planes.append(np.moveaxis(e, -1, 0).reshape(3, -1))
        sim.update_H()
        sim.time_steps_passed += 1
    if records:
        return J, np.stack(curls), np.stack(planes)
    return J
