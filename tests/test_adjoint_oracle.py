"""The gradient fixture and the gradient API's host logic (no GPU).

adjoint_g5.npz holds dJ/d permittivity of the 48x48x24 gradient scenario as
the reference's torch backend computes it with autograd. Here it is pinned
independently: its J and final E against the oracle's forward run, and its
gradient against central differences of the oracle's J.
"""

import numpy as np
import pytest
import torch

import adjoint_scenarios as A
import fdtd_amd as fdtd
from fdtd_amd.adjoint import IntensityGradient, _box_of
from fdtd_amd.backend import host_only
from fdtd_amd.sources import SoftArbitraryPointSource
from oracle import fdtd_oracle as orc


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.fixture(scope="module")
def g5(golden):
    return golden("adjoint_g5")


@pytest.fixture(scope="module")
def golden():
    import os

    from conftest import GOLDEN

    return lambda name: np.load(os.path.join(GOLDEN, f"{name}.npz"))


def test_fixture_inputs_regenerate(g5):
    assert np.array_equal(g5["permittivity"], A.design_permittivity(A.G5))


def test_oracle_forward_matches_fixture(g5):
    sc = A.G5
    sim = A.build_oracle(orc, sc, A.design_permittivity(sc))
    J = 0.0
    for _ in range(sc["steps"]):
        sim.step()
        J += float((sim.E[sc["plane"]] ** 2).sum())
    assert abs(J - float(g5["J"])) <= 1e-12 * float(g5["J"])
    assert rel(sim.E, g5["E"]) <= 1e-12


@pytest.mark.slow
def test_fixture_gradient_matches_oracle_central_differences(g5):
    """Two voxels (the one of largest |grad| and a seeded random one):
    (J(eps+h) - J(eps-h)) / 2h of the oracle against the fixture."""
    sc = A.G5
    perm = A.design_permittivity(sc)
    grad = g5["grad"]
    vox = [np.unravel_index(int(np.argmax(np.abs(grad))), grad.shape), A.sample_voxels(sc, 1, seed=3)[0]]
    for v in vox:
        v = tuple(int(i) for i in v)
        h = 1e-4 * perm[v]
        pp, pm = perm.copy(), perm.copy()
        pp[v] += h
        pm[v] -= h
        fd = (A.oracle_objective(orc, sc, pp) - A.oracle_objective(orc, sc, pm)) / (2 * h)
        assert abs(fd - grad[v]) <= 1e-6 * abs(fd), (v, fd, grad[v])


@pytest.fixture
def host():
    host_only(True)
    fdtd.set_backend("hip.float64")
    yield
    host_only(False)


def test_gradient_api_validation(host):
    if True:
        fdtd.set_backend("hip.float64")
        sc = A.TINY
        grid, design = A.build(fdtd, sc, torch.from_numpy(A.design_permittivity(sc)))
        prob = IntensityGradient(grid, design, sc["plane"], sc["steps"])
        assert prob.box == (15, 16, 6, 12, 5, 9)
        assert prob.nd == 5 * 6 * 6 and prob.np == 6 * 4
        assert _box_of(grid, (slice(-3, None), 2, slice(None))) == (17, 20, 2, 3, 0, 14)
        with pytest.raises(IndexError):
            _box_of(grid, (slice(0, 4, 2), 1, 1))
        with pytest.raises(IndexError):
            _box_of(grid, (1, 1))
        with pytest.raises(ValueError):
            IntensityGradient(grid, grid.pml_xlow, sc["plane"], 10)
        with pytest.raises(ValueError):
            IntensityGradient(grid, design, sc["plane"], 10, checkpoint="every")
        with pytest.raises(ValueError):
            IntensityGradient(grid, design, sc["plane"], 0)
        # an object overlapping the design region
        grid[9:11, 7:9, 5:7] = fdtd.Object(2.0, name="overlap")
        with pytest.raises(ValueError):
            IntensityGradient(grid, design, sc["plane"], 10)
        # an anisotropic object is transposed, but not where it overlaps an
        # absorbing one (the updates do not compose); a PML face registered
        # after a periodic boundary is a separate stage without a transpose
        g2, d2 = A.build(fdtd, sc, torch.from_numpy(A.design_permittivity(sc)))
        g2[0:2, 0:2, 0:2] = fdtd.AnisotropicObject(2.0, name="aniso")
        IntensityGradient(g2, d2, sc["plane"], 10)
        g2[1:3, 1:3, 1:3] = fdtd.AbsorbingObject(2.0, 1e4, name="absorber2")
        with pytest.raises(NotImplementedError):
            IntensityGradient(g2, d2, sc["plane"], 10)
        g5 = fdtd.Grid((16, 14, 12), grid_spacing=50e-9, courant_number=0.5)
        g5[:, :, 0] = fdtd.PeriodicBoundary(name="zp")
        g5[0:3, :, :] = fdtd.PML(name="late")
        d5 = fdtd.Object(2.0, name="d5")
        g5[6:9, 5:8, 4:7] = d5
        with pytest.raises(NotImplementedError):
            IntensityGradient(g5, d5, (12, slice(5, 8), slice(4, 7)), 10)
        # soft sources have no transposed step
        g4, d4 = A.build(fdtd, sc, torch.from_numpy(A.design_permittivity(sc)))
        g4[3, 3, 3] = SoftArbitraryPointSource(waveform_array=np.zeros(8), impedance=0.0, name="soft")
        with pytest.raises(NotImplementedError):
            IntensityGradient(g4, d4, sc["plane"], 10)
        fdtd.set_backend("hip.float32c")
        g3, d3 = A.build(fdtd, sc, torch.from_numpy(A.design_permittivity(sc)))
        with pytest.raises(NotImplementedError):
            IntensityGradient(g3, d3, sc["plane"], 10)
        fdtd.set_backend("hip.float64")


@pytest.fixture(scope="module")
def pa(golden):
    return golden("adjoint_pa")


def test_periodic_anisotropic_fixture_forward_matches_oracle(pa):
    """adjoint_pa.npz (periodic in z, a full-3x3 anisotropic object; written
    by the reference's torch autograd): its inputs regenerate, and its J and
    final E match the oracle's forward run."""
    sc = A.PA
    assert np.array_equal(pa["permittivity"], A.design_permittivity(sc))
    sim = A.build_oracle(orc, sc, A.design_permittivity(sc))
    J = 0.0
    for _ in range(sc["steps"]):
        sim.step()
        J += float((sim.E[sc["plane"]] ** 2).sum())
    assert abs(J - float(pa["J"])) <= 1e-12 * float(pa["J"])
    assert rel(sim.E, pa["E"]) <= 1e-12


def test_periodic_anisotropic_fixture_gradient_matches_oracle_central_differences(pa):
    """Two voxels (largest |grad| and a seeded one): central differences of
    the oracle's J against the reference-autograd gradient of the fixture;
    the same differences without the anisotropic object differ from it, so
    the fixture depends on that object's term."""
    sc = A.PA
    perm = A.design_permittivity(sc)
    grad = pa["grad"]
    voxels = [tuple(int(i) for i in np.unravel_index(np.argmax(np.abs(grad)), grad.shape)),
              A.sample_voxels(sc, 1, seed=4)[0]]
    h = 1e-3
    for v in voxels:
        pp, pm = perm.copy(), perm.copy()
        pp[v] += h
        pm[v] -= h
        fd = (A.oracle_objective(orc, sc, pp) - A.oracle_objective(orc, sc, pm)) / (2 * h)
        assert abs(fd - grad[v]) <= 1e-6 * np.abs(grad).max(), (v, fd, grad[v])
    plain = dict(sc, aniso=None)
    v = voxels[0]
    pp, pm = perm.copy(), perm.copy()
    pp[v] += h
    pm[v] -= h
    fd_plain = (A.oracle_objective(orc, plain, pp) - A.oracle_objective(orc, plain, pm)) / (2 * h)
    assert abs(fd_plain - grad[v]) > 1e-4 * np.abs(grad).max()


def test_design_permittivity_border_and_chain_rule(host):
    """The design's border entries keep the grid's values (the reference's
    registration, objects.py:79-90); the rest is 1/permittivity."""
    if True:
        fdtd.set_backend("hip.float64")
        sc = A.TINY
        perm = A.design_permittivity(sc)
        grid, design = A.build(fdtd, sc, torch.from_numpy(perm))
        prob = IntensityGradient(grid, design, sc["plane"], sc["steps"])
        before = design.inverse_permittivity.clone()
        newp = perm * 1.5
        prob._set_permittivity(torch.from_numpy(newp))
        inv = design.inverse_permittivity
        border = prob._border
        assert torch.equal(inv[border], before[border])
        expect = (1.0 / torch.from_numpy(newp))[..., None].expand(inv.shape)
        assert torch.allclose(inv[~border], expect[~border], rtol=0, atol=0)
        g, dcoef = prob._coefficients()
        assert dcoef is None
        S = grid.courant_number
        x, y, z = design.x, design.y, design.z
        assert torch.allclose(g[:, x, y, z], S * inv.permute(3, 0, 1, 2), rtol=1e-15, atol=0)
