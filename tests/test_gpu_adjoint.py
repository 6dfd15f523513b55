"""Gradients from the transposed HIP steps (fdtd_amd.adjoint).

  * 48x48x24: J, final E and dJ/d permittivity against the reference's own
    torch autograd (adjoint_g5.npz), both checkpoint modes
  * 36x32x12, periodic in z, with a full-3x3 anisotropic object: the same
    against the reference's autograd (adjoint_pa.npz)
  * the torch.autograd.Function wrapper: backward() equals the gradient
  * central finite differences of the HIP forward: every voxel of small
    cases with an absorbing object, with a hard plane source, and periodic
    with an anisotropic object, and 32 seeded voxels of the
    128x128x64 / 64x64x16-design / 2000-step case, in float64
  * the float32 gradient of the 128x128x64 case against the float64 one,
    error stated
  * tiles covering the (y, z) plane: the guard-free transposed H kernel
    against the guarded one, bit for bit

Tolerances: against the reference autograd 1e-10 relative L2 (the transposes
are exact; only rounding differs); against central differences 1e-4
relative (the target), with the measured values printed.
"""

import numpy as np
import pytest
import torch

import adjoint_scenarios as A
import fdtd_amd as fdtd
from fdtd_amd.adjoint import IntensityGradient, time_integrated_intensity

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _problem(sc, backend="hip.float64", checkpoint="sqrt"):
    fdtd.set_backend(backend)
    perm = A.design_permittivity(sc)
    grid, design = A.build(fdtd, sc, torch.from_numpy(perm))
    return IntensityGradient(grid, design, sc["plane"], sc["steps"], checkpoint), perm


@pytest.fixture(autouse=True)
def _fp64():
    fdtd.set_backend("hip.float64")
    yield
    fdtd.set_backend("hip.float64")


def test_g5_gradient_matches_reference_autograd(golden, monkeypatch):
    import fdtd_amd.adjoint as adjoint_mod

    f = golden("adjoint_g5")
    grads = {}
    # "sqrt2": the checkpointed mode with two record buffers, so that all but
    # the last two segments are recomputed
    for mode in ("record", "sqrt", "sqrt2"):
        _say(f"[gpu] adjoint g5 {mode}")
        monkeypatch.setattr(adjoint_mod, "RECORD_BUFFERS", 2 if mode == "sqrt2" else 16)
        prob, perm = _problem(A.G5, checkpoint=mode[:4] if mode != "record" else mode)
        assert np.array_equal(perm, f["permittivity"])
        J, g = prob.gradient(torch.from_numpy(perm))
        g = g.cpu().numpy()
        eJ = abs(J - float(f["J"])) / float(f["J"])
        eE = rel(prob.grid.E.cpu().numpy(), f["E"])
        eg = rel(g, f["grad"])
        _say(f"[gpu] adjoint g5 {mode}: J rel {eJ:.2e}, final E rel {eE:.2e}, grad rel L2 {eg:.2e} "
             f"({prob.last_info})")
        assert eJ <= 1e-12 and eE <= 1e-12
        assert eg <= 1e-10
        grads[mode] = g
    # the checkpointed reverse pass recomputes bit-identical records
    assert np.array_equal(grads["record"], grads["sqrt"])
    assert np.array_equal(grads["record"], grads["sqrt2"])


def test_periodic_anisotropic_gradient_matches_reference_autograd(golden, monkeypatch):
    """36x32x12, periodic in z (PML on x and y only), an anisotropic object
    with a per-cell full 3x3 inverse permittivity between the design region
    and the objective plane: J, final E and dJ/d permittivity against the
    reference's torch autograd (adjoint_pa.npz), both checkpoint modes."""
    import fdtd_amd.adjoint as adjoint_mod

    f = golden("adjoint_pa")
    grads = {}
    for mode in ("record", "sqrt", "sqrt2"):
        monkeypatch.setattr(adjoint_mod, "RECORD_BUFFERS", 2 if mode == "sqrt2" else 16)
        prob, perm = _problem(A.PA, checkpoint=mode[:4] if mode != "record" else mode)
        assert np.array_equal(perm, f["permittivity"])
        J, g = prob.gradient(torch.from_numpy(perm))
        g = g.cpu().numpy()
        eJ = abs(J - float(f["J"])) / float(f["J"])
        eE = rel(prob.grid.E.cpu().numpy(), f["E"])
        eg = rel(g, f["grad"])
        _say(f"[gpu] adjoint pa {mode}: J rel {eJ:.2e}, final E rel {eE:.2e}, grad rel L2 {eg:.2e} "
             f"({prob.last_info})")
        assert eJ <= 1e-12 and eE <= 1e-12
        assert eg <= 1e-10
        grads[mode] = g
    assert np.array_equal(grads["record"], grads["sqrt"])
    assert np.array_equal(grads["record"], grads["sqrt2"])


def test_small_periodic_anisotropic_case_matches_central_differences():
    """TINY with z periodic and an anisotropic object: the HIP gradient at
    every design voxel against central differences of the HIP forward."""
    sc = A.TINY_PA
    prob, perm = _problem(sc)
    J, g = prob.gradient(torch.from_numpy(perm))
    g = g.cpu().numpy()
    voxels = [tuple(int(i) for i in v) for v in np.ndindex(*perm.shape)]
    fd = _central_differences(prob, perm, voxels)
    ga = np.array([g[v] for v in voxels])
    e = rel(ga, fd)
    _say(f"[gpu] adjoint tiny_pa: {len(voxels)} voxels, rel L2 {e:.2e}")
    assert e <= 1e-6


def test_autograd_function_backward_equals_gradient():
    sc = A.TINY
    prob, perm = _problem(sc)
    J0, g0 = prob.gradient(torch.from_numpy(perm))
    p = torch.tensor(perm, dtype=torch.float64, requires_grad=True)
    J = time_integrated_intensity(p, prob.grid, prob.design, sc["plane"], sc["steps"])
    assert J.requires_grad
    (3.0 * J).backward()
    assert abs(J.detach().item() - J0) <= 1e-15 * J0
    assert torch.allclose(p.grad, 3.0 * g0.cpu(), rtol=1e-15, atol=0)


def _central_differences(prob, perm, voxels, rel_h=1e-4):
    out = []
    for v in voxels:
        h = rel_h * perm[v]
        pp, pm = perm.copy(), perm.copy()
        pp[v] += h
        pm[v] -= h
        out.append((prob.objective(torch.from_numpy(pp)) - prob.objective(torch.from_numpy(pm))) / (2 * h))
    return np.array(out)


def test_small_case_with_absorber_matches_central_differences():
    sc = A.TINY_ABSORBER
    prob, perm = _problem(sc)
    J, g = prob.gradient(torch.from_numpy(perm))
    g = g.cpu().numpy()
    voxels = [tuple(int(i) for i in v) for v in np.ndindex(*perm.shape)]
    fd = _central_differences(prob, perm, voxels)
    ga = np.array([g[v] for v in voxels])
    e = rel(ga, fd)
    emax = np.max(np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max()))
    _say(f"[gpu] adjoint tiny_absorber: {len(voxels)} voxels, rel L2 {e:.2e}, max per-voxel rel {emax:.2e}")
    assert e <= 1e-6


def test_small_case_with_plane_source_matches_central_differences():
    """A hard PlaneSource through the design region: its E and H assignments
    (sources.py:476-486) cut the dependence on the state at its cells, which
    the transposed half steps reproduce by zeroing the adjoint fields there;
    without that mask the gradient at the plane's voxels is wrong."""
    sc = A.TINY_PLANE
    prob, perm = _problem(sc)
    J, g = prob.gradient(torch.from_numpy(perm))
    g = g.cpu().numpy()
    voxels = [tuple(int(i) for i in v) for v in np.ndindex(*perm.shape)]
    fd = _central_differences(prob, perm, voxels)
    ga = np.array([g[v] for v in voxels])
    e = rel(ga, fd)
    on_plane = np.array([v[0] + sc["design"][0].start == 10 for v in voxels])
    e_plane = rel(ga[on_plane], fd[on_plane])
    _say(f"[gpu] adjoint tiny_plane: {len(voxels)} voxels, rel L2 {e:.2e}, on the plane {e_plane:.2e}")
    assert e <= 1e-6 and e_plane <= 1e-6


def test_c5_float64_gradient_matches_central_differences():
    sc = A.C5
    prob, perm = _problem(sc)
    _say("[gpu] adjoint c5 float64 gradient")
    J, g = prob.gradient(torch.from_numpy(perm))
    g = g.cpu().numpy()
    voxels = A.sample_voxels(sc, 32)
    _say("[gpu] adjoint c5 float64 central differences (64 runs)")
    fd = _central_differences(prob, perm, voxels)
    ga = np.array([g[v] for v in voxels])
    e = rel(ga, fd)
    per = np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    _say(f"[gpu] adjoint c5: J {J:.6e}, 32 voxels: rel L2 {e:.2e}, max per-voxel rel {per.max():.2e}")
    assert e <= 1e-4
    assert per.max() <= 1e-4


def test_c5_float32_gradient_error_against_float64():
    sc = A.C5
    p64, perm = _problem(sc, "hip.float64")
    J64, g64 = p64.gradient(torch.from_numpy(perm))
    del p64
    p32, _ = _problem(sc, "hip.float32")
    J32, g32 = p32.gradient(torch.from_numpy(perm))
    eJ = abs(J32 - J64) / J64
    eg = rel(g32.cpu().numpy(), g64.cpu().numpy())
    _say(f"[gpu] adjoint c5 float32 vs float64: J rel {eJ:.2e}, grad rel L2 {eg:.2e}")
    assert eJ <= 1e-4 and eg <= 1e-3


@pytest.mark.parametrize("backend", ["hip.float64", "hip.float32"])
def test_guard_free_transposed_kernel_matches_guarded(backend, monkeypatch):
    """Tiles that cover the (y, z) plane exactly run the transposed H kernel
    without its store guard; FDTD_HIP_ADJ_FULL=0 keeps the guard. Same
    gradient bit for bit."""
    sc = A.FULL_TILES
    prob, perm = _problem(sc, backend, checkpoint="record")
    J0, g0 = prob.gradient(torch.from_numpy(perm))
    monkeypatch.setenv("FDTD_HIP_ADJ_FULL", "0")
    J1, g1 = prob.gradient(torch.from_numpy(perm))
    _say(f"[gpu] adjoint full tiles {backend}: J {J0:.6e}")
    assert J0 == J1
    assert torch.equal(g0, g1)
    assert bool(torch.isfinite(g0).all()) and float(g0.abs().max()) > 0


def test_transposed_launch_shape_leaves_gradient_unchanged():
    """fdtd_adjoint_set_launch: x planes per workgroup of the transposed
    kernels change the schedule only, not the arithmetic; negative values are
    refused."""
    from fdtd_amd import _native

    sc = A.TINY_ABSORBER
    prob, perm = _problem(sc, checkpoint="record")
    J0, g0 = prob.gradient(torch.from_numpy(perm))
    orig = _native.Adjoint.set_launch
    grads = []
    try:
        for cx in (1, 3, 1000):
            _native.Adjoint.set_launch = lambda self, _cx, cx=cx: orig(self, cx)
            J, g = prob.gradient(torch.from_numpy(perm))
            assert J == J0
            grads.append(g.cpu().numpy())
    finally:
        _native.Adjoint.set_launch = orig
    for g in grads:
        assert np.array_equal(g, g0.cpu().numpy())
    adj = _native.Adjoint(prob._plan(), *prob._coefficients(), prob.grid._alloc_soa(prob.grid._E.dtype),
                          prob.grid._alloc_soa(prob.grid._H.dtype),
                          torch.zeros(_native.Adjoint.psi_size(prob._plan()), dtype=prob.grid.dtype, device="cuda"),
                          prob.design_box, prob.box, torch.zeros((3, prob.nd), dtype=torch.float64, device="cuda"))
    with pytest.raises(RuntimeError, match="cx"):
        adj.set_launch(-1)
