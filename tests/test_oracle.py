"""The numpy oracle against the reference's known-answer vectors, against the
golden fixtures written by the reference itself, and against closed forms."""

import math

import numpy as np
import pytest

import scenarios
from oracle import fdtd_oracle as orc

SCENARIOS = [sc["name"] for sc in scenarios.all_scenarios()]


def test_curl_known_answers(golden):
    g = golden("curl_golden")
    # reference tests/test_grid.py:47-164 compare with rel 1e-6; the oracle is exact
    np.testing.assert_allclose(orc.curl_E(g["E_in"]), g["curl_E"], rtol=1e-7, atol=1e-8)
    np.testing.assert_allclose(orc.curl_H(g["H_in"]), g["curl_H"], rtol=1e-7, atol=1e-8)


def test_curl_planted_mistake_is_caught(golden):
    """A sign slip in one term of the curl fails the known-answer check."""
    g = golden("curl_golden")
    bad = orc.curl_E(g["E_in"])
    bad[:, :-1, :, 0] -= 2 * (g["E_in"][:, 1:, :, 2] - g["E_in"][:, :-1, :, 2])
    assert not np.allclose(bad, g["curl_E"], rtol=1e-6)


def _run(name):
    sc = scenarios.get_scenario(name)
    sim, comps = scenarios.build_oracle(orc, sc)
    return sc, sim, comps


@pytest.mark.parametrize("name", SCENARIOS)
def test_oracle_matches_reference_fixture(golden, name):
    g = golden(name)
    sc, sim, comps = _run(name)
    assert scenarios.input_checksum(sc) == float(g["input_checksum"])
    assert [sim.Nx, sim.Ny, sim.Nz] == list(g["reg/Nxyz"])
    assert sim.courant_number == float(g["reg/courant_number"])
    assert sim.time_step == float(g["reg/time_step"])
    if "reg/grid_inverse_permittivity" in g.files:
        assert np.array_equal(sim.inverse_permittivity, g["reg/grid_inverse_permittivity"])
    for i, c in enumerate(comps):
        p = f"reg/{i}/"
        cls = str(g[p + "class"])
        if cls.startswith("_PML"):
            ax = c.axis
            assert c.t == int(g[p + "thickness"])
            for nm in ("bE", "cE", "bH", "cH"):
                arr = getattr(c, nm)
                idx = [0, 0, 0]
                idx[ax] = slice(None)
                assert np.array_equal(arr[tuple(idx) + (ax,)], g[p + nm]), nm
        elif cls in ("Object", "AbsorbingObject", "AnisotropicObject"):
            assert [c.x.start, c.x.stop] == list(g[p + "x"])
            if cls != "AnisotropicObject":
                assert np.array_equal(c.inverse_permittivity, g[p + "inverse_permittivity"])
            if cls == "AbsorbingObject":
                assert np.array_equal(c.absorption_factor, g[p + "absorption_factor"])
        elif cls == "LineSource":
            assert list(c.x) == list(g[p + "x"]) and list(c.y) == list(g[p + "y"]) and list(c.z) == list(g[p + "z"])
            assert c.period == int(g[p + "period"])
            assert np.array_equal(c.profile, g[p + "profile"])
        elif cls == "PointSource":
            assert [c.x, c.y, c.z] == list(g[p + "xyz"]) and c.period == int(g[p + "period"])
        elif cls == "PlaneSource":
            assert [c.Epol, c.Hpol] == list(g[p + "pol"])
        elif cls in ("LineDetector", "BlockDetector", "CurrentDetector"):
            assert list(c.x) == list(g[p + "x"]) and list(c.y) == list(g[p + "y"]) and list(c.z) == list(g[p + "z"])
        elif cls == "SoftArbitraryPointSource":
            assert [c.x, c.y, c.z] == list(g[p + "xyz"])

    scenarios.run_program(sim, sc)
    assert sim.time_steps_passed == int(g["time_steps_passed"])
    # bit-exact except where the reference itself uses a reduction whose
    # association order is numpy's (the anisotropic einsum)
    has_aniso = any(k[0] == "AnisotropicObject" for k in sc["components"])
    if has_aniso:
        np.testing.assert_allclose(sim.E, g["E"], rtol=0, atol=1e-13 * np.abs(g["E"]).max())
        np.testing.assert_allclose(sim.H, g["H"], rtol=0, atol=1e-13 * np.abs(g["H"]).max())
    else:
        assert np.array_equal(sim.E, g["E"])
        assert np.array_equal(sim.H, g["H"])
    for i, c in enumerate(comps):
        p = f"out/{i}/"
        cls = str(g[f"reg/{i}/class"])
        if cls.startswith("_PML"):
            names = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}[c.axis]
            for nm in names:
                arr = (c.psi_E if nm[0] == "E" else c.psi_H)["xyz".index(nm[1])][..., c.axis]
                ref = g[p + "psi_" + nm]
                if has_aniso:
                    np.testing.assert_allclose(arr, ref, rtol=0, atol=1e-13 * max(np.abs(ref).max(), 1e-30))
                else:
                    assert np.array_equal(arr, ref), nm
        elif cls in ("LineDetector", "BlockDetector"):
            e = np.array(c.E)
            h = np.array(c.H)
            tol = 1e-13 * max(np.abs(g[p + "E"]).max(), 1e-30) if has_aniso else 0
            np.testing.assert_allclose(e, g[p + "E"], rtol=0, atol=tol)
            np.testing.assert_allclose(h, g[p + "H"], rtol=0, atol=tol)
        elif cls == "CurrentDetector":
            assert np.array_equal(np.array(c.I), g[p + "I"])
        elif cls == "SoftArbitraryPointSource":
            assert np.array_equal(np.array(c.input_voltage), g[p + "input_voltage"])
            assert np.array_equal(np.array(c.source_voltage), g[p + "source_voltage"])
            assert np.array_equal(np.array(c.current_detector.I), g[p + "I"])


def test_courant_and_time_step_closed_forms():
    for shape, d in (((3, 1, 1), 1), ((3, 3, 1), 2), ((3, 3, 3), 3)):
        s = orc.Sim(shape, grid_spacing=1e-7)
        assert s.courant_number == 0.99 * float(d) ** -0.5
        assert s.time_step == s.courant_number * 1e-7 / 299792458.0
    with pytest.raises(ValueError):
        orc.Sim((3, 3, 3), courant_number=0.6)
    s = orc.Sim((10e-9, 10e-9, 3), grid_spacing=5e-9)
    assert (s.Nx, s.Ny, s.Nz) == (2, 2, 3)


def test_magic_time_step_1d_shifts_one_cell_per_step():
    """In 1D with courant number 1 the Yee scheme has no dispersion: with H
    at t=1/2 set to -f(i) (a wave travelling towards +x in these units), Ez
    moves exactly one cell per step."""
    n = 64
    s = orc.Sim((n, 1, 1), courant_number=1.0)
    x = np.arange(n)
    pulse = np.exp(-0.5 * ((x - 20) / 3.0) ** 2)
    s.E[:, 0, 0, 2] = pulse
    s.H[:, 0, 0, 1] = -pulse
    for _ in range(10):
        s.step()
    # away from the x=0 edge (where the 2e-10 tail meets the grid border)
    np.testing.assert_allclose(s.E[15:60, 0, 0, 2], pulse[5:50], rtol=0, atol=1e-12)


def test_uniform_absorbing_decay_closed_form():
    """With zero curl every absorbing cell scales E by (1-f)/(1+f) per step."""
    s = orc.Sim((6, 6, 6), courant_number=0.5, grid_spacing=1e-7)
    sigma = 3e4
    s.add("AbsorbingObject", (slice(None), slice(None), slice(None)), permittivity=2.0, conductivity=sigma)
    s.E[...] = 1.0
    f = 0.5 * 0.5 * 0.5 * sigma * 1e-7 * orc.ETA0
    r = (1 - f) / (1 + f)
    s.update_E()
    # interior cells (uniform E -> zero curl) decay by exactly r
    np.testing.assert_allclose(s.E[2:4, 2:4, 2:4], r, rtol=1e-15)


def test_pml_sigma_profile_closed_form():
    """sigma(v) = 40 v^3 / (t+1)^4, low face E-side at t-1/2 ... 1/2."""
    s = orc.Sim((10, 10, 10), courant_number=0.5)
    f = s.add("PML", (slice(0, 4), slice(None), slice(None)))
    v = np.array([3.5, 2.5, 1.5, 0.5])
    assert np.allclose(f.sigmaE[:, 0, 0, 0], 40 * v ** 3 / 5 ** 4, rtol=1e-15)
    # b = exp(-(sigma + a) * S); c = (b - 1) sigma / (sigma + a)
    b = np.exp(-(40 * v ** 3 / 625 + 1e-8) * 0.5)
    assert np.allclose(f.bE[:, 0, 0, 0], b, rtol=1e-15)
    # the H-side profile ends with a zero-conductivity cell
    assert f.sigmaH[-1, 0, 0, 0] == 0.0 and f.cH[-1, 0, 0, 0] == 0.0


def test_frequency_routines_match_reference_fixture(golden):
    """The impedance and FFT the reference's FrequencyRoutines computed on
    the soft-source run, restated with numpy (fourier.py:52-266)."""
    g = golden("soft_source")
    dt = float(g["reg/time_step"])
    steps = int(g["time_steps_passed"])
    f, z = orc.fr_impedance(g["out/8/source_voltage"], g["out/8/I"], dt, steps)
    assert np.array_equal(f, g["out/8/impedance_freqs"]) and np.array_equal(z, g["out/8/impedance"])
    f, sp = orc.fr_fft(g["out/8/source_voltage"], dt)
    assert np.array_equal(f, g["out/8/fft_freqs"]) and np.array_equal(sp, g["out/8/fft"])


def test_complex_path_pins_numpys_complex_product(golden, monkeypatch):
    """The complex fixture (written by the reference) fixes numpy's complex
    product, re = fma(ar, br, -(ai*bi)), im = fma(ar, bi, ai*br): the oracle
    (numpy) reproduces it bit for bit, and the textbook product without
    fused multiply-adds -- a plausible restatement -- does not."""
    g = golden("complex_object")
    sc, sim, comps = _run("complex_object")
    assert np.iscomplexobj(sim.E) and np.iscomplexobj(comps[0].psi_E[1])
    scenarios.run_program(sim, sc)
    assert np.array_equal(sim.E, g["E"]) and np.array_equal(sim.H, g["H"])

    def textbook(self, curl):
        s = self.sim
        loc = (self.x, self.y, self.z)
        k = s.courant_number * self.inverse_permittivity
        c = curl[loc]
        s.E[loc] = s.E[loc] + ((k.real * c.real - k.imag * c.imag) + 1j * (k.real * c.imag + k.imag * c.real))

    monkeypatch.setattr(orc._Object, "update_E", textbook)
    sc, sim, comps = _run("complex_object")
    scenarios.run_program(sim, sc)
    assert not np.array_equal(sim.E, g["E"])
    assert np.abs(sim.E - g["E"]).max() < 1e-12 * np.abs(g["E"]).max()


def _planted_absorbing(ratio, quotient):
    """An absorbing-object update with the E factor and the curl term's
    division computed by the given functions."""
    def update(self, curl):
        s = self.sim
        loc = (self.x, self.y, self.z)
        if self.kind != "absorbing":
            return orig_update(self, curl)
        f = self.absorption_factor
        s.E[loc] *= ratio(f)
        s.E[loc] += quotient(s.courant_number * self.inverse_permittivity * curl[loc], 1 + f)
    orig_update = orc._Object.update_E
    return update


def test_complex_absorbing_fixture_pins_numpys_division(golden, monkeypatch):
    """The complex_absorbing fixture (written by the reference) fixes numpy's
    complex division (Smith's method) for complex absorption factors and a
    real division for an absorbing object whose arrays are real on complex
    fields: the oracle (numpy) matches it bit for bit, and three plausible
    restatements do not -- multiplying by the reciprocal, dividing the real
    object's factor in complex arithmetic, and einsum products with fused
    multiply-adds (numpy's ufunc product) in the anisotropic mat-vec."""
    g = golden("complex_absorbing")
    sc, sim, comps = _run("complex_absorbing")
    kinds = [(o.kind, np.iscomplexobj(getattr(o, "absorption_factor", 0j))) for o in sim.objects]
    assert ("absorbing", True) in kinds and ("absorbing", False) in kinds
    scenarios.run_program(sim, sc)
    assert np.array_equal(sim.E, g["E"]) and np.array_equal(sim.H, g["H"])

    plants = {
        "reciprocal": _planted_absorbing(lambda f: (1 - f) * (1 / (1 + f)), lambda a, b: a * (1 / b)),
        "complex ratio of a real factor": _planted_absorbing(
            lambda f: (1 - f.astype(np.complex128)) / (1 + f.astype(np.complex128)), lambda a, b: a / b),
    }
    for what, fn in plants.items():
        monkeypatch.setattr(orc._Object, "update_E", fn)
        sc, sim, comps = _run("complex_absorbing")
        scenarios.run_program(sim, sc)
        assert not np.array_equal(sim.E, g["E"]), what
        assert np.abs(sim.E - g["E"]).max() < 1e-10 * np.abs(g["E"]).max(), what
        monkeypatch.undo()

    real_einsum = np.einsum

    def fused_einsum(spec, m, v):  # ufunc products (fused multiply-add), same order
        if spec != "ijk,ikl->ijl" or not np.iscomplexobj(m):
            return real_einsum(spec, m, v)
        v = v.astype(m.dtype)
        return ((m[:, :, 0] * v[:, 0]) + (m[:, :, 1] * v[:, 1]) + (m[:, :, 2] * v[:, 2]))[..., None]

    monkeypatch.setattr(orc.np, "einsum", fused_einsum)
    sc, sim, comps = _run("complex_absorbing")
    scenarios.run_program(sim, sc)
    assert not np.array_equal(sim.E, g["E"])
    assert np.abs(sim.E - g["E"]).max() < 1e-10 * np.abs(g["E"]).max()


def test_complex_permittivity_with_periodic_boundary_fails_like_the_reference():
    sim = orc.Sim((8, 8, 8), grid_spacing=50e-9)
    sim.add("PeriodicBoundary", (slice(None), slice(None), 0))
    with pytest.raises(AttributeError):
        sim.add("Object", (slice(2, 5), slice(2, 5), slice(2, 5)), permittivity=2.0 + 0.1j)
