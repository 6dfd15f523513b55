"""Parity of the HIP path (hip.float64 / hip.float32) with the reference.

  * every golden scenario: the device run against the fixture the reference
    wrote (fields, detector traces, PML state), bit-exact in float64
  * random scenarios: the device run against the numpy oracle, bit-exact
  * the 256^3 configuration with 16-cell PML: a few steps against the oracle,
    plus size-independent properties at full size (exact linearity)
  * float32: error against the float64 fixture, stated and bounded
  * the launch-shape overrides (tiles, x planes per workgroup, chunk
    dealing, workgroup size, non-temporal accesses): bit-identical results

Tolerances: the float64 bar is <= 1e-12 relative L2 on all fields; the
kernels are written to reproduce numpy's rounding sequence, so the tests
also demand bitwise equality.
"""

import math
import os

import numpy as np
import pytest
import torch

import fdtd_amd as fdtd
import scenarios
from oracle import fdtd_oracle as orc

pytestmark = pytest.mark.gpu

SCENARIOS = [sc["name"] for sc in scenarios.all_scenarios()]
PSI_NAMES = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}


def _say(msg):
    print(msg, flush=True)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    d = np.linalg.norm((a.astype(dt) - b.astype(dt)).ravel())
    n = np.linalg.norm(b.astype(dt).ravel())
    return d / n if n > 0 else d


def _build(sc):
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    return grid, comps


def _np(t):
    return t.detach().cpu().numpy()


def _assert_same(got, ref, what, bitwise=True):
    r = rel_l2(got, ref)
    assert r <= 1e-12, f"{what}: relative L2 {r:.3e}"
    if bitwise:
        assert np.array_equal(got, ref), f"{what}: not bit-identical (rel L2 {r:.3e}, max {np.abs(got - ref).max():.3e})"


@pytest.fixture(autouse=True)
def _fp64():
    fdtd.set_backend("hip.float64")
    yield
    fdtd.set_backend("hip.float64")


@pytest.mark.parametrize("name", SCENARIOS)
def test_float64_matches_reference_fixture(golden, name):
    _say(f"[gpu] fixture scenario {name}")
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, comps = _build(sc)
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    assert grid.time_steps_passed == int(g["time_steps_passed"])
    _assert_same(_np(grid.E), g["E"], "E")
    _assert_same(_np(grid.H), g["H"], "H")
    for i, c in enumerate(comps):
        p = f"out/{i}/"
        cls = str(g[f"reg/{i}/class"])
        if cls.startswith("_PML"):
            for q, nm in enumerate(PSI_NAMES[c.axis]):
                store = c._psiE if nm[0] == "E" else c._psiH
                _assert_same(_np(store[q % 2]), g[p + "psi_" + nm], f"psi_{nm} of face {i}")
        elif cls in ("LineDetector", "BlockDetector"):
            _assert_same(_np(c.trace("E")), g[p + "E"], f"detector {i} E")
            _assert_same(_np(c.trace("H")), g[p + "H"], f"detector {i} H")
        elif cls == "CurrentDetector":
            _assert_same(_np(c.trace("I")), g[p + "I"], f"current detector {i}")
        elif cls == "SoftArbitraryPointSource":
            _assert_same(_np(c.trace("input_voltage")), g[p + "input_voltage"], f"source {i} input voltage")
            _assert_same(_np(c.trace("source_voltage")), g[p + "source_voltage"], f"source {i} source voltage")
            _assert_same(_np(c.current_detector.trace("I")), g[p + "I"], f"source {i} current")


@pytest.mark.parametrize("seed", range(16))
def test_float64_random_scenarios_match_oracle(seed):
    sc = scenarios.random_scenario(seed)
    _say(f"[gpu] random scenario {seed}: shape {sc['shape']}, {[c[0] for c in sc['components']]}")
    sim, ocomps = scenarios.build_oracle(orc, sc)
    scenarios.run_program(sim, sc)
    grid, comps = _build(sc)
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    # numpy's einsum association is machine-dependent; anisotropic cases at tolerance
    bitwise = not any(k[0] == "AnisotropicObject" for k in sc["components"])
    _assert_same(_np(grid.E), sim.E, "E", bitwise)
    _assert_same(_np(grid.H), sim.H, "H", bitwise)
    dets = [c for c in comps if isinstance(c, (fdtd.LineDetector, fdtd.BlockDetector))]
    odets = sim.detectors
    for d, o in zip(dets, odets):
        _assert_same(_np(d.trace("E")).reshape(np.array(o.E).shape), np.array(o.E), "detector E", bitwise)
        _assert_same(_np(d.trace("H")).reshape(np.array(o.H).shape), np.array(o.H), "detector H", bitwise)


def test_save_data_writes_reference_readings(golden, tmp_path, monkeypatch):
    """Grid.save_simulation() + save_data() (grid.py:407-439, 490-514) on the
    G2 fixture run: fdtd_output/fdtd_output_<time> (<name>)/detector_readings.npz
    with keys "<name> (E)" / "<name> (H)", line readings (steps, cells, 3) and
    block readings (steps, nx, ny, nz, 3) -- the 3- and 5-axis shapes that
    plot_detection and dB_map_2D require (visualization.py:358, 420) -- equal
    to the reference's own readings bit for bit."""
    g = golden("g2")
    sc = scenarios.get_scenario("g2")
    grid, comps = _build(sc)
    scenarios.run_program(grid, sc)
    monkeypatch.chdir(tmp_path)
    with pytest.raises(Exception, match="Save location not initialized"):
        grid.save_data()
    folder = grid.save_simulation("readings")
    assert os.path.dirname(folder) == str(tmp_path / "fdtd_output")
    assert os.path.basename(folder).startswith("fdtd_output_") and folder.endswith(" (readings)")
    grid.save_data()
    dets = [(i, c) for i, c in enumerate(comps) if isinstance(c, (fdtd.LineDetector, fdtd.BlockDetector))]
    assert len(dets) == 3
    with np.load(os.path.join(folder, "detector_readings.npz")) as f:
        assert sorted(f.files) == sorted(f"{c.name} ({F})" for _, c in dets for F in "EH")
        for i, c in dets:
            for F in "EH":
                got = f[f"{c.name} ({F})"]
                assert got.ndim == (3 if isinstance(c, fdtd.LineDetector) else 5)
                assert got.shape[0] == grid.time_steps_passed
                _assert_same(got, g[f"out/{i}/{F}"], f"saved {c.name} ({F})")


LAUNCH_OVERRIDES = ["FDTD_HIP_FLAT", "FDTD_HIP_CX", "FDTD_HIP_CYCLIC", "FDTD_HIP_BLOCK", "FDTD_HIP_NT"]
LAUNCH_SHAPES = [
    dict(FDTD_HIP_FLAT="0"),
    dict(FDTD_HIP_CX="3"),
    dict(FDTD_HIP_CX="2", FDTD_HIP_CYCLIC="1"),
    dict(FDTD_HIP_CYCLIC="0", FDTD_HIP_BLOCK="512"),
    dict(FDTD_HIP_BLOCK="1024", FDTD_HIP_FLAT="0"),
    dict(FDTD_HIP_NT="0"),
]


@pytest.mark.parametrize("backend", ["hip.float64", "hip.float32", "hip.float32c"])
def test_launch_shape_overrides_leave_results_bit_identical(backend, monkeypatch):
    """The launch-shape overrides -- row or flat tiles, x planes per
    workgroup, chunk dealing, workgroup size, non-temporal accesses -- change
    the schedule only: on the all-feature fixture scenario (every object
    kind, point/line/plane sources, detectors, random initial fields) the
    fields and detector traces equal the default launch's bit for bit."""
    sc = scenarios.get_scenario("g2")

    def run(env):
        for k in LAUNCH_OVERRIDES:
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        fdtd.set_backend(backend)
        grid, comps = _build(sc)
        scenarios.run_program(grid, sc)
        torch.cuda.synchronize()
        out = [_np(grid.full_precision("E") if backend == "hip.float32c" else grid.E),
               _np(grid.full_precision("H") if backend == "hip.float32c" else grid.H)]
        for c in comps:
            if isinstance(c, fdtd.LineDetector):
                out += [_np(c.trace("E")), _np(c.trace("H"))]
        return out

    ref = run({})
    for env in LAUNCH_SHAPES:
        got = run(env)
        assert len(got) == len(ref)
        for i, (a, b) in enumerate(zip(got, ref)):
            assert np.array_equal(a, b), f"{backend} {env}: array {i} differs from the default launch"
    _say(f"[gpu] {backend}: {len(LAUNCH_SHAPES)} launch shapes bit-identical to the default")


@pytest.mark.parametrize("nz", [70, 100, 130])
def test_uneven_z_rows_match_oracle(nz):
    """Grids whose z edge is not a multiple of the 64-lane z-rows (70, 100,
    130 cells: partly filled last rows), with PML faces on every side,
    objects, a source, a z-line detector and random initial fields,
    bit-exact against the oracle."""
    s = slice
    sc = dict(
        name=f"uneven_z_{nz}", shape=(9, 7, nz), grid_spacing=50e-9, courant_number=0.5,
        permittivity=1.5, permeability=1.0,
        components=[
            ("PML", dict(), (s(0, 2), s(None), s(None)), {}),
            ("PML", dict(), (s(-3, None), s(None), s(None)), {}),
            ("PML", dict(), (s(None), s(0, 2), s(None)), {}),
            ("PML", dict(), (s(None), s(-2, None), s(None)), {}),
            ("PML", dict(), (s(None), s(None), s(0, 5)), {}),
            ("PML", dict(), (s(None), s(None), s(-6, None)), {}),
            ("Object", dict(permittivity=3.0), (s(3, 6), s(2, 5), s(nz // 3, nz // 2)), {}),
            ("AbsorbingObject", dict(permittivity=2.0, conductivity=3e4), (s(2, 4), s(3, 6), s(nz - 20, nz - 9)), {}),
            ("PointSource", dict(period=12.0), (4, 3, nz // 2 + 1), {}),
            ("LineDetector", dict(name="zline"), (4, 3, s(None)), {}),
        ],
        program=[("run", 25)],
    )
    rs = np.random.RandomState(nz)
    sc["init"] = {"E": rs.uniform(-1, 1, size=(9, 7, nz, 3)), "H": rs.uniform(-1, 1, size=(9, 7, nz, 3))}
    sim, _ = scenarios.build_oracle(orc, sc)
    scenarios.run_program(sim, sc)
    grid, comps = _build(sc)
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    _assert_same(_np(grid.E), sim.E, "E")
    _assert_same(_np(grid.H), sim.H, "H")
    det = [c for c in comps if isinstance(c, fdtd.LineDetector)][0]
    _assert_same(_np(det.trace("E")).reshape(np.array(sim.detectors[0].E).shape), np.array(sim.detectors[0].E),
                 "detector E")


def test_step_by_step_equals_run():
    """grid.step() in a loop, grid.run() and manual half steps give identical bits."""
    sc = scenarios.get_scenario("g2")
    out = []
    for mode in ("run", "step", "half"):
        grid, comps = _build(sc)
        for _ in range(2):
            if mode == "run":
                grid.run(7, progress_bar=False)
            elif mode == "step":
                for _ in range(7):
                    grid.step()
            else:
                for _ in range(7):
                    grid.update_E()
                    grid.update_H()
                    grid.time_steps_passed += 1
        torch.cuda.synchronize()
        out.append((_np(grid.E), _np(grid.H), _np(comps[-1].trace("E"))))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


def test_material_edits_take_effect():
    """Writes to grid materials and rebinding an object's array are honoured
    on the next step, as the reference reads them every step."""
    sc = scenarios.get_scenario("g2")
    sim, ocomps = scenarios.build_oracle(orc, sc)
    grid, comps = _build(sc)
    for g_ in (grid, sim):
        g_.run(5) if g_ is sim else g_.run(5, progress_bar=False)
    sim.inverse_permittivity[3:9, 2:6, 4:10] *= 0.5
    grid.inverse_permittivity[3:9, 2:6, 4:10] *= 0.5
    new_inv = np.asarray(ocomps[6].inverse_permittivity) * 0.75
    ocomps[6].inverse_permittivity = new_inv.copy()
    comps[6].inverse_permittivity = torch.from_numpy(new_inv.copy())
    sim.run(5)
    grid.run(5, progress_bar=False)
    torch.cuda.synchronize()
    _assert_same(_np(grid.E), sim.E, "E", bitwise=False)
    _assert_same(_np(grid.H), sim.H, "H", bitwise=False)


def _c2_grid(mod_grid, n=256, pml=16, amplitude=1.0, oracle=False):
    """The 256^3 benchmark configuration: 24 seeded dielectric spheres,
    16-cell PML on all six faces, pulsed 1550 nm point source at the centre."""
    from bench import sphere_permittivity
    eps = sphere_permittivity(n)
    kw = dict(grid_spacing=50e-9, permittivity=eps, courant_number=3 ** -0.5)
    grid = mod_grid.Sim((n, n, n), **kw) if oracle else fdtd.Grid((n, n, n), **kw)
    keys = [(slice(0, pml), slice(None), slice(None)), (slice(-pml, None), slice(None), slice(None)),
            (slice(None), slice(0, pml), slice(None)), (slice(None), slice(-pml, None), slice(None)),
            (slice(None), slice(None), slice(0, pml)), (slice(None), slice(None), slice(-pml, None))]
    src = dict(period=1550e-9 / 299792458.0, pulse=True, amplitude=amplitude)
    if oracle:
        for k in keys:
            grid.add("PML", k)
        grid.add("PointSource", (n // 2, n // 2, n // 2), **src)
    else:
        for k in keys:
            grid[k] = fdtd.PML()
        grid[n // 2, n // 2, n // 2] = fdtd.PointSource(**src)
    return grid


def test_full_size_256_matches_oracle_first_steps():
    """The benchmark configuration itself: 3 steps from random fields against
    the oracle at full size (bit-exact)."""
    _say("[gpu] 256^3 parity vs oracle (3 steps)")
    rs = np.random.RandomState(5)
    e0 = rs.randn(256, 256, 256, 3) * 0.01
    h0 = rs.randn(256, 256, 256, 3) * 0.01
    sim = _c2_grid(orc, oracle=True)
    sim.E[...] = e0
    sim.H[...] = h0
    grid = _c2_grid(None)
    grid.E = torch.from_numpy(e0)
    grid.H = torch.from_numpy(h0)
    del e0, h0
    sim.run(3)
    grid.run(3, progress_bar=False)
    torch.cuda.synchronize()
    _assert_same(_np(grid.E), sim.E, "E")
    _assert_same(_np(grid.H), sim.H, "H")


def test_full_size_256_linearity_is_exact():
    """Doubling the source amplitude doubles every field bit-for-bit (every
    operation of the step is linear, and scaling by 2 is exact)."""
    _say("[gpu] 256^3 linearity, 40 steps")
    a = _c2_grid(None, amplitude=1.0)
    a.run(40, progress_bar=False)
    Ea = a._E.clone()
    del a
    b = _c2_grid(None, amplitude=2.0)
    b.run(40, progress_bar=False)
    torch.cuda.synchronize()
    assert torch.count_nonzero(Ea).item() > 0
    assert torch.equal(b._E, 2 * Ea)


@pytest.mark.parametrize("name", ["g2", "c1", "periodic_y_last"])
def test_float32_error_against_float64_reference(golden, name):
    """float32 fields: relative L2 error against the float64 reference stated
    (printed) and bounded by 1e-4."""
    fdtd.set_backend("hip.float32")
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=lambda a: torch.from_numpy(np.asarray(a, np.float32)))
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    eE = rel_l2(_np(grid.E), g["E"])
    eH = rel_l2(_np(grid.H), g["H"])
    _say(f"[gpu] float32 {name}: relative L2 error E {eE:.2e} H {eH:.2e}")
    assert eE < 1e-4 and eH < 1e-4


@pytest.mark.parametrize("seed", range(16, 28))
def test_float64_folded_point_sources_match_oracle(seed):
    """Scenarios without detectors, line or plane sources: point sources are
    then added inside the E kernel (when no periodic boundary follows)."""
    sc = scenarios.random_scenario(seed)
    sc["components"] = [c for c in sc["components"]
                        if c[0] not in ("LineDetector", "BlockDetector", "LineSource", "PlaneSource")]
    sc["components"].append(("PointSource", dict(period=7, amplitude=0.3), (1, 1, 0), {}))
    _say(f"[gpu] folded-source scenario {seed}: shape {sc['shape']}, {[c[0] for c in sc['components']]}")
    sim, _ = scenarios.build_oracle(orc, sc)
    scenarios.run_program(sim, sc)
    grid, _ = _build(sc)
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    bitwise = not any(k[0] == "AnisotropicObject" for k in sc["components"])
    _assert_same(_np(grid.E), sim.E, "E", bitwise)
    _assert_same(_np(grid.H), sim.H, "H", bitwise)


@pytest.mark.parametrize("name", ["c1", "manual", "oned", "float_keys", "periodic_z_last"])
def test_material_palette_is_bit_invisible_in_float64(golden, name):
    """Reading the grid materials through the palette gives the same bits as
    reading the per-cell arrays (the palette stores the same float64 values);
    grids with more than 256 distinct materials fall back to the arrays."""
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, _ = _build(sc)
    grid.material_palette = True
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    _assert_same(_np(grid.E), g["E"], "E")
    _assert_same(_np(grid.H), g["H"], "H")


@pytest.mark.parametrize("name", ["g2", "c1", "manual", "periodic_x_first", "periodic_y_last", "oned"])
def test_compensated_float32_error_against_float64_reference(golden, name):
    """hip.float32c (float32 storage + 8 compensation bits, float64
    arithmetic): relative L2 error of the decoded fields against the float64
    reference stated and bounded by 1e-7 -- two orders below plain float32.
    Values the host writes through the float32 views (initial fields, the
    hard source of "manual") carry float32 rounding; the rest keeps 32
    significant bits."""
    fdtd.set_backend("hip.float32c")
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    eE = rel_l2(_np(grid.full_precision("E")), g["E"])
    eH = rel_l2(_np(grid.full_precision("H")), g["H"])
    _say(f"[gpu] float32c {name}: relative L2 error E {eE:.2e} H {eH:.2e}")
    assert eE < 1e-7 and eH < 1e-7
    for i, c in enumerate(comps):
        cls = str(g[f"reg/{i}/class"])
        if cls in ("LineDetector", "BlockDetector"):
            assert rel_l2(_np(c.trace("E")), g[f"out/{i}/E"]) < 1e-6


def test_compensated_words_are_the_record_slots():
    """hip.float32c fields are (Nx, Ny, Nz, 4) float32 records: the float32 E/H
    views are [..., :3] of them and the compensation words [..., 3];
    fdtd_plan_set_compensation refuses words anywhere else."""
    from fdtd_amd import _native

    fdtd.set_backend("hip.float32c")
    grid = fdtd.Grid((6, 5, 4), grid_spacing=50e-9, courant_number=0.5)
    assert grid._E_rec.shape == (6, 5, 4, 4) and grid._E_rec.stride() == (grid._px * 4, 16, 4, 1)
    assert grid.E.data_ptr() == grid._E_rec.data_ptr() and grid.E.stride() == (grid._px * 4, 16, 4, 1)
    assert grid._E_lo.data_ptr() == grid._E_rec.data_ptr() + 12
    plan = _native.Plan(_native.F32C, (6, 5, 4), grid.courant_number, grid._E, grid._H, grid._inv_eps, grid._inv_mu,
                        pitch=grid._px)
    with pytest.raises(RuntimeError, match="4th float"):
        plan.set_compensation(torch.zeros((6, 5, 4), dtype=torch.int32, device=grid._E.device), grid._H_lo)
    plan.set_compensation(grid._E_lo, grid._H_lo)
    fdtd.set_backend("hip.float64")


def test_frequency_routines_match_reference(golden):
    """FrequencyRoutines on the device (torch.fft, complex128) after the
    soft-source run, against what the reference's routines computed on its
    own run (numpy FFT): equal to 1e-12 relative."""
    g = golden("soft_source")
    sc = scenarios.get_scenario("soft_source")
    grid, comps = _build(sc)
    scenarios.run_program(grid, sc)
    src = comps[8]
    f, z = fdtd.FrequencyRoutines(grid, src).impedance()
    assert rel_l2(_np(f), g["out/8/impedance_freqs"]) < 1e-15
    assert rel_l2(_np(z), g["out/8/impedance"]) < 1e-12
    hist = src.trace("source_voltage").reshape(-1)
    f, sp = fdtd.FrequencyRoutines(grid, hist).FFT()
    assert rel_l2(_np(f), g["out/8/fft_freqs"]) < 1e-15
    assert rel_l2(_np(sp), g["out/8/fft"]) < 1e-12
    assert fdtd.FrequencyRoutines(fdtd.Grid((8, 8, 8), grid_spacing=1e-3), hist).FFT() == ([], [])


def test_complex_fields_float32_error_against_reference(golden):
    """The complex-field path in complex64 (hip.float32) against the
    reference's complex128 run: error stated, bounded at 1e-5 relative."""
    g = golden("complex_object")
    fdtd.set_backend("hip.float32")
    sc = scenarios.get_scenario("complex_object")
    grid, comps = _build(sc)
    assert grid.complex and grid.E.dtype == torch.complex64
    scenarios.run_program(grid, sc)
    eE, eH = rel_l2(_np(grid.E), g["E"]), rel_l2(_np(grid.H), g["H"])
    _say(f"[gpu] complex64 vs reference complex128: E rel {eE:.2e}, H rel {eH:.2e}")
    assert eE <= 1e-5 and eH <= 1e-5


def test_complex_fields_pml_registered_late_fails_like_the_reference():
    """A PML registered after the fields became complex keeps real CPML
    state; the reference's first step then fails casting complex into it."""
    grid = fdtd.Grid((12, 10, 8), grid_spacing=50e-9)
    grid[0:2, :, :] = fdtd.PML()
    grid[4:7, 3:6, 2:5] = fdtd.Object(permittivity=2.0 + 0.1j)
    grid[-2:, :, :] = fdtd.PML()
    with pytest.raises(TypeError):
        grid.step()


@pytest.mark.parametrize("seed", range(6))
def test_complex_random_scenarios_match_oracle(seed):
    """Random configurations on the complex-field path (complex128), objects
    of every kind (absorbing: numpy's complex division; anisotropic: its
    complex einsum), against the oracle's numpy complex arithmetic, bit for
    bit (anisotropic cases within 1e-12: einsum's association on the host
    running the oracle decides there); the energy plane of visualize()
    computed on the device against numpy's."""
    sc = scenarios.random_complex_scenario(seed)
    _say(f"[gpu] random complex scenario {seed}: shape {sc['shape']}, {[c[0] for c in sc['components']]}")
    sim, ocomps = scenarios.build_oracle(orc, sc)
    scenarios.run_program(sim, sc)
    grid, comps = _build(sc)
    assert grid.complex
    scenarios.run_program(grid, sc)
    torch.cuda.synchronize()
    bitwise = not any(k[0] == "AnisotropicObject" for k in sc["components"])
    _assert_same(_np(grid.E), sim.E, "E", bitwise)
    _assert_same(_np(grid.H), sim.H, "H", bitwise)
    dets = [c for c in comps if isinstance(c, (fdtd.LineDetector, fdtd.BlockDetector))]
    for d, o in zip(dets, sim.detectors):
        _assert_same(_np(d.trace("E")).reshape(np.array(o.E).shape), np.array(o.E), "detector E", bitwise)
        _assert_same(_np(d.trace("H")).reshape(np.array(o.H).shape), np.array(o.H), "detector H", bitwise)
    z = sc["shape"][2] // 2
    if sc["shape"][0] > 1 and sc["shape"][1] > 1:
        ref = np.sum(sim.E ** 2 + sim.H ** 2, -1)[:, :, z]
        np.testing.assert_allclose(_np(grid.grid_energy(z=z)), ref, rtol=1e-15, atol=1e-15 * np.abs(ref).max())


@pytest.mark.parametrize("backend", ["hip.float64", "hip.float32", "hip.float32c"])
def test_nonfinite_check_flags_a_diverged_run(backend):
    """Grid.has_nonfinite: False after a normal run; True after a NaN is
    written into E, and after a run whose courant_number was raised above
    the 3D limit after construction (the reference checks the limit only at
    construction, grid.py:123-127, and then diverges silently)."""
    fdtd.set_backend(backend)
    rs = np.random.RandomState(3)
    grid = fdtd.Grid((24, 20, 16), grid_spacing=50e-9, courant_number=0.5)
    grid[0:4, :, :] = fdtd.PML()
    grid[12, 10, 8] = fdtd.PointSource(period=10)
    grid.run(40, progress_bar=False)
    assert not grid.has_nonfinite()
    keep = grid.E[5, 6, 7, 2].clone()
    grid.E[5, 6, 7, 2] = float("nan")
    assert grid.has_nonfinite()
    grid.E[5, 6, 7, 2] = keep
    assert not grid.has_nonfinite()
    grid.E = torch.from_numpy(rs.uniform(-1, 1, size=(24, 20, 16, 3)))
    grid.courant_number = 1.5
    grid.run(1500, progress_bar=False)
    _say(f"[gpu] {backend}: courant 1.5 after construction, 1500 steps: max |E| "
         f"{float(grid.E.abs().max()) if not grid.has_nonfinite() else float('nan')}")
    assert grid.has_nonfinite()
