"""Host-side registration logic of fdtd_amd (no GPU): arrays are kept in host
memory and every resolved index, profile and coefficient is compared with
what the reference computed for the same scenario (tests/golden/*.npz)."""

import numpy as np
import pytest
import torch

import fdtd_amd as fdtd
import scenarios
from fdtd_amd.backend import host_only
from fdtd_amd.boundaries import DomainBorderPML, pml_profiles

SCENARIOS = [sc["name"] for sc in scenarios.all_scenarios()]


@pytest.fixture(autouse=True)
def _host_arrays():
    host_only(True)
    fdtd.set_backend("hip.float64")
    yield
    host_only(False)


def _np(t):
    return t.detach().cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


@pytest.mark.parametrize("name", SCENARIOS)
def test_registration_matches_reference(golden, name):
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    assert [grid.Nx, grid.Ny, grid.Nz] == list(g["reg/Nxyz"])
    assert grid.courant_number == float(g["reg/courant_number"])
    assert grid.time_step == float(g["reg/time_step"])
    if "reg/grid_inverse_permittivity" in g.files:
        assert np.array_equal(_np(grid.inverse_permittivity), g["reg/grid_inverse_permittivity"])
    for i, c in enumerate(comps):
        p = f"reg/{i}/"
        cls = str(g[p + "class"])
        assert type(c).__name__ == cls
        if cls.startswith("_PML"):
            assert c.thickness == int(g[p + "thickness"])
            for nm in ("bE", "cE", "bH", "cH"):
                assert np.array_equal(getattr(c, "_" + nm), g[p + nm]), nm
                full = _np(getattr(c, nm))
                other = [k for k in range(3) if k != c.axis][0]
                assert full.flat[other] == float(g[p + nm + "_other"]) or (
                    full.flat[other] == 0 and float(g[p + nm + "_other"]) == 0)
        elif cls in ("Object", "AbsorbingObject", "AnisotropicObject"):
            assert [c.x.start, c.x.stop] == list(g[p + "x"])
            assert [c.y.start, c.y.stop] == list(g[p + "y"])
            assert [c.z.start, c.z.stop] == list(g[p + "z"])
            if cls != "AnisotropicObject":
                assert np.array_equal(_np(c.inverse_permittivity), g[p + "inverse_permittivity"])
            if cls == "AbsorbingObject":
                assert np.array_equal(_np(c.absorption_factor), g[p + "absorption_factor"])
        elif cls == "PointSource":
            assert [c.x, c.y, c.z] == list(g[p + "xyz"])
            assert c.period == int(g[p + "period"])
        elif cls == "LineSource":
            assert (c.x, c.y, c.z) == (list(g[p + "x"]), list(g[p + "y"]), list(g[p + "z"]))
            assert c.period == int(g[p + "period"])
            assert np.array_equal(c.profile, g[p + "profile"])
        elif cls == "PlaneSource":
            assert [c.x.start, c.x.stop, c.y.start, c.y.stop, c.z.start, c.z.stop] == \
                list(g[p + "x"]) + list(g[p + "y"]) + list(g[p + "z"])
            assert [c._Epol, c._Hpol] == list(g[p + "pol"])
            assert c.period == int(g[p + "period"])
        elif cls in ("LineDetector", "BlockDetector"):
            assert (list(c.x), list(c.y), list(c.z)) == (list(g[p + "x"]), list(g[p + "y"]), list(g[p + "z"]))


@pytest.mark.parametrize("name", SCENARIOS)
def test_waveform_tables_match_oracle(name):
    """The per-step source scalars the device table holds equal what the oracle
    adds each step (same Python math expressions)."""
    from oracle import fdtd_oracle as orc

    sc = scenarios.get_scenario(name)
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    sim, ocomps = scenarios.build_oracle(orc, sc)
    for c, o in zip(comps, ocomps):
        if not hasattr(c, "_waveform"):
            continue
        for q in range(0, 400, 7):
            w = c._waveform(q)
            if not hasattr(o, "kind"):  # soft source: the waveform array, 0 past its end
                assert w == (float(o.waveform[q]) if q < o.waveform.shape[0] else 0.0)
                continue
            if o.kind == "point":
                if o.pulse:
                    h = o._pulse_value(q)
                    ref = o.amplitude * h if h is not None else 0.0
                else:
                    ref = o.amplitude * o._sine(q)
            elif o.kind == "line":
                h = o._pulse_value(q) if o.pulse else o._sine(q)
                ref = h if h is not None else 0.0
            else:
                ref = o._sine(q)
            assert w == ref


def test_pml_profiles_closed_form():
    bE, cE, bH, cH, sE, sH = pml_profiles(4, False, 1e-8, 1.0, 0.5)
    v = np.array([3.5, 2.5, 1.5, 0.5])
    np.testing.assert_allclose(sE, 40 * v ** 3 / 625, rtol=1e-15)
    np.testing.assert_allclose(bE, np.exp(-(sE + 1e-8) * 0.5), rtol=1e-15)
    assert sH[-1] == 0 and cH[-1] == 0
    bE2, _, _, _, sE2, _ = pml_profiles(4, True, 1e-8, 1.0, 0.5)
    np.testing.assert_array_equal(sE2, sE[::-1])


# ---- reference-style API tests (tests/test_grid.py, test_boundaries.py, test_sources.py)
def test_grid_shape_of_ints_and_floats():
    assert fdtd.Grid(shape=(3, 3, 3)).shape == (3, 3, 3)
    assert fdtd.Grid(shape=(10.0e-9, 10.0e-9, 10.0e-9), grid_spacing=5.0e-9).shape == (2, 2, 2)
    assert fdtd.Grid(shape=(10.0e-9, 10.0e-9, 3), grid_spacing=5.0e-9).shape == (2, 2, 3)


def test_default_courant_numbers():
    assert fdtd.Grid(shape=(3, 1, 1)).courant_number == pytest.approx(1.0, rel=0.02)
    assert fdtd.Grid(shape=(3, 3, 1)).courant_number == pytest.approx(0.5 ** 0.5, rel=0.02)
    assert fdtd.Grid(shape=(3, 3, 3)).courant_number == pytest.approx((1 / 3) ** 0.5, rel=0.02)
    with pytest.raises(ValueError):
        fdtd.Grid(shape=(3, 3, 3), courant_number=0.6)


def _grid():
    return fdtd.Grid(shape=(10, 10, 10), grid_spacing=100e-9)


def test_boundary_registration_rules():
    grid = _grid()
    pb = fdtd.PeriodicBoundary(name="periodic_boundary")
    grid[0, :, :] = pb
    assert pb in grid.boundaries and type(pb).__name__ == "_PeriodicBoundaryX"
    with pytest.raises(IndexError):
        _grid()[0:2, :, :] = fdtd.PeriodicBoundary()
    with pytest.raises(IndexError):
        _grid()[2, :, :] = fdtd.PeriodicBoundary()
    grid = _grid()
    pml = fdtd.PML(name="PML")
    grid[0:3, :, :] = pml
    assert pml in grid.boundaries and pml.thickness == 3
    with pytest.raises(IndexError):
        _grid()[2:4, :, :] = fdtd.PML()
    with pytest.raises(AttributeError):
        grid[0:2, :, :] = fdtd.PML()
    with pytest.raises(ValueError):
        grid[-3:, :, :] = fdtd.PML(name="PML")


def test_domain_border_pml():
    grid = _grid()
    with pytest.raises(IndexError):
        DomainBorderPML(grid, grid.Nx // 2 + 1)
    DomainBorderPML(_grid(), 3)
    grid = _grid()
    DomainBorderPML(grid, 3)
    assert sorted(type(b).__name__ for b in grid.boundaries) == sorted(
        ["_PMLZlow", "_PMLZhigh", "_PMLXlow", "_PMLXhigh", "_PMLYlow", "_PMLYhigh"])


def test_plane_source_polarization():
    with pytest.raises(ValueError):
        _grid()[0, :, :] = fdtd.PlaneSource(polarization="x")
    with pytest.raises(ValueError):
        _grid()[:, 0, :] = fdtd.PlaneSource(polarization="y")
    with pytest.raises(ValueError):
        _grid()[:, :, 0] = fdtd.PlaneSource(polarization="z")
    table = [((0, slice(None), slice(None)), "y", 1, 2), ((0, slice(None), slice(None)), "z", 2, 1),
             ((slice(None), 0, slice(None)), "x", 0, 2), ((slice(None), 0, slice(None)), "z", 2, 0),
             ((slice(None), slice(None), 0), "x", 0, 1), ((slice(None), slice(None), 0), "y", 1, 0)]
    for key, pol, e, h in table:
        ps = fdtd.PlaneSource(polarization=pol)
        _grid()[key] = ps
        assert (ps._Epol, ps._Hpol) == (e, h)


def test_point_and_line_source_errors():
    with pytest.raises(ValueError):
        _grid()[1:3, 2, 2] = fdtd.PointSource()
    with pytest.raises(ValueError):
        _grid()[2, 2, 2:3] = fdtd.LineSource()


def test_set_backend_names():
    fdtd.set_backend("hip")
    assert fdtd.backend.float == torch.float64
    fdtd.set_backend("hip.float32")
    assert fdtd.backend.float == torch.float32
    fdtd.set_backend("hip.float64")
    with pytest.raises(ValueError):
        fdtd.set_backend("numpy")
    with pytest.raises(ValueError):
        fdtd.set_backend("hip.float16")


def test_stepping_refuses_host_arrays():
    grid = _grid()
    with pytest.raises(RuntimeError):
        grid.step()


def test_complex_permittivity_promotes_fields_and_cpml_state():
    """objects.py:51-52 / grid.py:353-356: an Object with a complex
    permittivity makes E, H and the PML state registered before it complex;
    a later absorbing object with a real permittivity keeps real arrays (its
    factor (1 - f)/(1 + f) is then a real division), one with a complex
    permittivity gets a complex absorption factor; periodic boundaries are
    refused as in the reference."""
    grid = fdtd.Grid((12, 10, 8), grid_spacing=50e-9)
    grid[0:2, :, :] = fdtd.PML()
    grid[4:7, 3:6, 2:5] = fdtd.Object(permittivity=2.0 + 0.1j)
    assert grid.complex and grid.E.dtype == torch.complex128
    assert grid.boundaries[0]._psiE.dtype == torch.complex128
    inv = grid.objects[0].inverse_permittivity
    assert inv.dtype == torch.complex128
    assert inv[0, 0, 0, 0].item() == np.complex128(1) / np.complex128(2.0 + 0.1j)
    assert grid.objects[0]._real_arrays is False
    grid[8:10, 3:6, 2:5] = fdtd.AbsorbingObject(permittivity=1.5, conductivity=1e3)
    real = grid.objects[1]
    assert real._real_arrays and not real.absorption_factor.is_complex()
    grid[1:3, 6:9, 2:5] = fdtd.AbsorbingObject(permittivity=1.5 + 0.02j, conductivity=1e3)
    lossy = grid.objects[2]
    assert not lossy._real_arrays and lossy.absorption_factor.dtype == torch.complex128
    f = 0.5 * grid.courant_number * lossy._inv_host * 1e3 * grid.grid_spacing * fdtd.constants.eta0
    assert np.array_equal(lossy.absorption_factor.numpy(), f)
    g2 = fdtd.Grid((12, 10, 8), grid_spacing=50e-9)
    g2[:, :, 0] = fdtd.PeriodicBoundary()
    with pytest.raises(AttributeError):
        g2[4:7, 3:6, 2:5] = fdtd.Object(permittivity=2.0 + 0.1j)


@pytest.mark.parametrize("name", ["g2", "complex_object"])
def test_grid_energy_is_the_visualize_reduction(golden, name):
    """grid_energy(x|y|z) equals the reference's visualize() energy,
    sum(E**2 + H**2, -1) sliced (transposed for y), on the fixture fields:
    bit for bit for real fields; for complex ones within 1e-15 relative
    (numpy squares complex numbers with a fused multiply-add, torch does not)."""
    g = golden(name)
    sc = scenarios.get_scenario(name)
    grid, _ = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    grid.E = torch.from_numpy(g["E"])
    grid.H = torch.from_numpy(g["H"])
    ref = np.sum(g["E"] ** 2 + g["H"] ** 2, -1)
    for got, want in ((grid.grid_energy(x=3), ref[3, :, :]), (grid.grid_energy(y=4), ref[:, 4, :].T),
                      (grid.grid_energy(z=5), ref[:, :, 5])):
        if np.iscomplexobj(want):
            np.testing.assert_allclose(_np(got), want, rtol=1e-15, atol=1e-15 * np.abs(want).max())
        else:
            assert np.array_equal(_np(got), want)
    with pytest.raises(ValueError):
        grid.grid_energy(x=1, z=2)
    with pytest.raises(ValueError):
        grid.grid_energy()


def _check_reverse_schedule(ops, nseg, nbuf):
    """Replay a checkpointed reverse sweep's issue order against its record
    buffers: every transposed segment finds its own record in its buffer,
    no recomputation overwrites a buffer before the transposed steps of the
    segment it holds were issued, the forward-recorded segments are never
    recomputed and the others exactly once, transposed steps latest first."""
    holder = {s % nbuf: s for s in range(max(0, nseg - nbuf), nseg)}
    done, recomputed, order = set(), [], []
    for op, s in ops:
        b = s % nbuf
        if op == "recompute":
            assert holder.get(b) is None or holder[b] in done, f"buffer {b} of segment {holder.get(b)} overwritten"
            holder[b] = s
            recomputed.append(s)
        else:
            assert holder.get(b) == s, f"segment {s} transposed without its record"
            done.add(s)
            order.append(s)
    assert order == list(range(nseg - 1, -1, -1))
    assert sorted(recomputed) == list(range(max(0, nseg - nbuf)))


def test_checkpointed_reverse_schedule():
    """fdtd_amd.adjoint.reverse_schedule for 1-80 segments and 2-8 buffers;
    a schedule that recomputes one segment further ahead than the buffers
    allow is caught."""
    from fdtd_amd.adjoint import reverse_schedule

    for nseg in range(1, 81):
        for nbuf in range(2, 9):
            _check_reverse_schedule(reverse_schedule(nseg, nbuf), nseg, nbuf)

    def too_far_ahead(nseg, nbuf):
        ops, pending = [], nseg - nbuf - 1
        for s in range(nseg - 1, -1, -1):
            while pending >= 0 and pending >= s - nbuf:
                ops.append(("recompute", pending))
                pending -= 1
            ops.append(("adjoint", s))
        return ops

    with pytest.raises(AssertionError, match="overwritten"):
        _check_reverse_schedule(too_far_ahead(20, 4), 20, 4)


def test_bench_algorithmic_bytes():
    """bench.algorithmic_bytes_per_half_step at 256^3 with 16-cell PML: the
    per-step figures of the measurement plan (24 values per cell plus 8 per
    face slab cell: 3,623,878,656 B in float64, 1,811,939,328 B in float32);
    the palette form reads a 1-byte material index instead of 3 values; the
    compensated form 9 values of 5 bytes and the index, float64 psi."""
    import bench

    hb = bench.algorithmic_bytes_per_half_step
    assert 2 * hb(256, 16, 8) == 3_623_878_656
    assert 2 * hb(256, 16, 4) == 1_811_939_328
    assert hb(256, 16, 8) - hb(256, 16, 8, palette=True) == 256 ** 3 * (3 * 8 - 1)
    assert hb(256, 16, 4, compensated=True) == 256 ** 3 * 46 + 6 * 16 * 256 * 256 * 4 * 8


def test_checkpointed_schedule(monkeypatch):
    """IntensityGradient._sqrt_schedule: segments of ceil(sqrt(T)) to twice
    that many steps and 2 to RECORD_BUFFERS record buffers, within the memory
    of the classic schedule's checkpoints plus a quarter; among those the pair
    that records the most steps (checked by brute force), the classic
    schedule with two buffers when none fits, SEGMENT_STEPS fixing the length."""
    from types import SimpleNamespace

    import fdtd_amd.adjoint as A

    E = torch.zeros((3, 128, 128, 64), dtype=torch.float32)
    psi = [torch.zeros(2 * 12 * 128 * 64, dtype=torch.float32) for _ in range(12)]
    state = sum(t.numel() * 4 for t in [E, E] + psi)
    T, K0 = 2000, 45
    budget = 1.25 * K0 * state

    def sched(nd):
        fake = SimpleNamespace(grid=SimpleNamespace(_E=E), steps=T, nd=nd, np=256, _state=lambda: [E, E] + psi)
        return A.IntensityGradient._sqrt_schedule(fake)

    def memory(K, B, nd):
        nseg = -(-T // K)
        return max(nseg - B, 0) * state + min(B, nseg) * K * 3 * (nd + 256) * 4

    # the differentiation workload (64x64x16 design): small records
    nd = 64 * 64 * 16
    K, B = sched(nd)
    assert K0 <= K <= 2 * K0 and 2 <= B <= A.RECORD_BUFFERS
    assert memory(K, B, nd) <= budget
    assert -(-T // K) > B  # still checkpointed and recomputed
    best = max(min(B2 * K2, T) for B2 in range(2, A.RECORD_BUFFERS + 1) for K2 in range(K0, 2 * K0 + 1)
               if memory(K2, B2, nd) <= budget)
    assert min(B * K, T) == best
    assert T - B * K <= T // 3
    # a larger design: fewer buffers fit
    K2, B2 = sched(128 * 128 * 16)
    assert memory(K2, B2, 128 * 128 * 16) <= budget and 2 <= B2 < B
    # a whole-grid design: records as large as half a state, nothing fits
    assert sched(128 * 128 * 64) == (K0, 2)
    # a fixed segment length
    monkeypatch.setattr(A, "SEGMENT_STEPS", K0)
    K3, B3 = sched(nd)
    assert K3 == K0 and 2 <= B3 <= A.RECORD_BUFFERS and memory(K3, B3, nd) <= budget


def test_compensated_records_and_host_writes():
    """hip.float32c storage (host arrays): one (Nx, Ny, Nz, 4) float32 record
    per cell; a whole-field assignment keeps 32 significant bits (the float
    and the next 8 mantissa bits, rounded at bit 21 of the double), a write
    into grid.E[...] changes the written floats only and leaves every other
    cell's 40-bit value untouched."""
    fdtd.set_backend("hip.float32c")
    try:
        grid = fdtd.Grid((5, 4, 3), grid_spacing=50e-9, courant_number=0.5)
        assert grid._E_rec.shape == (5, 4, 3, 4)
        rs = np.random.RandomState(7)
        v = rs.uniform(-1, 1, size=(5, 4, 3, 3))
        grid.E = torch.from_numpy(v)
        # the 32-bit rounding of v, from its bit pattern (numpy, independent of the grid code)
        bits = v.view(np.int64)
        want = ((bits + (1 << 20)) & ~np.int64((1 << 21) - 1)).view(np.float64)
        got = grid.full_precision("E").numpy()
        assert np.array_equal(got, want)
        assert np.abs(got - v).max() <= np.abs(v).max() * 2.0 ** -32
        # the float part: the top 24 significant bits of that rounding (exact in float32)
        hi = (want.view(np.int64) & ~np.int64((1 << 29) - 1)).view(np.float64)
        assert np.array_equal(grid.E.numpy(), hi.astype(np.float32))
        grid.E[1:3, :, 2, 0] = 0.25
        after = grid.full_precision("E").numpy()
        mask = np.zeros(v.shape, bool)
        mask[1:3, :, 2, 0] = True
        assert np.array_equal(after[~mask], want[~mask])
        assert np.all(np.abs(after[mask] - 0.25) < 0.25 * 2.0 ** -23)
        # the words are the records' 4th floats
        assert grid._E_lo.data_ptr() == grid._E_rec.data_ptr() + 12
    finally:
        fdtd.set_backend("hip.float64")


def test_compensated_material_assignment_keeps_fields_and_precision():
    """On hip.float32c, assigning a grid material writes the float64 material
    array exactly and leaves the fields' compensation words alone (only E and
    H are stored as float + 8 bits)."""
    fdtd.set_backend("hip.float32c")
    try:
        grid = fdtd.Grid((4, 5, 6), grid_spacing=50e-9, courant_number=0.5)
        grid.H = torch.full((4, 5, 6, 3), 1.0 + 2.0 ** -30, dtype=torch.float64)
        grid.E = torch.full((4, 5, 6, 3), -3.0 - 2.0 ** -28, dtype=torch.float64)
        h_before = grid.full_precision("H").clone()
        e_before = grid.full_precision("E").clone()
        lo_before = (grid._E_lo.clone(), grid._H_lo.clone())
        assert float(h_before[0, 0, 0, 0]) == 1.0 + 2.0 ** -30
        grid.inverse_permittivity = 1.0 / 3.0
        grid.inverse_permeability = torch.full((4, 5, 6, 3), 0.7, dtype=torch.float64)
        assert grid.inverse_permittivity.dtype == torch.float64
        assert torch.all(grid.inverse_permittivity == 1.0 / 3.0)
        assert torch.all(grid.inverse_permeability == 0.7)
        assert torch.equal(grid._E_lo, lo_before[0]) and torch.equal(grid._H_lo, lo_before[1])
        assert torch.equal(grid.full_precision("H"), h_before)
        assert torch.equal(grid.full_precision("E"), e_before)
    finally:
        fdtd.set_backend("hip.float64")


def test_padded_plane_layout(monkeypatch):
    """x planes of 4096 cells and more are padded to a multiple of 32 cells
    (FDTD_HIP_PITCH: 0 never, 1 every plane); the (Nx, Ny, Nz, 3) views and
    the cell indices the kernels get follow the padded storage, and writes
    through the views never touch the padding."""
    from fdtd_amd.grid import plane_pitch

    monkeypatch.delenv("FDTD_HIP_PITCH", raising=False)
    assert plane_pitch(250, 250) == 62528 and plane_pitch(256, 256) == 65536
    assert plane_pitch(18, 14) == 252 and plane_pitch(100, 100) == 10016
    monkeypatch.setenv("FDTD_HIP_PITCH", "1")
    assert plane_pitch(18, 14) == 256 and plane_pitch(8, 4) == 32 and plane_pitch(97, 1) == 128
    monkeypatch.setenv("FDTD_HIP_PITCH", "0")
    assert plane_pitch(250, 250) == 62500
    monkeypatch.setenv("FDTD_HIP_PITCH", "1")
    for backend in ("hip.float64", "hip.float32c"):
        fdtd.set_backend(backend)
        g = fdtd.Grid((5, 6, 7), grid_spacing=50e-9, courant_number=0.5)
        assert g._px == 64
        assert g._cell(2, 3, 4) == 2 * 64 + 3 * 7 + 4 and g._cell(-1, -1, -1) == 4 * 64 + 5 * 7 + 6
        rng = np.random.default_rng(0)
        E = rng.uniform(-1, 1, (5, 6, 7, 3))
        g.E = torch.from_numpy(E)
        if g.compensated:
            store = g._E_rec._base if g._E_rec._base is not None else g._E_rec
            assert store.shape == (5, 64, 4)
            assert not store[:, 42:].any()  # the padding of every plane stays zero
            assert np.abs(_np(g.full_precision("E")) - E).max() <= 2.0 ** -31
        else:
            store = g._E._base
            assert store.shape == (3, 5, 64) and not store[:, :, 42:].any()
            assert np.array_equal(_np(g.E), E)
        g.inverse_permittivity = torch.full((5, 6, 7, 3), 0.25, dtype=torch.float64)
        assert np.all(_np(g.inverse_permittivity) == 0.25)
    fdtd.set_backend("hip.float64")
