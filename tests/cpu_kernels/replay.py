"""Child-process driver of the host copy of libfdtd_hip (test infrastructure).

    LD_PRELOAD=<libasan.so> python tests/cpu_kernels/replay.py CASE [CASE ...]

Loads tests/cpu_kernels/_build/libfdtd_cpu.so in place of the device library
(fdtd_amd._native's handle is replaced in this process only), keeps every
grid array in host memory (fdtd_amd.backend.host_only) and runs each case
through the package's own Python layer, so the kernels see exactly the plans
the device library gets. Prints one JSON line per case: {"case", "ok",
"detail", "launches"}. Cases:

  fixture:<name>      a golden scenario written by the reference, float64:
                      fields, CPML state and detector/source traces bit for
                      bit against the fixture
  random:<seed>       a random scenario against the oracle, float64 (bit for
                      bit unless it has an anisotropic object: 1e-12)
  random_complex:<seed>  the same on the complex-field path
  compensated:<name>  a golden scenario in hip.float32c: relative L2 error of
                      the fields against the fixture <= 1e-7
  nonfinite:<name>    Grid.has_nonfinite after a golden scenario, and with a
                      NaN / an infinity written into E / H
  adjoint:<name>      the gradient of a gradient scenario (record mode)
                      against the reference-autograd fixture, 1e-10
  adjoint_full:<backend>  the transposed H kernel without its store guard
                      against the guarded form, bit for bit
  adjoint_pitch:<scenario>  a gradient scenario with padded x planes against
                      unpadded ones, bit for bit
  batch:<seed>        grids of uneven shapes stepped by fdtd_batch_run
                      (direct launches, graph replays and a persistent
                      launch) and by fdtd_batch_run_static, with the x-chunks
                      dealt to the XCDs contiguously and cyclically, bit for
                      bit against the same grids stepped alone
"""

import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import fdtd_amd as fdtd  # noqa: E402
import fdtd_amd.grid as grid_mod  # noqa: E402
from fdtd_amd import _native  # noqa: E402
from fdtd_amd.backend import host_only  # noqa: E402
import scenarios  # noqa: E402

LIB = os.path.join(HERE, "_build", "libfdtd_cpu.so")
PSI_NAMES = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}


def load():
    L = ctypes.CDLL(LIB)
    for name, (args, res) in _native._SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    L.fdtd_shim_launches.restype = ctypes.c_int64
    _native._lib = L
    host_only(True)
    # host arrays and a null stream: the host library runs launches at once
    grid_mod.Grid._stream = staticmethod(lambda: None)
    real_bd, real_ensure = grid_mod.bd, grid_mod.Grid._ensure_plan

    class _Dev:
        type = "cuda"

    class _Bd:
        device = _Dev()

        def __getattr__(self, name):
            return getattr(real_bd, name)

    def ensure_plan(self):
        grid_mod.bd = _Bd()
        try:
            return real_ensure(self)
        finally:
            grid_mod.bd = real_bd

    grid_mod.Grid._ensure_plan = ensure_plan
    # the gradient's plan and stream
    import types

    import fdtd_amd.adjoint as adj_mod

    adj_mod._require_device = lambda t: None
    torch.cuda.current_stream = lambda *a, **k: types.SimpleNamespace(cuda_stream=None)
    return L


def _np(t):
    return t.detach().cpu().numpy()


def _same(got, ref, what, bitwise=True):
    got, ref = np.asarray(got), np.asarray(ref)
    if got.shape != ref.shape:
        return f"{what}: shape {got.shape} != {ref.shape}"
    if bitwise:
        if not np.array_equal(got, ref):
            return f"{what}: not bit-identical (max |diff| {np.abs(got - ref).max():.3e})"
        return None
    n = np.linalg.norm(ref.ravel())
    r = np.linalg.norm((got - ref).ravel()) / (n if n > 0 else 1.0)
    return None if r <= 1e-12 else f"{what}: relative L2 {r:.3e}"


def case_fixture(name):
    fdtd.set_backend("hip.float64")
    g = np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"))
    sc = scenarios.get_scenario(name)
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    scenarios.run_program(grid, sc)
    errs = [_same(_np(grid.E), g["E"], "E"), _same(_np(grid.H), g["H"], "H")]
    if grid.time_steps_passed != int(g["time_steps_passed"]):
        errs.append("time_steps_passed")
    for i, c in enumerate(comps):
        p = f"out/{i}/"
        cls = str(g[f"reg/{i}/class"])
        if cls.startswith("_PML"):
            for q, nm in enumerate(PSI_NAMES[c.axis]):
                store = c._psiE if nm[0] == "E" else c._psiH
                errs.append(_same(_np(store[q % 2]), g[p + "psi_" + nm], f"psi_{nm} of face {i}"))
        elif cls in ("LineDetector", "BlockDetector"):
            errs.append(_same(_np(c.trace("E")), g[p + "E"], f"detector {i} E"))
            errs.append(_same(_np(c.trace("H")), g[p + "H"], f"detector {i} H"))
        elif cls == "CurrentDetector":
            errs.append(_same(_np(c.trace("I")), g[p + "I"], f"current detector {i}"))
        elif cls == "SoftArbitraryPointSource":
            errs.append(_same(_np(c.trace("input_voltage")), g[p + "input_voltage"], f"source {i} vin"))
            errs.append(_same(_np(c.trace("source_voltage")), g[p + "source_voltage"], f"source {i} vout"))
    return [e for e in errs if e]


def case_random(seed, make=None):
    from oracle import fdtd_oracle as orc

    fdtd.set_backend("hip.float64")
    sc = (make or scenarios.random_scenario)(int(seed))
    sim, _ = scenarios.build_oracle(orc, sc)
    scenarios.run_program(sim, sc)
    grid, comps = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    scenarios.run_program(grid, sc)
    bitwise = not any(k[0] == "AnisotropicObject" for k in sc["components"])
    errs = [_same(_np(grid.E), sim.E, "E", bitwise), _same(_np(grid.H), sim.H, "H", bitwise)]
    dets = [c for c in comps if isinstance(c, (fdtd.LineDetector, fdtd.BlockDetector))]
    for d, o in zip(dets, sim.detectors):
        errs.append(_same(_np(d.trace("E")).reshape(np.array(o.E).shape), np.array(o.E), "detector E", bitwise))
        errs.append(_same(_np(d.trace("H")).reshape(np.array(o.H).shape), np.array(o.H), "detector H", bitwise))
    return [e for e in errs if e]


def case_random_complex(seed):
    return case_random(seed, scenarios.random_complex_scenario)


def case_compensated(name):
    fdtd.set_backend("hip.float32c")
    g = np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"))
    sc = scenarios.get_scenario(name)
    grid, _ = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
    scenarios.run_program(grid, sc)
    errs = []
    for f in ("E", "H"):
        got, ref = _np(grid.full_precision(f)), g[f]
        r = np.linalg.norm((got - ref).ravel()) / np.linalg.norm(ref.ravel())
        if not r <= 1e-7:
            errs.append(f"{f}: relative L2 {r:.3e} against the float64 fixture")
    return errs


def case_adjoint(name):
    """The gradient (record mode) against the reference-autograd fixture
    adjoint_<name>.npz: J and dJ/d permittivity within 1e-10."""
    import adjoint_scenarios as A
    from fdtd_amd.adjoint import IntensityGradient

    fdtd.set_backend("hip.float64")
    sc = {"g5": A.G5, "pa": A.PA}[name]
    g = np.load(os.path.join(ROOT, "tests", "golden", f"adjoint_{name}.npz"))
    perm = A.design_permittivity(sc)
    grid, design = A.build(fdtd, sc, torch.from_numpy(perm))
    prob = IntensityGradient(grid, design, sc["plane"], sc["steps"], checkpoint="record")
    J, grad = prob.gradient(torch.from_numpy(perm))
    grad = grad.numpy()
    errs = []
    if abs(J - float(g["J"])) > 1e-10 * float(g["J"]):
        errs.append(f"J {J!r} against {float(g['J'])!r}")
    r = np.linalg.norm(grad - g["grad"]) / np.linalg.norm(g["grad"])
    if not r <= 1e-10:
        errs.append(f"gradient: relative L2 {r:.3e} against the reference autograd")
    return errs


def case_adjoint_pitch(name):
    """A gradient scenario (record mode) with unpadded x planes and with every
    plane padded to a multiple of 32 cells (FDTD_HIP_PITCH 0 and 1): the same
    J, final E and gradient bit for bit."""
    import adjoint_scenarios as A
    from fdtd_amd.adjoint import IntensityGradient

    sc = getattr(A, name)
    out, pitches = [], []
    keep = os.environ.get("FDTD_HIP_PITCH")
    try:
        for pitch in ("0", "1"):
            os.environ["FDTD_HIP_PITCH"] = pitch
            fdtd.set_backend("hip.float64")
            perm = A.design_permittivity(sc)
            grid, design = A.build(fdtd, sc, torch.from_numpy(perm))
            prob = IntensityGradient(grid, design, sc["plane"], sc["steps"], checkpoint="record")
            J, grad = prob.gradient(torch.from_numpy(perm))
            out.append((J, grad.numpy().copy(), _np(grid.E).copy()))
            pitches.append(grid._px)
    finally:
        if keep is None:
            os.environ.pop("FDTD_HIP_PITCH", None)
        else:
            os.environ["FDTD_HIP_PITCH"] = keep
    errs = []
    if pitches[0] == pitches[1]:
        errs.append(f"the scenario's planes were not padded (pitch {pitches})")
    if out[0][0] != out[1][0]:
        errs.append(f"J {out[0][0]!r} != {out[1][0]!r}")
    for what, a, b in (("gradient", out[0][1], out[1][1]), ("final E", out[0][2], out[1][2])):
        if not np.array_equal(a, b):
            errs.append(f"{what} differs with padded planes")
    return errs


def case_adjoint_full(backend):
    """The transposed H kernel without its store guard (tiles covering the
    (y, z) plane) against the guarded form (FDTD_HIP_ADJ_FULL=0): the same
    gradient bit for bit."""
    import adjoint_scenarios as A
    from fdtd_amd.adjoint import IntensityGradient

    sc = A.FULL_TILES
    out = []
    for full in (None, "0"):
        if full is None:
            os.environ.pop("FDTD_HIP_ADJ_FULL", None)
        else:
            os.environ["FDTD_HIP_ADJ_FULL"] = full
        try:
            fdtd.set_backend(backend)
            perm = A.design_permittivity(sc)
            grid, design = A.build(fdtd, sc, torch.from_numpy(perm))
            prob = IntensityGradient(grid, design, sc["plane"], sc["steps"], checkpoint="record")
            out.append(prob.gradient(torch.from_numpy(perm)))
        finally:
            os.environ.pop("FDTD_HIP_ADJ_FULL", None)
    (J0, g0), (J1, g1) = out
    errs = []
    if J0 != J1:
        errs.append(f"J {J0!r} against {J1!r}")
    if not torch.equal(g0, g1):
        errs.append("gradients differ")
    if not bool(torch.isfinite(g0).all()) or float(g0.abs().max()) == 0.0:
        errs.append("gradient not finite or all zero")
    return errs


def case_nonfinite(name):
    """Grid.has_nonfinite on a golden scenario: False after its run, True
    after a NaN or an infinity is written into one E or H value."""
    errs = []
    sc = scenarios.get_scenario(name)
    complex_ = name.startswith("complex")
    for backend in ("hip.float64", "hip.float32" if complex_ else "hip.float32c"):
        fdtd.set_backend(backend)
        grid, _ = scenarios.build_fdtd(fdtd, sc, to_array=torch.from_numpy)
        scenarios.run_program(grid, sc)
        if grid.has_nonfinite():
            errs.append(f"{backend}: finite run flagged")
        for field, bad in (("E", float("nan")), ("H", float("inf"))):
            f = getattr(grid, field)
            keep = f[1, 2, 3, 1].clone()
            f[1, 2, 3, 1] = bad
            if not grid.has_nonfinite():
                errs.append(f"{backend}: {bad} in {field} not flagged")
            f[1, 2, 3, 1] = keep
        if grid.has_nonfinite():
            errs.append(f"{backend}: flag stuck after restoring the value")
    return errs


def _batch_grids(seed):
    rs = np.random.RandomState(int(seed))
    grids = []
    for g in range(4):
        shape = tuple(int(v) for v in rs.randint(9, 23, size=3))
        grid = fdtd.Grid(shape, grid_spacing=50e-9, permittivity=rs.uniform(1.0, 4.0, size=shape + (1,)),
                         courant_number=0.5)
        for axis in range(3):
            for high in (False, True):
                if rs.rand() < 0.8:
                    key = [slice(None)] * 3
                    key[axis] = slice(-3, None) if high else slice(0, 3)
                    grid[tuple(key)] = fdtd.PML()
        if g % 2 == 1:
            grid[2:6, 3:7, 1:5] = fdtd.Object(permittivity=2.5)
            grid[6:8, 1:4, 4:8] = fdtd.AbsorbingObject(permittivity=1.5, conductivity=1e5)
        if g % 3 == 2:
            grid[1:3, 5:8, 5:8] = fdtd.AnisotropicObject(permittivity=rs.uniform(1.5, 3.0, size=(2, 3, 3)))
        p = [int(rs.randint(0, s)) for s in shape]
        grid[p[0], p[1], p[2]] = fdtd.PointSource(period=float(rs.uniform(8, 20)), pulse=bool(g % 2))
        grid.E = torch.from_numpy(rs.uniform(-1, 1, size=shape + (3,)))
        grid.H = torch.from_numpy(rs.uniform(-1, 1, size=shape + (3,)))
        grids.append(grid)
    return grids


def _state(grids):
    out = []
    for g in grids:
        out += [g._E, g._H]
        for b in g.boundaries:
            out += [b._psiE, b._psiH]
    return [t.clone() for t in out]


def _run_batch(mode, launch, grids, steps):
    """Step `grids` by `steps` through one of the batched entry points."""
    plans = [g._prepare(steps, steps, steps) for g in grids]
    if mode == "static":
        _native.batch_run_static(plans, *launch, 0, steps, None)
        return []
    b = _native.Batch(plans)
    b.set_launch(*launch)
    if mode == "graph":
        b.run_graph(0, steps, 4, None)
    else:
        b.run(0, steps, None)
    return []


def case_batch(seed):
    fdtd.set_backend("hip.float64")
    steps = 9
    alone = _batch_grids(seed)
    for g in alone:
        g.run(steps, progress_bar=False)
    want = _state(alone)
    errs = []
    runs = [(cyc, mode, launch) for cyc in ("0", "1")
            for mode, launch in (("direct", (256, 1, 0)), ("graph", (256, 2, 0)), ("persistent", (512, 1, 3)),
                                 ("persistent", (256, 3, 16)), ("static", (256, 2, 0)))]
    try:
        for cyc, mode, launch in runs:
            # small batches deal contiguously unless told otherwise
            os.environ["FDTD_HIP_CYCLIC"] = cyc
            grids = _batch_grids(seed)
            errs += _run_batch(mode, launch, grids, steps)
            for i, (x, y) in enumerate(zip(_state(grids), want)):
                if not torch.equal(x, y):
                    errs.append(f"cyclic={cyc} {mode} {launch}: state tensor {i} differs")
    finally:
        os.environ.pop("FDTD_HIP_CYCLIC", None)
    return errs


CASES = {"fixture": case_fixture, "random": case_random, "random_complex": case_random_complex,
         "compensated": case_compensated, "batch": case_batch, "adjoint": case_adjoint,
         "adjoint_full": case_adjoint_full, "adjoint_pitch": case_adjoint_pitch, "nonfinite": case_nonfinite}


def main(argv):
    L = load()
    for case in argv:
        kind, _, arg = case.partition(":")
        before = L.fdtd_shim_launches()
        try:
            errs = CASES[kind](arg)
        except Exception as e:  # reported to the parent test
            errs = [f"{type(e).__name__}: {e}"]
        print(json.dumps({"case": case, "ok": not errs, "detail": errs,
                          "launches": L.fdtd_shim_launches() - before}), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
