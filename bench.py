"""Throughput of the fused FDTD step on one MI355X.

Workload (the 256^3 plain-method configuration): Grid 256x256x256, spacing
50 nm, courant 1/sqrt(3), 24 seeded dielectric spheres (radius 8-32 cells,
eps uniform in [1, 12]) rasterised into the grid permittivity, 16-cell PML on
all six faces, pulsed 1550 nm PointSource at the centre. A "step" is one full
Yee step (E then H half step) over the whole grid.

    python bench.py [--steps K] [--warmup W] [--dtype float64|float32]

Prints one JSON line: Mcells/s over K timed steps (HIP events bracketing
grid.run(K) on the current stream, after W warm-up steps), the HBM roofline of
the dominant kernel (algorithmic bytes per launch / average launch time,
timed live with HIP events on the launch stream), and a CPU baseline: the
numpy oracle on the same configuration, a bounded number of steps, one core.
"""

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "Mcells/s + % HBM roofline (256³ PML, fp64 & fp32); adjoint/forward time ratio"


def sphere_permittivity(n=256, count=24, seed=0):
    """Relative permittivity with `count` seeded spheres (centres uniform in
    the grid, integer radius in [8, 32], eps uniform in [1, 12]; later spheres
    overwrite earlier ones)."""
    rng = np.random.default_rng(seed)
    eps = np.ones((n, n, n))
    ax = np.arange(n)
    for _ in range(count):
        c = rng.uniform(0, n, size=3)
        r = int(rng.integers(8, 33))
        e = float(rng.uniform(1.0, 12.0))
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, n)
        x = ax[lo[0]:hi[0], None, None]
        y = ax[None, lo[1]:hi[1], None]
        z = ax[None, None, lo[2]:hi[2]]
        mask = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
        eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]][mask] = e
    return eps


PML_KEYS = lambda t: [(slice(0, t), slice(None), slice(None)), (slice(-t, None), slice(None), slice(None)),
                      (slice(None), slice(0, t), slice(None)), (slice(None), slice(-t, None), slice(None)),
                      (slice(None), slice(None), slice(0, t)), (slice(None), slice(None), slice(-t, None))]


PALETTE = "auto"  # --palette: how the grid material is read (auto | on | off)


def build_grid(fdtd, n=256, pml=16, palette=None):
    eps = sphere_permittivity(n)
    grid = fdtd.Grid((n, n, n), grid_spacing=50e-9, permittivity=eps, courant_number=3 ** -0.5)
    grid.material_palette = {"auto": "auto", "on": True, "off": False}[palette or PALETTE]
    for k in PML_KEYS(pml):
        grid[k] = fdtd.PML()
    grid[n // 2, n // 2, n // 2] = fdtd.PointSource(period=1550e-9 / 299792458.0, pulse=True)
    return grid


def algorithmic_bytes_per_half_step(n, pml, itemsize, compensated=False, palette=False):
    """12 values per cell (read field, other field, material; write field) plus
    4 per (face, slab cell) (read and write the two psi components). The
    compensated form stores a field value in 5 bytes (float + int8), reads the
    material as a 1-byte palette index and keeps psi in float64; with a
    palette alone the material is that 1-byte index."""
    cells = n ** 3
    if compensated:
        slab = pml * n * n
        return cells * (9 * 5 + 1) + 6 * slab * 4 * 8
    slab = pml * n * n
    if palette:
        return cells * (9 * itemsize + 1) + 6 * slab * 4 * itemsize
    return cells * 12 * itemsize + 6 * slab * 4 * itemsize


def copy_bandwidth(torch, nbytes=1 << 30, reps=10):
    """Plain streaming device copy of `nbytes` (libfdtd_hip's k_copy): the
    practical HBM bandwidth next to the 8 TB/s peak, and the reference for
    scaling byte counters."""
    from fdtd_amd import _native

    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream().cuda_stream
    _native.bench_copy(dst, src, nbytes, 2, stream)
    ms = _native.bench_copy(dst, src, nbytes, reps, stream)
    return 2 * nbytes * reps / (ms / 1e3) / 1e9


def cpu_baseline(n, pml, steps):
    """The numpy oracle (single-threaded ufuncs) on the same workload: `steps`
    timed steps after one untimed one; the value is the median step, the
    spread the fastest and slowest."""
    from oracle import fdtd_oracle as orc

    sim = orc.Sim((n, n, n), grid_spacing=50e-9, permittivity=sphere_permittivity(n), courant_number=3 ** -0.5)
    for k in PML_KEYS(pml):
        sim.add("PML", k)
    sim.add("PointSource", (n // 2, n // 2, n // 2), period=1550e-9 / 299792458.0, pulse=True)
    sim.step()  # first step outside the timing (page faults of fresh arrays)
    per = []
    for _ in range(steps):
        t0 = time.perf_counter()
        sim.step()
        per.append(time.perf_counter() - t0)
    rate = lambda dt: n ** 3 / dt / 1e6
    cpu = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=rate(float(np.median(per))), unit="Mcells/s", cores=1, kind="port",
                spread={"min": round(rate(max(per)), 3), "max": round(rate(min(per)), 3), "steps": steps},
                sample=f"{steps} steps of the {n}^3 workload (after 1 untimed step), median step; numpy oracle, "
                       f"ufuncs single-threaded; {os.cpu_count()} host cores visible ({cpu})")


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "k_main_256cube_pml16_hbm_traffic.md")


def profiled_traffic(kernel, ctype, n, pml):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary (counters scaled by a plain copy of known size), for
    the configuration that summary was taken on; None otherwise."""
    if (n, pml) != (256, 16) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip().strip("`") for c in line.split("|")]
        if len(cells) > 5 and cells[1] == f"{kernel}<{ctype}>":
            return int(float(cells[4]))
    return None


WINDOWS = 3  # timed windows of K steps each; the median is reported


def field_errors(got, ref):
    """Error of a run's final fields against the float64 run of the same
    workload (device tensors of every E and H value, float64): relative L2
    norm, max-norm relative error (max |difference| / max |float64 value|)
    and the pointwise relative error with a floor of 1e-3 max |float64 value|."""
    d = (got - ref).abs()
    mx = ref.abs().max()
    return {"rel_l2": float((got - ref).norm() / ref.norm()), "max_norm_rel": float(d.max() / mx),
            "pointwise_rel_floor_1e-3max": float((d / ref.abs().clamp_min(1e-3 * mx)).max())}


def measure(torch, fdtd, dtype, n, pml, steps, warmup, kernel_reps, keep_fields=False, palette=None):
    """Mcells/s over `steps` timed steps (median of WINDOWS windows) and the
    roofline of the dominant kernel; with keep_fields, also the final E and H
    (float64, decoded) for error figures. `palette` ("on"/"off") overrides
    --palette for this run."""
    fdtd.set_backend(f"hip.{dtype}")
    grid = build_grid(fdtd, n, pml, palette)
    grid.run(warmup, progress_bar=False)
    pal = grid._palette_wanted() and grid._material_palette() is not None  # palette in use (<= 256 materials)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    windows = []
    for _ in range(WINDOWS):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        grid.run(steps, progress_bar=False)
        ev1.record(stream)
        torch.cuda.synchronize()
        windows.append(ev0.elapsed_time(ev1))
    ms = float(np.median(windows))
    ms_step = ms / steps
    itemsize = 8 if dtype == "float64" else 4
    half_bytes = algorithmic_bytes_per_half_step(n, pml, itemsize, compensated=dtype == "float32c",
                                                 palette=pal and dtype != "float32c")
    fields = None
    if keep_fields:
        fields = torch.cat([grid.full_precision("E").reshape(-1), grid.full_precision("H").reshape(-1)])
    # dominant kernel: the fused main kernels alone, HIP events on the launch stream
    plan = grid._ensure_plan()
    plan.bench_main_kernel(5, stream.cuda_stream)
    tE, tH = plan.bench_main_kernel(kernel_reps, stream.cuda_stream)
    per_launch = {"k_main_E": tE, "k_main_H": tH}
    dom = max(per_launch, key=per_launch.get)
    achieved = half_bytes / (per_launch[dom] / 1e3) / 1e9
    ctype = {"float64": "double", "float32": "float"}.get(dtype, "compensated")
    if pal and dtype != "float32c":
        ctype += ", palette"
    total_steps = warmup + WINDOWS * steps
    del grid, plan
    torch.cuda.empty_cache()
    return dict(
        mcells=n ** 3 * steps / (ms / 1e3) / 1e6,
        ms_step=ms_step,
        windows_ms_per_step=[round(w / steps, 4) for w in windows],
        steps_run=total_steps,
        step_frac=2 * half_bytes / (ms_step / 1e3) / (HBM_PEAK_GBS * 1e9),
        fields=fields,
        roofline={"bound": "hbm", "kernel": f"{dom}<{ctype}>",
                  "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(achieved / HBM_PEAK_GBS, 4),
                  "traffic": profiled_traffic(dom, ctype, n, pml),
                  "algorithmic_bytes_per_launch": half_bytes,
                  "avg_launch_ms": {k: round(v, 5) for k, v in per_launch.items()}})


def spd_inverse_block(shape, seed=7, lo=1.0, hi=9.0):
    """Per-cell inverse of a random SPD 3x3 permittivity with eigenvalues in
    [lo, hi]: R diag(1/lambda) R^T with R a random rotation (unit quaternion)."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)
    lam = rng.uniform(lo, hi, size=(n, 3))
    inv = np.einsum("nij,nj,nkj->nik", R, 1.0 / lam, R)
    return inv.reshape(tuple(shape) + (3, 3))


def precision_components(shape=(512, 512, 384), pml=20):
    """Registration list of the precision configuration, (kind, key, params,
    attributes set after registration): `pml`-cell PML on all faces, an
    absorbing slab (conductivity 1e4 S/m, 24 cells thick at 512 cells along
    x), an anisotropic block of one eighth of each edge with a seeded SPD
    permittivity (eigenvalues 1-9) and a continuous 1550 nm point source at
    the centre. The object sizes scale with the shape, so a small copy of the
    configuration keeps its structure."""
    nx, ny, nz = shape
    comps = [("PML", k, {}, {}) for k in PML_KEYS(pml)]
    x0, sx = int(0.65 * nx), max(2, round(24 * nx / 512))
    comps.append(("AbsorbingObject", (slice(x0, x0 + sx), slice(pml, ny - pml), slice(pml, nz - pml)),
                  dict(permittivity=2.25, conductivity=1e4), {}))
    bx, by, bz = nx // 8, ny // 8, nz // 8
    box = (slice(int(0.25 * nx), int(0.25 * nx) + bx), slice(int(0.4 * ny), int(0.4 * ny) + by),
           slice(int(0.35 * nz), int(0.35 * nz) + bz))
    comps.append(("AnisotropicObject", box, dict(permittivity=1.0),
                  {"inverse_permittivity": spd_inverse_block((bx, by, bz))}))
    comps.append(("PointSource", (nx // 2, ny // 2, nz // 2), dict(period=1550e-9 / 299792458.0), {}))
    return comps


PRECISION_COURANT = 0.99 * 3 ** -0.5


def build_precision_grid(fdtd, shape=(512, 512, 384), pml=20, amplitude=1.0):
    """The precision configuration on the fdtd_amd API (see
    precision_components); `amplitude` scales the source."""
    import torch

    grid = fdtd.Grid(shape, grid_spacing=50e-9, courant_number=PRECISION_COURANT)
    for kind, key, params, attrs in precision_components(shape, pml):
        if kind == "PointSource":
            params = dict(params, amplitude=amplitude)
        comp = getattr(fdtd, kind)(**params)
        grid[key] = comp
        for name, value in attrs.items():
            setattr(comp, name, torch.from_numpy(value))
    return grid


def build_precision_oracle(orc, shape, pml=20):
    """The same configuration through the oracle's registration API."""
    sim = orc.Sim(shape, 50e-9, 1.0, 1.0, PRECISION_COURANT)
    for kind, key, params, attrs in precision_components(shape, pml):
        comp = sim.add(kind, key, **params)
        for name, value in attrs.items():
            setattr(comp, name, value)
    return sim


def precision_form(args):
    """float32 (and compensated float32) against the float64 path on the
    precision configuration: throughput and field error after `steps`."""
    import torch

    import fdtd_amd as fdtd

    shape = tuple(args.shape) if args.shape else (512, 512, 384)
    steps = args.steps
    out = {"form": "precision", "config": {"workload": "precision: 20-cell PML, absorbing slab, SPD anisotropic block",
                                           "grid": list(shape), "steps": steps}, "runs": {}}
    ref = None
    for name in ["float64", "float32", "float32c"]:
        fdtd.set_backend(f"hip.{name}")
        grid = build_precision_grid(fdtd, shape)
        grid.run(1, progress_bar=False)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        grid.run(steps - 1, progress_bar=False)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / (steps - 1)
        run = {"ms_per_step": round(ms, 4), "Mcells/s": round(np.prod(shape) / (ms / 1e3) / 1e6, 1)}
        fields = torch.cat([grid.full_precision("E").reshape(-1), grid.full_precision("H").reshape(-1)])
        tE, tH = grid._ensure_plan().bench_main_kernel(10, torch.cuda.current_stream().cuda_stream)
        run["main_kernel_ms"] = {"E": round(tE, 4), "H": round(tH, 4)}
        if ref is None:
            ref = fields
        else:
            run["error_vs_f64"] = field_errors(fields, ref)
        out["runs"][name] = run
        del grid, fields
        torch.cuda.empty_cache()
    f64 = out["runs"]["float64"]["ms_per_step"]
    for name, run in out["runs"].items():
        run["speedup_vs_f64_general_kernel"] = round(f64 / run["ms_per_step"], 3)
    print(json.dumps(out), flush=True)


def adjoint_measure(dtype="float32", checkpoints=("sqrt", "record"), reps=3, steps=None):
    """Gradient cost on the differentiation configuration (128x128x64, 12-cell
    PML, 64x64x16 design Object with seeded eps in [1, 12], pulsed 1550 nm
    line source, J = time-integrated |E|^2 over a 16x16 plane, 2000 steps):
    wall time of one gradient evaluation (forward with records/checkpoints,
    recomputation, transposed steps, all host work) over the wall time of the
    plain forward run of the same steps; the best of `reps` of each."""
    import torch

    import fdtd_amd as fdtd
    from fdtd_amd.adjoint import IntensityGradient

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden"))
    import adjoint_scenarios as A

    sc = A.C5
    fdtd.set_backend(f"hip.{dtype}")
    perm = torch.from_numpy(A.design_permittivity(sc))
    grid, design = A.build(fdtd, sc, perm)
    T = steps or sc["steps"]

    def timed(fn):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt_ = time.perf_counter() - t0
            best = dt_ if best is None else min(best, dt_)
        return best * 1e3

    probs = {c: IntensityGradient(grid, design, sc["plane"], T, c) for c in checkpoints}
    first = probs[checkpoints[0]]
    first._set_permittivity(perm)

    def forward():
        first._reset()
        grid.run(T, progress_bar=False)

    forward()
    fwd_ms = timed(forward)
    out = {"metric": "adjoint/forward time ratio", "unit": "x", "dtype": "f32" if dtype == "float32" else "f64",
           "config": {"workload": "differentiation: 128x128x64, 12-cell PML x6, 64x64x16 design Object, "
                                  "pulsed line source, sum_t sum_16x16 plane |E|^2", "steps": T},
           "forward_ms": round(fwd_ms, 2), "gradient_ms": {}, "ratio": {}}
    for c, prob in probs.items():
        prob.gradient(perm)
        ms = timed(lambda: prob.gradient(perm))
        out["gradient_ms"][c] = round(ms, 2)
        out["ratio"][c] = round(ms / fwd_ms, 3)
        out.setdefault("segments", {})[c] = prob.last_info
    out["value"] = out["ratio"][checkpoints[0]]
    del grid, design, probs, first
    torch.cuda.empty_cache()
    return out


def batch_specs(count=64, seed=0):
    """The batching configuration's grids as (shape, permittivity, source
    cell): edges drawn uniformly from 64-192 cells, a per-voxel permittivity
    of smoothed Gaussian noise mapped to [1, 12] (noise on a grid 8x coarser,
    trilinearly interpolated) and a random source position at least 10 cells
    from every face."""
    import torch
    import torch.nn.functional as Fn

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        shape = tuple(int(v) for v in rng.integers(64, 193, size=3))
        coarse = torch.from_numpy(rng.normal(size=tuple(n // 8 + 2 for n in shape)))
        fine = Fn.interpolate(coarse[None, None], size=shape, mode="trilinear", align_corners=True)[0, 0]
        eps = 1.0 + 11.0 * (fine - fine.min()) / (fine.max() - fine.min())
        p = tuple(int(rng.integers(BATCH_PML, n - BATCH_PML)) for n in shape)
        out.append((shape, eps, p))
    return out


BATCH_PML = 10


def batch_grid(fdtd, spec):
    """One grid of the batching configuration: 10-cell PML on all faces and a
    pulsed 1550 nm point source."""
    shape, eps, p = spec
    grid = fdtd.Grid(shape, grid_spacing=50e-9, permittivity=eps[..., None], courant_number=3 ** -0.5)
    for k in PML_KEYS(BATCH_PML):
        grid[k] = fdtd.PML()
    grid[p[0], p[1], p[2]] = fdtd.PointSource(period=1550e-9 / 299792458.0, pulse=True)
    return grid


def batch_oracle(orc, spec):
    """The same grid through the oracle's registration API."""
    shape, eps, p = spec
    sim = orc.Sim(shape, 50e-9, eps.numpy()[..., None], 1.0, 3 ** -0.5)
    for k in PML_KEYS(BATCH_PML):
        sim.add("PML", k)
    sim.add("PointSource", p, period=1550e-9 / 299792458.0, pulse=True)
    return sim


def batch_grids(fdtd, torch, count=64, seed=0):
    """The batching configuration: `count` grids (batch_specs, batch_grid)."""
    return [batch_grid(fdtd, spec) for spec in batch_specs(count, seed)]


def batch_measure(steps=2000, seq_steps=200, count=64, launch=None):
    """Aggregate throughput of the batched step (autotuned launch shape, HIP
    graph replays) on the batching configuration, beside the same grids run
    one after another through the single-grid path."""
    import torch

    import fdtd_amd as fdtd
    from fdtd_amd.batch import GridBatch

    fdtd.set_backend("hip.float32")
    grids = batch_grids(fdtd, torch, count)
    cells = sum(g.Nx * g.Ny * g.Nz for g in grids)
    slab_cells = sum(2 * 10 * (g.Ny * g.Nz + g.Nx * g.Nz + g.Nx * g.Ny) for g in grids)
    step_bytes = 2 * (cells * 12 * 4 + slab_cells * 4 * 4)
    batch = GridBatch(grids, graph_steps=min(100, steps))
    if launch:  # a fixed launch shape (block threads, x planes, workgroups): no autotuning
        batch.set_launch(*launch)
        tune = {}
    else:
        tune = batch.autotune(steps=10)
    stream = torch.cuda.current_stream()

    def timed(fn):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1)

    batch.run(min(100, steps))  # graph capture and warm-up
    ms = timed(lambda: batch.run(steps))
    ms_direct = timed(lambda: batch.run(seq_steps, graph=False)) / seq_steps
    ms_seq = timed(lambda: [g.run(seq_steps, progress_bar=False) for g in grids]) / seq_steps
    ms_step = ms / steps
    achieved = step_bytes / (ms_step / 1e3) / 1e9
    out = {"metric": "batched Mcells/s (64 uneven fp32 grids)", "unit": "Mcells/s", "dtype": "f32",
           "value": round(cells / (ms_step / 1e3) / 1e6, 1),
           "config": {"workload": f"batching: {count} grids, edges U[64,192], smoothed-noise eps in [1,12], "
                                  "10-cell PML x6, pulsed point source", "cells": cells, "steps": steps},
           "ms_per_step": round(ms_step, 4),
           "graph_vs_direct": round(ms_direct / ms_step, 3),
           "sequential_single_grid": {"Mcells/s": round(cells / (ms_seq / 1e3) / 1e6, 1),
                                      "ms_per_step": round(ms_seq, 4), "steps": seq_steps},
           "speedup_vs_sequential": round(ms_seq / ms_step, 3),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_step": step_bytes},
           "launch": {"block_threads": batch.launch[0], "cx": batch.launch[1], "workgroups": batch.launch[2]},
           "autotune_ms_per_step": {f"{k[0]}x{k[1]}x{k[2]}": round(v, 4) for k, v in sorted(tune.items())}}
    del batch, grids
    torch.cuda.empty_cache()
    return out


def adjoint_form(args):
    import torch

    copy_gbs = copy_bandwidth(torch, 1 << 28, 2)  # also the scale of byte counters in a profiled run
    out = adjoint_measure("float64" if args.adjoint_f64 else "float32", reps=args.adjoint_reps,
                          steps=args.adjoint_steps)
    out["copy_GBps"] = round(copy_gbs, 1)
    out["higher_is_better"] = False
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--dtype", default="float64", choices=["float64", "float32", "float32c"])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--pml", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the float32, compensated and float64-palette legs of the plain line")
    ap.add_argument("--form", default="plain", choices=["plain", "precision", "adjoint", "batch"],
                    help="plain: the 256^3 headline line; precision: float32 forms vs float64; "
                         "adjoint: gradient/forward time on the differentiation configuration; "
                         "batch: 64 uneven grids stepped together")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement of the plain line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching measurement of the plain line")
    ap.add_argument("--palette", default="auto", choices=["auto", "on", "off"],
                    help="grid material as a <=256-entry palette (auto: in the compensated form only)")
    ap.add_argument("--adjoint-f64", action="store_true", help="--form adjoint in float64")
    ap.add_argument("--shape", type=int, nargs=3, default=None, help="grid shape for --form precision")
    ap.add_argument("--adjoint-steps", type=int, default=None, help="--form adjoint: steps (default: the workload's 2000)")
    ap.add_argument("--adjoint-reps", type=int, default=3, help="--form adjoint: timed repetitions (best of)")
    ap.add_argument("--batch-steps", type=int, default=2000, help="--form batch: timed steps")
    ap.add_argument("--batch-launch", default=None,
                    help="--form batch: fixed launch shape 'threads,cx,workgroups' instead of autotuning")
    args = ap.parse_args()
    global PALETTE
    PALETTE = args.palette
    if args.gpus != 1:
        sys.stderr.write("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)\n")
        sys.exit(2)

    if args.form == "precision":
        return precision_form(args)
    if args.form == "adjoint":
        return adjoint_form(args)
    if args.form == "batch":
        import torch

        copy_gbs = copy_bandwidth(torch, 1 << 28, 2)  # also the scale of byte counters in a profiled run
        launch = tuple(int(v) for v in args.batch_launch.split(",")) if args.batch_launch else None
        out = batch_measure(steps=args.batch_steps, seq_steps=min(200, args.batch_steps), launch=launch)
        out["copy_GBps"] = round(copy_gbs, 1)
        out["higher_is_better"] = True
        print(json.dumps(out), flush=True)
        return

    import torch

    import fdtd_amd as fdtd

    n, pml = args.n, args.pml
    legs = args.dtype == "float64" and not args.no_fp32
    main = measure(torch, fdtd, args.dtype, n, pml, args.steps, args.warmup, args.kernel_reps, keep_fields=legs)
    others = {}
    if legs:
        # the float32 forms on the same workload and step count, each with its
        # error against the float64 run's final fields
        for name in ("float32", "float32c"):
            m = measure(torch, fdtd, name, n, pml, args.steps, args.warmup, args.kernel_reps, keep_fields=True)
            m["error_vs_f64"] = field_errors(m.pop("fields"), main["fields"])
            others[name] = m
        if PALETTE == "auto":
            # a specialised float64 kernel: the grid materials as a 1-byte
            # palette index into their float64 values (<= 256 distinct), the
            # same arithmetic on the same values, so the fields must be equal
            m = measure(torch, fdtd, "float64", n, pml, args.steps, args.warmup, args.kernel_reps, keep_fields=True,
                        palette="on")
            m["bit_identical"] = bool(torch.equal(m.pop("fields"), main["fields"]))
            others["float64_palette"] = m
    main.pop("fields", None)
    copy_gbs = copy_bandwidth(torch)

    base = None
    if not args.no_cpu_baseline:
        base = cpu_baseline(n, pml, args.cpu_steps)

    # arithmetic type; the compensated storage (float + 8 bits) computes in float64
    dt = "f32" if args.dtype == "float32" else "f64"
    storage = {"float64": "f64", "float32": "f32", "float32c": "f32+8 compensated"}[args.dtype]
    out = {
        "metric": METRIC,
        "value": round(main["mcells"], 1),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main["ms_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dt,
        "data": "synthetic",
        "config": {"workload": f"{n}^3 grid, {pml}-cell PML x6, 24 dielectric spheres, pulsed point source, {storage}",
                   "grid": [n, n, n], "pml": pml, "storage": storage, "parallelism": "single GPU"},
        "step_hbm_roofline_frac": round(main["step_frac"], 4),
        "windows_ms_per_step": main["windows_ms_per_step"],
        "copy_GBps": round(copy_gbs, 1),
        "roofline": main["roofline"],
        "cpu_baseline": base,
    }
    for name, key, storage_ in (("float32", "fp32", "f32 fields, materials and CPML state, f32 arithmetic"),
                                ("float32c", "fp32c", "f32 + 8 compensation bits per field value, f64 arithmetic "
                                                      "and CPML state, 1-byte material palette")):
        if name not in others:
            continue
        o = others[name]
        out[key] = {"value": round(o["mcells"], 1), "unit": "Mcells/s", "ms_per_step": round(o["ms_step"], 4),
                    "windows_ms_per_step": o["windows_ms_per_step"], "storage": storage_,
                    "step_hbm_roofline_frac": round(o["step_frac"], 4), "roofline": o["roofline"],
                    "speedup_vs_f64": round(o["mcells"] / main["mcells"], 3),
                    "error_vs_f64": dict(o["error_vs_f64"], steps=o["steps_run"],
                                         against="hip.float64 run of the same workload and steps")}
    if "float64_palette" in others:
        o = others["float64_palette"]
        out["fp64_palette"] = {
            "value": round(o["mcells"], 1), "unit": "Mcells/s", "ms_per_step": round(o["ms_step"], 4),
            "windows_ms_per_step": o["windows_ms_per_step"],
            "storage": "f64 fields, CPML state and arithmetic; grid materials as a 1-byte index into a float64 "
                       "palette (a specialised kernel: at most 256 distinct materials), roofline against its own "
                       "reduced bytes",
            "step_hbm_roofline_frac": round(o["step_frac"], 4), "roofline": o["roofline"],
            "speedup_vs_f64": round(o["mcells"] / main["mcells"], 3),
            "bit_identical_to_f64": o["bit_identical"]}
    if not args.no_adjoint:
        out["adjoint"] = adjoint_measure("float32")
    if not args.no_batch:
        out["batch"] = batch_measure()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
