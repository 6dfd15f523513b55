"""Many independent grids stepped together on one GPU.

The reference steps one Grid at a time (fdtd/grid.py:250-273); running many
small simulations that way leaves most of the GPU idle and pays two or more
launches per grid and step. ``GridBatch`` steps any number of grids of
different shapes with one launch per half step (libfdtd_hip:
fdtd_batch_run / fdtd_batch_run_graph): the tiles of all grids form one work
list (one workgroup per tile, or a persistent set of workgroups walking it),
and the step loop is captured in a HIP graph and replayed. Each grid's arithmetic is exactly its own plan's, so
a batched run matches running the grids one by one, bit for bit.

Eligible grids: PML boundaries, objects and at most 4 point sources (no
periodic boundaries, no detectors, no line or plane sources), all of one
backend dtype and at the same time step.

    batch = GridBatch(grids)
    batch.autotune()          # pick the launch shape by timing candidates
    batch.run(2000)
"""

import itertools

import torch

from . import _native

# launch shapes chosen by autotune(), per batch signature
_TUNED = {}


def _same_key(a, b):
    """Entry by entry: the same plan and waveform objects at the same table
    address and length."""
    if b is None or len(a) != len(b):
        return False
    return all(pa is pb and wa is wb and ka == kb for (pa, wa, ka), (pb, wb, kb) in zip(a, b))


def _signature(grids):
    g0 = grids[0]
    return (str(g0.dtype), bool(g0.compensated),
            tuple((g.Nx, g.Ny, g.Nz, len(g.boundaries), len(g.objects), len(g.sources)) for g in grids))


class GridBatch:
    """Step `grids` together; see the module docstring."""

    def __init__(self, grids, graph_steps=100):
        grids = list(grids)
        if not grids:
            raise ValueError("GridBatch needs at least one grid")
        g0 = grids[0]
        for g in grids:
            if g.dtype != g0.dtype or g.compensated != g0.compensated:
                raise ValueError("batched grids must share one backend dtype")
            if g.detectors:
                raise ValueError("batched grids cannot have detectors")
            if g.time_steps_passed != g0.time_steps_passed:
                raise ValueError("batched grids must be at the same time step")
        self.grids = grids
        self.graph_steps = int(graph_steps)
        self._batch = None
        self._plan_ids = None
        self.launch = _TUNED.get(_signature(grids), (256, 1, 0))

    # ------------------------------------------------------------ plumbing
    def _ready(self, nsteps):
        plans = [g._prepare(nsteps, nsteps, nsteps) for g in self.grids]
        # The batch holds copies of each plan's arguments, the waveform
        # table's address and length included. The key keeps references to
        # the plans and tables it was built from, so that none of them can be
        # freed (and its address reused) while the batch may still point there.
        key = tuple((p, g._wave, None if g._wave is None else (g._wave.data_ptr(), g._wave.shape[1]))
                    for p, g in zip(plans, self.grids))
        if self._batch is None or not _same_key(key, self._plan_ids):
            self._batch = _native.Batch(plans)
            self._batch.set_launch(*self.launch)
            self._plan_ids = key
        return self._batch

    def _state(self):
        out = []
        for g in self.grids:
            out += [g._E, g._H]
            if g.compensated:
                out += [g._E_lo, g._H_lo]
            for b in g.boundaries:
                out += [t for t in (getattr(b, "_psiE", None), getattr(b, "_psiH", None)) if t is not None]
        return out

    # -------------------------------------------------------------- public
    @property
    def time_steps_passed(self):
        return self.grids[0].time_steps_passed

    def set_launch(self, block_threads=256, cx=1, workgroups=0):
        """Threads per workgroup (256/512/1024), x planes per work item and
        the number of workgroups (0: one per work item; fewer: persistent)."""
        self.launch = (int(block_threads), int(cx), int(workgroups))
        if self._batch is not None:
            self._batch.set_launch(*self.launch)

    def run(self, total_time, graph=True):
        """Advance every grid by `total_time` steps (an int, or a float time
        in seconds when all grids share one time step)."""
        if isinstance(total_time, float):
            dt = self.grids[0].time_step
            if any(abs(g.time_step - dt) > 1e-12 * dt for g in self.grids):
                raise ValueError("a time in seconds needs grids with one time step; pass a step count")
            total_time /= dt
        n = int(total_time)
        if n <= 0:
            return
        q0 = self.time_steps_passed
        if any(g.time_steps_passed != q0 for g in self.grids):
            raise ValueError("batched grids must be at the same time step "
                             f"(found {sorted({g.time_steps_passed for g in self.grids})})")
        batch = self._ready(n)
        stream = torch.cuda.current_stream().cuda_stream
        if graph and n >= self.graph_steps:
            batch.run_graph(q0, n, self.graph_steps, stream)
        else:
            batch.run(q0, n, stream)
        for g in self.grids:
            g.time_steps_passed += n

    def autotune(self, steps=10, candidates=None, verbose=False):
        """Time `steps` steps for each candidate launch shape (block threads,
        x planes per item, workgroups) on a copy of the state, keep the
        fastest, and restore the state. Returns {shape: ms per step}."""
        if candidates is None:
            cus = torch.cuda.get_device_properties(self.grids[0]._E.device).multi_processor_count
            candidates = list(itertools.product((256, 512), (1, 2, 4), (0, cus * 8, cus * 32)))
        saved = [t.clone() for t in self._state()]
        q_saved = self.time_steps_passed
        times = {}
        stream = torch.cuda.current_stream()
        try:
            for shape in candidates:
                self.set_launch(*shape)
                self.run(2, graph=False)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                self.run(steps, graph=False)
                e1.record(stream)
                e1.synchronize()
                times[shape] = e0.elapsed_time(e1) / steps
                if verbose:
                    print(shape, f"{times[shape]:.4f} ms/step", flush=True)
        finally:
            for dst, src in zip(self._state(), saved):
                dst.copy_(src)
            for g in self.grids:
                g.time_steps_passed = q_saved
        best = min(times, key=times.get)
        self.set_launch(*best)
        _TUNED[_signature(self.grids)] = best
        return times
