"""Reverse-mode gradients of HIP-backend runs with respect to a design region.

The reference differentiates a run by stepping its torch backend with
autograd enabled (fdtd/backend.py:207-361): torch records every array
operation of Grid.update_E/update_H (fdtd/grid.py:275-325), PML.update_phi_*
/update_* (fdtd/boundaries.py:409-487) and Object.update_E (fdtd/objects.py:
118-129), then replays their transposes. Here the same derivative comes from
hand-written transposed half-step kernels (libfdtd_hip: fdtd_run_recorded,
fdtd_adjoint_run) for one objective,

    J = sum over the T steps of sum over an objective box of |E|^2,

with E taken after each step, and one design region: a plain ``Object``
registered in the grid whose permittivity is the variable. The run starts from
rest (E = H = CPML state = 0, time index 0), as a fresh reference grid does.

Memory. The gradient needs, for every step, curl H inside the design box and
E inside the objective box. ``checkpoint="record"`` keeps them for all T
steps (one forward pass plus one reverse pass). ``checkpoint="sqrt"`` keeps
a copy of the whole state (E, H, every CPML psi) every K steps instead; the
reverse pass restores each checkpoint, steps its segment forward again
recording K steps, and runs the transposed steps over it. There are B record
buffers of K steps; the last B segments, one per buffer, are recorded by the
first forward pass into the buffers the reverse pass reads, so they are
neither checkpointed nor recomputed. K and B come from the memory: the
budget is that of checkpoints every ceil(sqrt(T)) steps plus a quarter,
O(sqrt(T)) states, and within it the pair that records the most steps
(records are much smaller than a state when the design and objective boxes
are small, so fewer, longer segments and more buffers pay; see
``_sqrt_schedule``).
``"auto"`` picks "record" when the records fit in a quarter of the free
device memory.
"""

import math

import numpy as np
import torch

from . import _native
from .boundaries import PML, PeriodicBoundary
from .objects import AbsorbingObject, AnisotropicObject, Object

# checkpointed mode: at most RECORD_BUFFERS record buffers (the last ones'
# segments are recorded by the forward pass, the others recomputed into them
# in turn), SEGMENT_STEPS a fixed segment length instead of the one chosen
# from the memory (None), and x planes per workgroup of the transposed
# kernels running beside the recomputation (0: the plan's own; on MI355X at
# 128x128x64, 2000 steps, where that is 1: 1 and 3 2.45x the forward, 2
# 2.52x, 4 2.56x).
RECORD_BUFFERS = 16
SEGMENT_STEPS = None
SQRT_MODE_CX = 0


def _box_of(grid, key):
    """Half-open box (x0, x1, y0, y1, z0, z1) of a grid index tuple of ints
    and unit-step slices."""
    if not isinstance(key, tuple) or len(key) != 3:
        raise IndexError("the objective box is indexed as grid[x, y, z] with ints or unit-step slices")
    out = []
    for k, n in zip(key, (grid.Nx, grid.Ny, grid.Nz)):
        if isinstance(k, (int, np.integer)):
            k = int(k)
            if k < 0:
                k += n
            lo, hi = k, k + 1
        elif isinstance(k, slice):
            if k.step not in (None, 1):
                raise IndexError("the objective box needs unit-step slices")
            lo, hi, _ = k.indices(n)
        else:
            raise IndexError(f"invalid index {k!r} for the objective box")
        if not 0 <= lo < hi <= n:
            raise IndexError("empty or out-of-range objective box")
        out += [lo, hi]
    return tuple(out)


class IntensityGradient:
    """Objective J = sum_t sum_box |E|^2 of a T-step run from rest and its
    gradient with respect to the permittivity of the design object.

    grid:        a HIP-backend grid with PML faces (at most one per side),
                 periodic boundaries registered after every PML face, plain,
                 absorbing and anisotropic objects (an anisotropic object
                 may not overlap an absorbing one) and point, line or plane
                 sources (no soft sources), float64 or float32 storage;
                 detectors are ignored. A plane source assigns E and H at its
                 cells (sources.py:476-486): those values do not depend on
                 the state, so the transposed half steps zero the adjoint
                 fields there. A periodic boundary copies a plane after each
                 half step (boundaries.py:184-219): its transpose adds the
                 copied plane's adjoint to the source plane's and zeroes it.
    design:      a plain ``Object`` registered in ``grid``
    box:         the objective box, a grid index tuple (x, y, z)
    steps:       T
    checkpoint:  "sqrt" | "record" | "auto"
    """

    def __init__(self, grid, design, box, steps, checkpoint="sqrt"):
        if grid.compensated or grid.complex:
            raise NotImplementedError("gradients need real hip.float64 or hip.float32 fields")
        if type(design) is not Object or design.grid is not grid:
            raise ValueError("the design region must be a plain Object registered in the grid")
        seen_periodic = False
        for b in grid.boundaries:
            if isinstance(b, PeriodicBoundary):
                seen_periodic = True
            elif not isinstance(b, PML):
                raise NotImplementedError(f"gradients do not support {type(b).__name__}")
            elif seen_periodic:
                raise NotImplementedError("gradients need every PML face registered before the periodic boundaries")
        for obj in grid.objects:
            if obj is not design and _overlap(obj, design):
                raise ValueError("another object overlaps the design region")
            if isinstance(obj, AnisotropicObject):
                for other in grid.objects:
                    if isinstance(other, AbsorbingObject) and _overlap(obj, other):
                        raise NotImplementedError("gradients do not support an anisotropic object overlapping an "
                                                  "absorbing one (the two updates do not compose)")
        for src in grid.sources:
            if type(src).__name__ == "SoftArbitraryPointSource":
                raise NotImplementedError("soft sources are not supported by the gradient (a resistive one feeds "
                                          "the fields back into the source; its transpose is not implemented)")
        if checkpoint not in ("sqrt", "record", "auto"):
            raise ValueError("checkpoint must be 'sqrt', 'record' or 'auto'")
        self.grid = grid
        self.design = design
        self.box = _box_of(grid, box)
        self.steps = int(steps)
        if self.steps < 1:
            raise ValueError("steps must be positive")
        d = design
        self.design_box = (d.x.start, d.x.stop, d.y.start, d.y.stop, d.z.start, d.z.stop)
        self.nd = d.Nx * d.Ny * d.Nz
        b = self.box
        self.np = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])
        self.checkpoint = checkpoint
        # the design's border entries hold the grid's values, not the
        # permittivity's (objects.py:79-90): keep them
        inv = design._device_array("inverse_permittivity", (d.Nx, d.Ny, d.Nz, 3))
        self._border = torch.zeros(inv.shape, dtype=torch.bool, device=inv.device)
        if d.Nx > 1:
            self._border[-1, :, :, 0] = True
        if d.Ny > 1:
            self._border[:, -1, :, 1] = True
        if d.Nz > 1:
            self._border[:, :, -1, 2] = True
        self._border_values = inv[self._border].clone()
        self.last_info = {}

    # ---------------------------------------------------------- plumbing
    def _set_permittivity(self, permittivity):
        """design.inverse_permittivity <- 1/permittivity in place (border
        entries kept); returns the broadcast permittivity it used."""
        d = self.design
        inv = d._device_array("inverse_permittivity", (d.Nx, d.Ny, d.Nz, 3))
        p = torch.as_tensor(permittivity).detach().to(device=inv.device, dtype=torch.float64)
        if p.dim() == 3:
            p = p[..., None]
        p = p.expand(d.Nx, d.Ny, d.Nz, 3)
        new = (1.0 / p).to(inv.dtype)
        new[self._border] = self._border_values
        inv.copy_(new)
        d.permittivity = permittivity
        return p

    def _coefficients(self):
        """(gcoef, dcoef) of the composed E update E' = dcoef E + gcoef curl H
        (the diagonal part: the library adds the transposes of the
        anisotropic objects' 3x3 updates itself)."""
        g = self.grid
        S = g.courant_number
        gcoef = (S * g._inv_eps.to(torch.float64))
        dcoef = None
        for obj in g.objects:
            if isinstance(obj, AnisotropicObject):
                continue
            sl = (slice(None), obj.x, obj.y, obj.z)
            inv = obj.inverse_permittivity.to(torch.float64).permute(3, 0, 1, 2)
            if isinstance(obj, AbsorbingObject):
                f = obj.absorption_factor.to(torch.float64).permute(3, 0, 1, 2)
                a = (1 - f) / (1 + f)
                if dcoef is None:
                    dcoef = torch.ones_like(gcoef)
                dcoef[sl] = dcoef[sl] * a
                gcoef[sl] = gcoef[sl] * a + S * inv / (1 + f)
            else:
                gcoef[sl] = gcoef[sl] + S * inv
        dt = g.dtype

        def pitched(t):  # the fields' (pitched) layout, which the kernels index
            out = g._alloc_soa(dt)
            out.copy_(t)
            return out

        return pitched(gcoef), (pitched(dcoef) if dcoef is not None else None)

    def _plan(self):
        """A plan of the grid's boundaries, objects and sources (no detectors)
        with a waveform table covering the T steps, built once (the object
        arrays it reads are updated in place)."""
        if getattr(self, "_cached_plan", None) is None:
            self._cached_plan = self._new_plan()
        return self._cached_plan

    def _new_plan(self):
        g = self.grid
        _require_device(g._E)
        code = _native.F64 if g.dtype == torch.float64 else _native.F32
        plan = _native.Plan(code, (g.Nx, g.Ny, g.Nz), g.courant_number, g._E, g._H, g._inv_eps, g._inv_mu,
                            pitch=g._px)
        for comp in list(g.boundaries) + list(g.objects) + list(g.sources):
            comp._add_to_plan(plan)
        if g.sources:
            table = np.empty((len(g.sources), self.steps), dtype=np.float64)
            for r, src in enumerate(g.sources):
                table[r] = [src._waveform(q) for q in range(self.steps)]
            wave = torch.from_numpy(table).to(g._E.device)
            plan.set_waveforms(wave)
        else:
            plan.set_waveforms(None)
        plan.finalize()
        return plan

    def _state(self):
        g = self.grid
        out = [g._E, g._H]
        for b in g.boundaries:
            if isinstance(b, PML):
                out += [b._psiE, b._psiH]
        return out

    def _reset(self):
        for t in self._state():
            t.zero_()
        self.grid.time_steps_passed = 0

    def _sqrt_schedule(self):
        """(K, B) of the checkpointed mode: segments of K steps and B record
        buffers. With C the bytes of a state (E, H, every CPML psi) and r of
        one step's records (curl in the design box, E in the objective box),
        its memory is (ceil(T/K) - B) checkpoints and B buffers of K steps,
        and the recomputation covers the T - B K steps that the forward pass
        does not record. Segments are K0 = ceil(sqrt(T)) to 2 K0 steps long
        and the memory at most that of the classic scheme's checkpoints (every
        K0 steps) plus a quarter, 1.25 ceil(T/K0) C: O(sqrt(T)) either way.
        Within that, the pair with the most recorded steps (2 <= B <=
        RECORD_BUFFERS; ties: less memory); K0 and 2 buffers when none fits.
        SEGMENT_STEPS fixes K."""
        T = self.steps
        esz = self.grid._E.element_size()
        r = 3 * (self.nd + self.np) * esz
        C = sum(t.numel() * t.element_size() for t in self._state())
        K0 = max(1, math.isqrt(T - 1) + 1)
        budget = 1.25 * (-(-T // K0)) * C
        if SEGMENT_STEPS is None:
            Ks = np.arange(K0, min(2 * K0, T) + 1, dtype=np.float64)
        else:
            Ks = np.array([float(SEGMENT_STEPS)])
        nseg = np.ceil(T / Ks)
        best = None  # (recorded steps, -memory, K, B)
        for B in range(2, max(2, RECORD_BUFFERS) + 1):
            mem = np.maximum(nseg - B, 0) * C + np.minimum(B, nseg) * Ks * r
            rec = np.minimum(B * Ks, T)
            ok = mem <= budget
            if not ok.any():
                continue
            i = np.lexsort((mem[ok], -rec[ok]))[0]  # most recorded steps, then least memory
            cand = (float(rec[ok][i]), -float(mem[ok][i]), int(Ks[ok][i]), B)
            if best is None or cand[:2] > best[:2]:
                best = cand
        if best is None:
            return (int(SEGMENT_STEPS) if SEGMENT_STEPS else K0), 2
        K, B = best[2], best[3]
        return K, int(min(B, -(-T // K)))

    def _mode(self):
        if self.checkpoint != "auto":
            return self.checkpoint
        esz = self.grid._E.element_size()
        need = self.steps * 3 * (self.nd + self.np) * esz
        free, _ = torch.cuda.mem_get_info(self.grid._E.device)
        return "record" if need <= free // 4 else "sqrt"

    # ------------------------------------------------------------- public
    def objective(self, permittivity):
        """J for the given design permittivity (forward only)."""
        return self._run(permittivity, gradient=False)[0]

    def gradient(self, permittivity):
        """(J, dJ/d permittivity) for the given design permittivity; the
        gradient has the shape of ``permittivity``, float64."""
        return self._run(permittivity, gradient=True)

    def _reverse_checkpointed(self, adj, segments, checkpoints, curl, plane):
        """Reverse sweep of the checkpointed mode on two streams: the segments
        are recomputed from their checkpoints (forward kernels, recording) on
        a second stream while the transposed steps of later segments run on
        the current one. Both are latency-bound kernels on disjoint data (the
        recomputation writes the grid state and one record buffer, the
        transposed steps read another buffer and write the adjoint state), so
        they overlap. Segment s uses record buffer s % len(curl); the last
        len(curl) segments were recorded by the forward pass on the current
        stream, checkpoints[s] restores each earlier one. Events order each
        buffer's reuse: a recomputation is issued only after the transposed
        steps that last read its buffer, so that its wait sees them."""
        cur = torch.cuda.current_stream()
        side = getattr(self, "_side", None)
        if side is None or side.device != cur.device:
            side = self._side = torch.cuda.Stream(device=cur.device)
        nbuf = len(curl)
        final = [t.clone() for t in self._state()]
        side.wait_stream(cur)  # the forward pass and the final-state copy are done
        recorded = [None] * len(segments)
        consumed = [None] * nbuf  # per buffer: the transposed steps that last read it

        def recompute(s):
            b = s % nbuf
            with torch.cuda.stream(side):
                if consumed[b] is not None:
                    side.wait_event(consumed[b])
                for dst, src in zip(self._state(), checkpoints[s]):
                    dst.copy_(src)
                q, m = segments[s]
                adj.run_recorded(q, m, curl[b], plane[b], side.cuda_stream)
                recorded[s] = torch.cuda.Event()
                recorded[s].record(side)

        for op, s in reverse_schedule(len(segments), nbuf):
            if op == "recompute":
                recompute(s)
                continue
            b = s % nbuf
            if recorded[s] is not None:
                cur.wait_event(recorded[s])
            adj.run_adjoint(segments[s][1], curl[b], plane[b], cur.cuda_stream)
            consumed[b] = torch.cuda.Event()
            consumed[b].record(cur)
        cur.wait_stream(side)
        for dst, src in zip(self._state(), final):
            dst.copy_(src)
        # the checkpoints were read on the side stream
        for cp in checkpoints:
            for t in cp:
                t.record_stream(side)

    def _run(self, permittivity, gradient):
        g = self.grid
        perm_b = self._set_permittivity(permittivity)
        plan = self._plan()
        dev, dt = g._E.device, g.dtype
        stream = torch.cuda.current_stream().cuda_stream
        T = self.steps
        mode = self._mode() if gradient else "sqrt"
        sqrt_grad = gradient and mode == "sqrt"
        if mode == "record":
            K, nbuf = T, 1
        elif sqrt_grad:
            K, nbuf = self._sqrt_schedule()
        else:
            K, nbuf = max(1, math.isqrt(T - 1) + 1), 1
        gcoef, dcoef = self._coefficients()
        lamE = g._alloc_soa(g._E.dtype)  # the fields' (pitched) layout
        lamH = g._alloc_soa(g._H.dtype)
        psi = torch.zeros(_native.Adjoint.psi_size(plan), dtype=dt, device=dev)
        grad = torch.zeros((3, self.nd), dtype=torch.float64, device=dev)
        adj = _native.Adjoint(plan, gcoef, dcoef, lamE, lamH, psi, self.design_box, self.box, grad)
        # in the checkpointed mode several buffers of K steps: the
        # recomputation of earlier segments fills some while the transposed
        # steps of a later one read another; fewer and longer workgroups of
        # the transposed kernels leave room for the recomputation beside them
        adj.set_launch(SQRT_MODE_CX if sqrt_grad else 0)
        curl = [torch.empty((K, 3, self.nd), dtype=dt, device=dev) for _ in range(nbuf)]
        plane = [torch.empty((K, 3, self.np), dtype=dt, device=dev) for _ in range(nbuf)]
        J = torch.zeros((), dtype=torch.float64, device=dev)

        self._reset()
        segments = [(q, min(K, T - q)) for q in range(0, T, K)]
        last = len(segments) - 1
        checkpoints = []
        for s, (q, m) in enumerate(segments):
            b = s % nbuf
            # checkpointed gradient: the last nbuf segments are recorded now,
            # into the buffers the reverse sweep reads them from, and are
            # not recomputed; the others keep a checkpoint instead
            keep = mode == "record" or (sqrt_grad and s > last - nbuf)
            if sqrt_grad and not keep:
                checkpoints.append([t.clone() for t in self._state()])
            adj.run_recorded(q, m, curl[b] if keep else None, plane[b], stream)
            J += plane[b][:m].to(torch.float64).square().sum()
        g.time_steps_passed = T
        self.last_info = {"mode": mode, "segment": K, "checkpoints": len(checkpoints), "record_buffers": nbuf}
        if not gradient:
            return float(J), None
        if mode == "record":
            for s in range(len(segments) - 1, -1, -1):
                adj.run_adjoint(segments[s][1], curl[0], plane[0], stream)
        else:
            self._reverse_checkpointed(adj, segments, checkpoints, curl, plane)
        d = self.design
        g_inv = grad.reshape(3, d.Nx, d.Ny, d.Nz).permute(1, 2, 3, 0).clone()
        g_inv[self._border] = 0
        # inv = 1/perm: dJ/dperm = -dJ/dinv / perm^2, summed over broadcast axes
        g_perm_b = -g_inv / perm_b.square()
        return float(J), _reduce_to(g_perm_b, permittivity)


def reverse_schedule(nseg, nbuf):
    """Host issue order of the checkpointed reverse sweep over nseg segments
    with nbuf record buffers: ("recompute", s) and ("adjoint", s) pairs.
    Segment s uses buffer s % nbuf; the last nbuf segments (recorded by the
    forward pass) are not recomputed. Each segment's transposed steps come
    after its recomputation, and a recomputation into buffer b comes after the
    transposed steps of the segment that last held b (s + nbuf), so that a
    wait issued with it sees them; within that bound recomputations run as
    far ahead as the buffers allow."""
    ops = []
    pending = nseg - nbuf - 1  # the next segment to recompute, downwards
    for s in range(nseg - 1, -1, -1):
        # buffer (s' % nbuf) was last read by segment s' + nbuf >= s + 1
        while pending >= 0 and pending >= s + 1 - nbuf:
            ops.append(("recompute", pending))
            pending -= 1
        ops.append(("adjoint", s))
    return ops


def _reduce_to(g, permittivity):
    """Sum a (Nx, Ny, Nz, 3) gradient down to the shape of `permittivity`."""
    shape = tuple(torch.as_tensor(permittivity).shape)
    if len(shape) == 3:
        return g.sum(dim=3)
    if len(shape) == 4:
        return g if shape[3] == 3 else g.sum(dim=3, keepdim=True)
    if len(shape) == 0:
        return g.sum()
    raise ValueError(f"unsupported permittivity shape {shape}")


def _require_device(t):
    if t.device.type != "cuda":
        raise RuntimeError("gradients need a HIP device")


def _overlap(a, b):
    return all(max(sa.start, sb.start) < min(sa.stop, sb.stop) for sa, sb in ((a.x, b.x), (a.y, b.y), (a.z, b.z)))


class _IntensityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, permittivity, problem):
        J, grad = problem.gradient(permittivity)
        ctx.save_for_backward(grad.to(device=permittivity.device, dtype=permittivity.dtype))
        return torch.tensor(J, dtype=torch.float64, device=permittivity.device)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ.to(grad.dtype) * grad, None


def time_integrated_intensity(permittivity, grid, design, box, steps, checkpoint="sqrt"):
    """J = sum over `steps` steps from rest of sum over `box` of |E|^2, as a
    differentiable function of the design object's permittivity (a torch
    tensor, typically with requires_grad=True); backward() yields dJ/d
    permittivity from the transposed HIP steps."""
    problem = IntensityGradient(grid, design, box, steps, checkpoint)
    perm = permittivity if torch.is_tensor(permittivity) else torch.as_tensor(permittivity, dtype=torch.float64)
    return _IntensityFunction.apply(perm, problem)
