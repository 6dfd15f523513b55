"""ctypes binding of libfdtd_hip.so (C ABI: include/fdtd_hip.h).

The library is loaded after torch so that it binds to the HIP runtime torch
already mapped (same SONAME); a second HIP runtime in the process would make
torch's stream handles meaningless to the library, so that case is refused.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_hip.so")

F64, F32, F32C, C128, C64 = 0, 1, 2, 3, 4
OBJ_PLAIN, OBJ_ABSORBING, OBJ_ANISOTROPIC = 0, 1, 2
OBJ_REAL_ARRAYS = 16  # flag: the object's arrays are real in the reference
SRC_POINT, SRC_LINE, SRC_PLANE = 0, 1, 2

_lib = None

_c_int, _c_i64, _c_dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pdbl = ctypes.POINTER(ctypes.c_double)
_pi32 = ctypes.POINTER(ctypes.c_int32)

_SIGNATURES = {
    "fdtd_version": ([], ctypes.c_char_p),
    "fdtd_last_error": ([], ctypes.c_char_p),
    "fdtd_plan_create": ([_c_int, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)], _c_int),
    "fdtd_plan_destroy": ([_vp], _c_int),
    "fdtd_plan_set_pitch": ([_vp, _c_i64], _c_int),
    "fdtd_plan_set_compensation": ([_vp, _vp, _vp], _c_int),
    "fdtd_plan_set_material_palette": ([_vp, _vp, _vp, _vp, _vp], _c_int),
    "fdtd_plan_add_pml": ([_vp, _c_int, _c_int, _c_int, _pdbl, _pdbl, _pdbl, _pdbl, _vp, _vp], _c_int),
    "fdtd_plan_add_periodic": ([_vp, _c_int], _c_int),
    "fdtd_plan_add_object": ([_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp], _c_int),
    "fdtd_plan_add_source": ([_vp, _c_int, _c_int, _pi64, _pdbl, _c_dbl, _c_int, _c_int], _c_int),
    "fdtd_plan_set_waveforms": ([_vp, _vp, _c_i64], _c_int),
    "fdtd_plan_add_detector": ([_vp, _c_int, _pi64, _vp, _vp, _c_i64, _c_i64, _c_i64], _c_int),
    "fdtd_plan_finalize": ([_vp], _c_int),
    "fdtd_plan_detector_counts": ([_vp, _c_int, _pi64, _pi64], _c_int),
    "fdtd_plan_add_current_detector": ([_vp, _c_int, _pi32, _pi32, _pi32, _c_dbl, _vp, _c_i64, _c_i64], _c_int),
    "fdtd_plan_add_soft_source": ([_vp, _c_i64, _c_dbl, _c_dbl, _c_int, _vp, _vp, _c_i64, _c_i64], _c_int),
    "fdtd_plan_current_detector_count": ([_vp, _c_int, _pi64], _c_int),
    "fdtd_plan_soft_source_count": ([_vp, _c_int, _pi64], _c_int),
    "fdtd_half_step_E": ([_vp, _c_i64, _vp], _c_int),
    "fdtd_half_step_H": ([_vp, _c_i64, _vp], _c_int),
    "fdtd_run": ([_vp, _c_i64, _c_i64, _vp], _c_int),
    "fdtd_check_finite": ([_vp, _vp, _vp], _c_int),
    "fdtd_bench_main_kernel": ([_vp, _c_i64, _vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
                               _c_int),
    "fdtd_bench_copy": ([_vp, _vp, _c_i64, _c_i64, _vp, ctypes.POINTER(ctypes.c_float)], _c_int),
    "fdtd_batch_create": ([_c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)], _c_int),
    "fdtd_batch_destroy": ([_vp], _c_int),
    "fdtd_batch_set_launch": ([_vp, _c_int, _c_int, _c_int], _c_int),
    "fdtd_batch_work_items": ([_vp, _pi64], _c_int),
    "fdtd_batch_run": ([_vp, _c_i64, _c_i64, _vp], _c_int),
    "fdtd_batch_run_graph": ([_vp, _c_i64, _c_i64, _c_i64, _vp], _c_int),
    "fdtd_batch_run_static": ([_c_int, ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_i64, _c_i64, _vp], _c_int),
    "fdtd_adjoint_psi_size": ([_vp, _pi64], _c_int),
    "fdtd_adjoint_create": ([_vp, _vp, _vp, _vp, _vp, _vp, _pi32, _pi32, _vp, ctypes.POINTER(_vp)], _c_int),
    "fdtd_adjoint_destroy": ([_vp], _c_int),
    "fdtd_run_recorded": ([_vp, _c_i64, _c_i64, _vp, _vp, _vp], _c_int),
    "fdtd_adjoint_run": ([_vp, _c_i64, _vp, _vp, _vp], _c_int),
    "fdtd_adjoint_set_launch": ([_vp, _c_int], _c_int),
}


def bench_copy(dst, src, nbytes, reps, stream):
    """Milliseconds for `reps` plain device copies of nbytes (measurement hook)."""
    ms = ctypes.c_float()
    check(lib().fdtd_bench_copy(dst.data_ptr(), src.data_ptr(), int(nbytes), int(reps), stream, ctypes.byref(ms)))
    return ms.value


def batch_run_static(plans, block_threads, cx, workgroups, q0, nsteps, stream):
    """fdtd_batch_run_static over finalized Plans: a batched run whose tables
    live in the library's static device array (no device allocation)."""
    arr = (_vp * len(plans))(*[p._h.value for p in plans])
    check(lib().fdtd_batch_run_static(len(plans), arr, int(block_threads), int(cx), int(workgroups), int(q0),
                                      int(nsteps), stream))


def exported_symbols():
    return sorted(_SIGNATURES)


def _hip_runtimes_mapped():
    paths = set()
    try:
        with open("/proc/self/maps") as fh:
            for line in fh:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return paths


def lib():
    """The loaded library; raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (maps torch's HIP runtime first)

    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m fdtd_amd.csrc.build` "
            "(or __graft_entry__.build()); the HIP backend has no other implementation")
    L = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    rts = _hip_runtimes_mapped()
    if len(rts) > 1:
        raise RuntimeError(f"two HIP runtimes are mapped in this process: {sorted(rts)}")
    _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"libfdtd_hip: {lib().fdtd_last_error().decode()} (code {rc})")


def _i64_array(values):
    arr = (ctypes.c_int64 * len(values))(*[int(v) for v in values])
    return arr


def _dbl_array(values):
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr


class Plan:
    """Owner of one FdtdPlan*; every device buffer it points at is a torch
    tensor that the caller keeps alive for the plan's lifetime."""

    def __init__(self, dtype_code, shape, courant, E, H, inv_eps, inv_mu, pitch=None):
        L = lib()
        self._L = L
        h = _vp()
        check(L.fdtd_plan_create(dtype_code, shape[0], shape[1], shape[2], float(courant),
                                 E.data_ptr(), H.data_ptr(), inv_eps.data_ptr(), inv_mu.data_ptr(),
                                 ctypes.byref(h)))
        self._h = h
        self._keep = [E, H, inv_eps, inv_mu]
        if pitch is not None and pitch != shape[1] * shape[2]:
            check(L.fdtd_plan_set_pitch(h, int(pitch)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.fdtd_plan_destroy(h)
            self._h = None

    def set_compensation(self, E_lo, H_lo):
        check(self._L.fdtd_plan_set_compensation(self._h, E_lo.data_ptr(), H_lo.data_ptr()))
        self._keep += [E_lo, H_lo]

    def set_material_palette(self, eps_id, eps_tab, mu_id, mu_tab):
        check(self._L.fdtd_plan_set_material_palette(self._h, eps_id.data_ptr(), eps_tab.data_ptr(),
                                                     mu_id.data_ptr(), mu_tab.data_ptr()))
        self._keep += [eps_id, eps_tab, mu_id, mu_tab]

    def add_pml(self, axis, high, thickness, bE, cE, bH, cH, psiE, psiH):
        check(self._L.fdtd_plan_add_pml(self._h, axis, int(high), thickness, _dbl_array(bE), _dbl_array(cE),
                                        _dbl_array(bH), _dbl_array(cH), psiE.data_ptr(), psiH.data_ptr()))
        self._keep += [psiE, psiH]

    def add_periodic(self, axis):
        check(self._L.fdtd_plan_add_periodic(self._h, axis))

    def add_object(self, kind, box, inv, absorption=None):
        (x0, x1), (y0, y1), (z0, z1) = box
        check(self._L.fdtd_plan_add_object(self._h, kind, x0, x1, y0, y1, z0, z1, inv.data_ptr(),
                                           absorption.data_ptr() if absorption is not None else None))
        self._keep += [inv] + ([absorption] if absorption is not None else [])

    def add_source(self, kind, cells, profile, amplitude, comp_E, comp_H):
        prof = _dbl_array(profile) if profile is not None else None
        check(self._L.fdtd_plan_add_source(self._h, kind, len(cells), _i64_array(cells), prof, float(amplitude),
                                           comp_E, comp_H))

    def set_waveforms(self, table):
        check(self._L.fdtd_plan_set_waveforms(self._h, table.data_ptr() if table is not None else None,
                                              0 if table is None else table.shape[1]))
        self._wave = table

    def add_detector(self, cells, bufE, bufH, count_E, count_H):
        check(self._L.fdtd_plan_add_detector(self._h, len(cells), _i64_array(cells), bufE.data_ptr(),
                                             bufH.data_ptr(), bufE.shape[0], count_E, count_H))
        self._keep += [bufE, bufH]

    def add_current_detector(self, px, py, pz, spacing, buf, count):
        n = len(px)
        arr = lambda v: (ctypes.c_int32 * n)(*[int(a) for a in v])
        check(self._L.fdtd_plan_add_current_detector(self._h, n, arr(px), arr(py), arr(pz), float(spacing),
                                                     buf.data_ptr(), buf.shape[0], count))
        self._keep.append(buf)

    def add_soft_source(self, cell, impedance, spacing, detector, vin, vout, count):
        check(self._L.fdtd_plan_add_soft_source(self._h, int(cell), float(impedance), float(spacing), int(detector),
                                                vin.data_ptr(), vout.data_ptr(), vin.shape[0], count))
        self._keep += [vin, vout]

    def finalize(self):
        check(self._L.fdtd_plan_finalize(self._h))

    def detector_counts(self, index):
        e, h = ctypes.c_int64(), ctypes.c_int64()
        check(self._L.fdtd_plan_detector_counts(self._h, index, ctypes.byref(e), ctypes.byref(h)))
        return e.value, h.value

    def half_step_E(self, q, stream):
        check(self._L.fdtd_half_step_E(self._h, q, stream))

    def half_step_H(self, q, stream):
        check(self._L.fdtd_half_step_H(self._h, q, stream))

    def run(self, q0, nsteps, stream):
        check(self._L.fdtd_run(self._h, q0, nsteps, stream))

    def check_finite(self, flag, stream):
        """OR 1 into the device int32 tensor `flag` if an E or H value is
        non-finite (queued on `stream`)."""
        check(self._L.fdtd_check_finite(self._h, flag.data_ptr(), stream))

    def bench_main_kernel(self, reps, stream):
        """Average milliseconds of the E and H main kernels over `reps` steps."""
        e, h = ctypes.c_float(), ctypes.c_float()
        check(self._L.fdtd_bench_main_kernel(self._h, reps, stream, ctypes.byref(e), ctypes.byref(h)))
        return e.value, h.value


class Adjoint:
    """Owner of one FdtdAdjoint* over a Plan (reverse-mode steps for the
    objective sum_steps sum_box |E|^2); the arrays are torch tensors kept
    alive here."""

    def __init__(self, plan, gcoef, dcoef, lamE, lamH, psi_adj, design_box, objective_box, grad):
        L = plan._L
        self._L = L
        h = _vp()
        db = (ctypes.c_int32 * 6)(*[int(v) for v in design_box])
        ob = (ctypes.c_int32 * 6)(*[int(v) for v in objective_box])
        check(L.fdtd_adjoint_create(plan._h, gcoef.data_ptr(), dcoef.data_ptr() if dcoef is not None else None,
                                    lamE.data_ptr(), lamH.data_ptr(),
                                    psi_adj.data_ptr() if psi_adj is not None and psi_adj.numel() else None,
                                    db, ob, grad.data_ptr(), ctypes.byref(h)))
        self._h = h
        self._keep = [plan, gcoef, dcoef, lamE, lamH, psi_adj, grad]

    @staticmethod
    def psi_size(plan):
        n = ctypes.c_int64()
        check(plan._L.fdtd_adjoint_psi_size(plan._h, ctypes.byref(n)))
        return n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.fdtd_adjoint_destroy(h)
            self._h = None

    def run_recorded(self, q0, nsteps, curl_records, plane_records, stream):
        check(self._L.fdtd_run_recorded(self._h, q0, nsteps,
                                        curl_records.data_ptr() if curl_records is not None else None,
                                        plane_records.data_ptr(), stream))

    def set_launch(self, cx):
        """x planes per workgroup of the transposed kernels (0: the plan's)."""
        check(self._L.fdtd_adjoint_set_launch(self._h, int(cx)))

    def run_adjoint(self, nsteps, curl_records, plane_records, stream):
        check(self._L.fdtd_adjoint_run(self._h, nsteps, curl_records.data_ptr(), plane_records.data_ptr(), stream))


class Batch:
    """Owner of one FdtdBatch* over finalized Plans (kept alive here)."""

    def __init__(self, plans):
        L = plans[0]._L
        self._L = L
        arr = (_vp * len(plans))(*[p._h.value for p in plans])
        h = _vp()
        check(L.fdtd_batch_create(len(plans), arr, ctypes.byref(h)))
        self._h = h
        self._plans = list(plans)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.fdtd_batch_destroy(h)
            self._h = None

    def set_launch(self, block_threads, cx, workgroups):
        check(self._L.fdtd_batch_set_launch(self._h, int(block_threads), int(cx), int(workgroups)))

    def work_items(self):
        n = ctypes.c_int64()
        check(self._L.fdtd_batch_work_items(self._h, ctypes.byref(n)))
        return n.value

    def run(self, q0, nsteps, stream):
        check(self._L.fdtd_batch_run(self._h, q0, nsteps, stream))

    def run_graph(self, q0, nsteps, steps_per_graph, stream):
        check(self._L.fdtd_batch_run_graph(self._h, q0, nsteps, steps_per_graph, stream))
