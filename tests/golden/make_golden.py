"""Write the golden fixtures under tests/golden/ by running the reference
package (flaport/fdtd 0.3.6, numpy fp64 backend) on the scenarios of
scenarios.py.

Run once, in the build container only (the reference tree is not available
anywhere else):

    python tests/golden/make_golden.py /root/reference [scenario ...]

Outputs (data only: inputs are regenerated from scenarios.py and checked
against the stored checksum):

  <scenario>.npz   final E/H, detector traces, live PML psi components, PML
                   coefficient profiles and the resolved registration data
  curl_golden.npz  the curl known-answer vectors held in the reference's
                   tests/test_grid.py (test_curl_E / test_curl_H)
"""

import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import scenarios  # noqa: E402

_NORMAL_PSI = {
    # axis -> (E-side live psi, H-side live psi), names follow the reference
    # PML attributes psi_<F><comp>[..., axis]
    0: (("psi_Ey", "psi_Ez"), ("psi_Hy", "psi_Hz")),
    1: (("psi_Ex", "psi_Ez"), ("psi_Hx", "psi_Hz")),
    2: (("psi_Ex", "psi_Ey"), ("psi_Hx", "psi_Hy")),
}
_FACE_AXIS = {"_PMLXlow": 0, "_PMLXhigh": 0, "_PMLYlow": 1, "_PMLYhigh": 1, "_PMLZlow": 2, "_PMLZhigh": 2}


def _profile(arr, axis):
    """1-D profile of a (.., 3) PML coefficient array along its normal axis."""
    a = np.asarray(arr)
    idx = [0, 0, 0]
    idx[axis] = slice(None)
    prof = a[tuple(idx) + (axis,)]
    # the profile must be uniform across the two tangential axes
    shp = [1, 1, 1]
    shp[axis] = -1
    assert np.array_equal(a[..., axis], np.broadcast_to(prof.reshape(shp), a.shape[:3]))
    other = [c for c in range(3) if c != axis]
    return prof.copy(), float(a[0, 0, 0, other[0]])


def _slice_pair(s):
    return np.array([s.start, s.stop], dtype=np.int64)


def record(fdtd, sc):
    grid, comps = scenarios.build_fdtd(fdtd, sc)
    out = {"input_checksum": np.float64(scenarios.input_checksum(sc))}
    out["reg/Nxyz"] = np.array([grid.Nx, grid.Ny, grid.Nz])
    out["reg/courant_number"] = np.float64(grid.courant_number)
    out["reg/time_step"] = np.float64(grid.time_step)
    if sc["name"] != "c1":
        out["reg/grid_inverse_permittivity"] = np.asarray(grid.inverse_permittivity).copy()
    for i, c in enumerate(comps):
        p = f"reg/{i}/"
        cls = type(c).__name__
        out[p + "class"] = np.array(cls)
        if cls in _FACE_AXIS:
            ax = _FACE_AXIS[cls]
            out[p + "thickness"] = np.int64(c.thickness)
            for nm in ("bE", "cE", "bH", "cH"):
                prof, other = _profile(getattr(c, nm), ax)
                out[p + nm] = prof
                out[p + nm + "_other"] = np.float64(other)
        elif cls in ("Object", "AbsorbingObject", "AnisotropicObject"):
            out[p + "x"] = _slice_pair(c.x)
            out[p + "y"] = _slice_pair(c.y)
            out[p + "z"] = _slice_pair(c.z)
            if cls != "AnisotropicObject":
                out[p + "inverse_permittivity"] = np.asarray(c.inverse_permittivity).copy()
            if cls == "AbsorbingObject":
                out[p + "absorption_factor"] = np.asarray(c.absorption_factor).copy()
        elif cls == "PointSource":
            out[p + "xyz"] = np.array([c.x, c.y, c.z])
            out[p + "period"] = np.int64(c.period)
        elif cls == "LineSource":
            out[p + "x"] = np.array(c.x)
            out[p + "y"] = np.array(c.y)
            out[p + "z"] = np.array(c.z)
            out[p + "period"] = np.int64(c.period)
            out[p + "profile"] = np.asarray(c.profile).copy()
        elif cls == "PlaneSource":
            out[p + "x"] = _slice_pair(c.x)
            out[p + "y"] = _slice_pair(c.y)
            out[p + "z"] = _slice_pair(c.z)
            out[p + "pol"] = np.array([c._Epol, c._Hpol])
            out[p + "period"] = np.int64(c.period)
        elif cls in ("LineDetector", "BlockDetector", "CurrentDetector"):
            out[p + "x"] = np.array(c.x)
            out[p + "y"] = np.array(c.y)
            out[p + "z"] = np.array(c.z)
        elif cls == "SoftArbitraryPointSource":
            out[p + "xyz"] = np.array([c.x, c.y, c.z])

    scenarios.run_program(grid, sc)

    out["time_steps_passed"] = np.int64(grid.time_steps_passed)
    out["E"] = np.asarray(grid.E).copy()
    out["H"] = np.asarray(grid.H).copy()
    for i, c in enumerate(comps):
        cls = type(c).__name__
        p = f"out/{i}/"
        if cls in _FACE_AXIS:
            ax = _FACE_AXIS[cls]
            live_e, live_h = _NORMAL_PSI[ax]
            for nm in live_e + live_h:
                out[p + nm] = np.asarray(getattr(c, nm))[..., ax].copy()
            dead = 0.0
            for nm in ("psi_Ex", "psi_Ey", "psi_Ez", "psi_Hx", "psi_Hy", "psi_Hz"):
                a = np.asarray(getattr(c, nm)).copy()
                if nm in live_e + live_h:
                    a[..., ax] = 0.0
                dead = max(dead, float(np.max(np.abs(a))))
            out[p + "dead_psi_maxabs"] = np.float64(dead)
        elif cls == "LineDetector":
            out[p + "E"] = np.stack([np.asarray(v) for v in c.E]) if c.E else np.zeros((0,))
            out[p + "H"] = np.stack([np.asarray(v) for v in c.H]) if c.H else np.zeros((0,))
        elif cls == "BlockDetector":
            out[p + "E"] = np.array(c.E, dtype=np.float64)
            out[p + "H"] = np.array(c.H, dtype=np.float64)
        elif cls == "CurrentDetector":
            out[p + "I"] = np.array(c.I, dtype=np.float64)
        elif cls == "SoftArbitraryPointSource":
            out[p + "input_voltage"] = np.array(c.input_voltage, dtype=np.float64)
            out[p + "source_voltage"] = np.array(c.source_voltage, dtype=np.float64)
            out[p + "I"] = np.array(c.current_detector.I, dtype=np.float64)
            # the reference's frequency routines on this run (fourier.py:172-266)
            fr = fdtd.FrequencyRoutines(grid, c)
            f, z = fr.impedance()
            out[p + "impedance_freqs"], out[p + "impedance"] = np.asarray(f), np.asarray(z)
            hist = np.array(c.source_voltage, dtype=np.float64).reshape(-1)
            f, sp = fdtd.FrequencyRoutines(grid, hist).FFT()
            out[p + "fft_freqs"], out[p + "fft"] = np.asarray(f), np.asarray(sp)
    return out


def _np_array_literal(fn_node):
    """Return the literal list passed to np.array(...) inside a test function."""
    for node in ast.walk(fn_node):
        if isinstance(node, ast.Call) and getattr(node.func, "attr", None) == "array" and node.args:
            try:
                return ast.literal_eval(node.args[0])
            except ValueError:
                continue
    raise RuntimeError("no literal found")


def curl_golden(ref_root):
    path = os.path.join(ref_root, "tests", "test_grid.py")
    tree = ast.parse(open(path).read())
    fns = {n.name: n for n in tree.body if isinstance(n, ast.FunctionDef)}
    return {
        "E_in": np.random.RandomState(0).randn(3, 3, 3, 3),
        "curl_E": np.array(_np_array_literal(fns["test_curl_E"]), dtype=np.float64),
        "H_in": np.random.RandomState(1).randn(3, 3, 3, 3),
        "curl_H": np.array(_np_array_literal(fns["test_curl_H"]), dtype=np.float64),
    }


def main(ref_root, only=None):
    sys.path.insert(0, ref_root)
    import fdtd  # the reference package

    fdtd.set_backend("numpy")
    if not only:
        np.savez_compressed(os.path.join(HERE, "curl_golden.npz"), **curl_golden(ref_root))
    for sc in scenarios.all_scenarios():
        if only and sc["name"] not in only:
            continue
        out = record(fdtd, sc)
        fn = os.path.join(HERE, f"{sc['name']}.npz")
        np.savez_compressed(fn, **out)
        print(f"{fn}: {os.path.getsize(fn) / 1024:.0f} KiB")


if __name__ == "__main__":
    # make_golden.py [reference root] [scenario names...]: all scenarios by default
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", sys.argv[2:])
