"""The kernels on the CPU: fdtd_amd/csrc/fdtd_hip.hip built unchanged with g++
and AddressSanitizer against a host stand-in for the HIP runtime
(tests/cpu_kernels), each workgroup's work-items run as lockstep fibers.

Every golden scenario the reference wrote is replayed through the package's
own Python layer and must match bit for bit (fields, CPML state, detector and
source traces), as on the GPU; random scenarios, real and complex, match the
oracle; the compensated form stays within its bound; batched runs (direct,
graph replay, persistent, and the static-table entry point) equal the grids
stepped alone; the transposed steps give the reference-autograd gradients of
both gradient fixtures (all PML faces; periodic z with an anisotropic
object), and the guard-free transposed H kernel of tiles that cover the
(y, z) plane equals the guarded one bit for bit; the non-finite check flags a NaN or an infinity in E or H and
nothing else.
Both tile shapes of the main kernels run: the default (flat tiles for these
grids, whose nz is not a multiple of 64) and row tiles forced with
FDTD_HIP_FLAT=0; and the padded plane layout, forced on these small grids with
FDTD_HIP_PITCH=1 (every x plane whose ny*nz is not a multiple of 32 padded to
one), must give the same bits. Any out-of-bounds access aborts the child
process with an AddressSanitizer report.
"""

import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def _cases():
    import scenarios

    out = [f"fixture:{sc['name']}" for sc in scenarios.all_scenarios()]
    out += [f"random:{s}" for s in range(6)]
    out += [f"random_complex:{s}" for s in range(6)]
    out += ["compensated:g2", "compensated:periodic_x_first", "batch:1", "batch:2", "adjoint:g5", "adjoint:pa",
            "adjoint_full:hip.float64", "adjoint_full:hip.float32", "adjoint_pitch:TINY_PA", "adjoint_pitch:TINY_ABSORBER",
            "nonfinite:g2", "nonfinite:complex_object"]
    return out


@pytest.fixture(scope="module")
def host_library():
    sys.path.insert(0, os.path.join(ROOT, "tests", "cpu_kernels"))
    import build

    return build.build(), build.asan_runtime()


@pytest.mark.parametrize("layout", ["default", "row_tiles", "padded_planes"])
def test_kernels_on_the_host_match_fixtures_and_oracle(host_library, layout):
    lib, asan = host_library
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    env.pop("FDTD_HIP_PITCH", None)
    if layout == "row_tiles":
        env["FDTD_HIP_FLAT"] = "0"
    elif layout == "padded_planes":
        env["FDTD_HIP_PITCH"] = "1"
    cases = _cases()
    res = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "cpu_kernels", "replay.py")] + cases,
                         capture_output=True, text=True, env=env, timeout=900)
    lines = [json.loads(x) for x in res.stdout.splitlines() if x.startswith("{")]
    assert res.returncode == 0, res.stderr[-4000:]
    assert [r["case"] for r in lines] == cases
    bad = [r for r in lines if not r["ok"]]
    assert not bad, bad
    assert all(r["launches"] > 0 for r in lines)
