"""Compile libfdtd_hip.so for gfx950 in-tree (fdtd_amd/libfdtd_hip.so).

    python -m fdtd_amd.csrc.build          # or build() from Python

hipcc cross-compiles without a GPU. Flags:
  -ffp-contract=off    no fused multiply-add: every float64 operation rounds
                       like the numpy reference, so results match bit for bit
  -target-feature -packed-fp32-ops
                       keep the compiler from emitting packed fp32 VALU ops
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
SOURCES = [os.path.join(HERE, "fdtd_hip.hip")]
OUTPUT = os.path.join(PKG, "libfdtd_hip.so")
ARCH = os.environ.get("FDTD_AMD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++20",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
    ]


def build(verbose=False):
    tmp = OUTPUT + ".tmp"
    cmd = [hipcc()] + flags() + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libfdtd_hip.so")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
