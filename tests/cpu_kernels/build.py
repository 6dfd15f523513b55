"""Build the host (CPU) copy of libfdtd_hip used by tests/test_cpu_kernels.py.

g++ compiles fdtd_amd/csrc/fdtd_hip.hip, unchanged, as C++ against the host
stand-in for <hip/hip_runtime.h> in this directory, with AddressSanitizer:
every kernel runs on the host with each workgroup's work-items as lockstep
fibers, and every memory access the kernels make is bounds-checked. The
library has the C ABI of the real one, so the product's Python layer can
drive it through ctypes with host arrays.

Output: tests/cpu_kernels/_build/libfdtd_cpu.so (rebuilt when a source is
newer). Loading an ASan library needs the ASan runtime preloaded, so the
tests run it in a child process with LD_PRELOAD=asan_runtime().
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "fdtd_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "_build")
OUTPUT = os.path.join(OUT_DIR, "libfdtd_cpu.so")
WRAPPER = os.path.join(HERE, "fdtd_hip_cpu.cpp")
SOURCES = [WRAPPER, os.path.join(CSRC, "fdtd_hip.hip"), os.path.join(CSRC, "fdtd_kernels.hpp"),
           os.path.join(INCLUDE, "fdtd_hip.h"), os.path.join(HERE, "hip", "hip_runtime.h"), __file__]


def asan_runtime():
    out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    return out.stdout.strip()


def flags():
    return ["-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-omit-frame-pointer",
            "-fsanitize=address", "-Wno-attributes", "-Wno-ignored-attributes",
            "-I", HERE, "-I", INCLUDE, "-I", CSRC]


def build(force=False):
    newest = max(os.path.getmtime(s) for s in SOURCES)
    if not force and os.path.exists(OUTPUT) and os.path.getmtime(OUTPUT) >= newest:
        return OUTPUT
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUTPUT + ".tmp"
    cmd = ["g++"] + flags() + [WRAPPER, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the host copy of libfdtd_hip")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
