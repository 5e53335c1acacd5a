"""Builds the gfx950 shared library libomega_swe.so in-tree with hipcc.

Every .hip / .cpp file under csrc/ is compiled to an object in _build/ and the
objects are linked into omega_playground_amd/libomega_swe.so.  The flags keep
every floating-point operation separately rounded (-ffp-contract=off), which
the bit-for-bit parity with the reference's arithmetic depends on.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libomega_swe.so")
ARCH = os.environ.get("OMEGA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
]
_NOISE = "'-packed-fp32-ops' is not a recognized feature"


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if _NOISE not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    if msg.strip():
        sys.stderr.write(msg + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


REPO = os.path.dirname(HERE)
CPP_TESTS = os.path.join(REPO, "tests", "cpp")


def build_cpp_programs(verbose: bool = False) -> list:
    """Compiles the C++ programs written against the drop-in C++ API
    (tests/cpp/*.cpp) with g++ into tests/cpp/bin/, linked to
    libomega_swe.so through a relative rpath."""
    out_dir = os.path.join(CPP_TESTS, "bin")
    os.makedirs(out_dir, exist_ok=True)
    built = []
    for src in sorted(glob.glob(os.path.join(CPP_TESTS, "*.cpp"))):
        exe = os.path.join(out_dir, os.path.splitext(os.path.basename(src))[0])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Wno-unused-parameter",
               "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "cpp"),
               src, "-o", exe, "-L", HERE, "-lomega_swe",
               "-Wl,-rpath,$ORIGIN/../../../omega_playground_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}{r.stderr}")
        if r.stderr.strip():
            sys.stderr.write(r.stderr)
        built.append(exe)
        if verbose:
            print(f"built {exe}")
    return built


if __name__ == "__main__":
    build(verbose=True)
