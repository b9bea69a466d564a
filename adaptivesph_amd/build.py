"""In-tree build of the HIP library (adaptivesph_amd/lib/libasph.so) for gfx950.

Every source under csrc/ is compiled by hipcc as HIP with IEEE FP32 division
and square root, no FMA contraction and no packed-FP32 instructions, then
linked into one shared library exporting the C ABI of include/*.h.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "lib" / "obj"
LIB = PKG / "lib" / "libasph.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ASPH_OFFLOAD_ARCH", "gfx950")

SOURCES = ["sort.hip", "sph.hip", "walls.hip", "engine.cpp", "boundary.cpp", "abi.cpp"]

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> Path:
    out = OBJ / (src + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn.strip():
        print(f"[build] {src}: {warn[:4000]}", file=sys.stderr)
    return out


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB} ({LIB.stat().st_size // 1024} KiB)")
    return LIB


def _host_program(src: Path, out: Path, verbose: bool) -> Path:
    root = PKG.parent
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", str(root / "include"),
           str(src), "-L", str(LIB.parent), "-lasph", "-Wl,-rpath,$ORIGIN", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build of {src.name} failed:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        print(f"[build] {src.name}: {r.stderr[:4000]}", file=sys.stderr)
    if verbose:
        print(f"[build] {out}")
    return out


def build_runner(verbose: bool = True) -> Path:
    """The headless runner (runner/asph_run.cpp) over the C++ interface."""
    return _host_program(PKG / "runner" / "asph_run.cpp", LIB.parent / "asph_run", verbose)


def build_cpp_test(verbose: bool = True) -> Path:
    """The C++ interface test program (tests/cpp/facade_test.cpp), compiled by
    the host compiler against include/ and linked to the library."""
    return _host_program(PKG.parent / "tests" / "cpp" / "facade_test.cpp", LIB.parent / "facade_test", verbose)


if __name__ == "__main__":
    build()
    build_runner()
    build_cpp_test()
