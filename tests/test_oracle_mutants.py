"""The oracle's pins catch plausible slips: each mutant below is a copy of
oracle/sph_oracle.cpp with one deliberate mistake of the kind a restatement
of the reference could contain (a swapped h_i / h_j, a dropped factor, a tie
broken the other way, a wrong reset value, an edge made non-strict ...),
built into a temporary folder with the oracle's own flags; the oracle's pin
tests (test_oracle_pins.py, test_oracle_physics.py, test_oracle_math.py) run
against each mutant in a separate process and must fail. The mistakes are
planted in the oracle only, on the CPU; the product library is not involved.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "sph_oracle.cpp"
PIN_TESTS = ["tests/test_oracle_pins.py", "tests/test_oracle_physics.py", "tests/test_oracle_math.py"]

# name: (text of the oracle, replacement); each text occurs in the oracle
MUTANTS = {
    "viscosity nu with h_j": ("const T nu = (((T(2) * p.alpha) * hi) * p.c0) / (rhoi + rhoj);",
                              "const T nu = (((T(2) * p.alpha) * hj) * p.c0) / (rhoi + rhoj);"),
    "viscosity softening with h_j": (
        "const T piv = ((-nu) * comp) / (d2 + ((eps * hi) * hi));\n                    const V4<T> dir",
        "const T piv = ((-nu) * comp) / (d2 + ((eps * hj) * hj));\n                    const V4<T> dir"),
    "viscosity factor 2 dropped": ("const T nu = (((T(2) * p.alpha) * hi) * p.c0)", "const T nu = ((p.alpha * hi) * p.c0)"),
    "surface tension W with h_i": ("const T s = mj * wendland(d, hj);", "const T s = mj * wendland(d, hi);"),
    "pressure term W' with h_i": ("const T dw = wendlandDerivative(d, hj);", "const T dw = wendlandDerivative(d, hi);"),
    "density W with h_i": ("density = density + m[j] * wendland(d, hj);", "density = density + m[j] * wendland(d, hi);"),
    "boundary friction factor 2 dropped": ("const T nu = ((f * hi) * p.c0) / (T(2) * rhoi);",
                                           "const T nu = ((f * hi) * p.c0) / rhoi;"),
    "boundary density with h0": ("density = density + bpsi[b] * wendland(d, hi);",
                                 "density = density + bpsi[b] * wendland(d, p.h0 * T(1.1));"),
    "boundary pressure direction flipped": ("abp.x = abp.x - s * dir.x;", "abp.x = abp.x + s * dir.x;"),
    "merge tie replaces the first hit": ("d2 < hmin * hmin && d2 < closest && m[j] < massLimit",
                                         "d2 < hmin * hmin && d2 <= closest && m[j] < massLimit"),
    "merge candidates reset to noCandidate": ("std::vector<uint32_t> cand(n, 0U);",
                                              "std::vector<uint32_t> cand(n, noCandidate);"),
    "merged position not mass-weighted": ("const V4<T> np {(ma * pa.x + mb * pb.x) / M,",
                                          "const V4<T> np {(pa.x + pb.x) / T(2),"),
    "merged radius ratio inverted": ("rad[a] = p.r0 * Const<T>::cbrt(M / p.m0);",
                                     "rad[a] = p.r0 * Const<T>::cbrt(p.m0 / M);"),
    "split count without the block round-up": ("const uint32_t launched = ((toSplit + tpb - 1U) / tpb) * tpb;",
                                               "const uint32_t launched = toSplit;"),
    "split capacity check off by one": ("if (static_cast<uint64_t>(n) + 12U > ref.maxCount)",
                                        "if (static_cast<uint64_t>(n) + 12U >= ref.maxCount)"),
    "vorticity curl sign": ("const V3<T> om {dzy - dyz, dxz - dzx, dyx - dxy};",
                            "const V3<T> om {dzy - dyz, dxz - dzx, dyx + dxy};"),
    "vorticity W' with h_j": ("const T gw = wendlandDerivative(dist, hh);",
                              "const T gw = wendlandDerivative(dist, h[j]);"),
    "interface split with the merge ratio": ("const V3<T> thr {ref.ifSplit * (p.dmax.x - p.dmin.x)",
                                             "const V3<T> thr {ref.ifMerge * (p.dmax.x - p.dmin.x)"),
    "split-in-domain test non-strict": ("return !(q.x - mo < p.dmin.x", "return !(q.x - mo <= p.dmin.x"),
    "velocity merge threshold non-strict": ("if (mag < ref.velMerge)", "if (mag <= ref.velMerge)"),
}


def _build(name_text, tmp: Path) -> Path:
    name, (old, new) = name_text
    src = SRC.read_text()
    assert src.count(old) >= 1, f"mutant {name!r}: text not in the oracle"
    d = tmp / f"m{abs(hash(name)) % 10**8}"
    d.mkdir()
    (d / "sph_oracle.cpp").write_text(src.replace(old, new))
    lib = d / "liboracle.so"
    cmd = [os.environ.get("CXX", "g++"), "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", "-iquote", str(ROOT / "oracle"), "-shared", "-o", str(lib), str(d / "sph_oracle.cpp")]
    subprocess.run(cmd, check=True, capture_output=True)
    return lib


def _pins_fail(lib: Path) -> bool:
    env = dict(os.environ, ASPH_ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *PIN_TESTS],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode == 1  # 1: tests ran and failed (not a collection or usage error)


def test_every_mutant_fails_a_pin(tmp_path):
    items = list(MUTANTS.items())
    workers = max(1, min(8, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(workers) as ex:
        libs = list(ex.map(lambda it: _build(it, tmp_path), items))
        failed = list(ex.map(_pins_fail, libs))
    missed = [name for (name, _), f in zip(items, failed) if not f]
    assert not missed, f"mutants not caught by the pins: {missed}"


def test_unmutated_copy_passes(tmp_path):
    """The same build recipe without a mutation passes the pins (so a mutant's
    failure is the mutation's doing)."""
    lib = _build(("none", ("#include <omp.h>", "#include <omp.h>")), tmp_path)
    env = dict(os.environ, ASPH_ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *PIN_TESTS], cwd=ROOT,
                       env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
