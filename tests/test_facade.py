"""The C++ interface (include/sph/Simulation.hpp), driven by a C++ program
the way the reference's application drives the reference: host-side checks
on the CPU, and on the GPU a run whose result must equal, bit for bit, the
same scene stepped through the C ABI from Python."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from adaptivesph_amd import abi, scene
from adaptivesph_amd.simulation import Simulation

BIN = Path(abi.PKG_DIR) / "lib" / "facade_test"


def _bin() -> Path:
    if not BIN.exists():
        raise FileNotFoundError(f"{BIN} is not built (run __graft_entry__.build())")
    return BIN


def test_cpp_interface_host_checks():
    r = subprocess.run([str(_bin()), "host"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host ok" in r.stdout


@pytest.mark.gpu
def test_cpp_interface_run_equals_c_abi(tmp_path):
    s = scene.dam_break((10, 12, 8), 0.01, (0.3, 0.25, 0.2), seed=31, steps=15, name="facade")
    pfile = tmp_path / "params.bin"
    xfile = tmp_path / "positions.bin"
    out = tmp_path / "out.bin"
    pfile.write_bytes(bytes(s.params))
    xfile.write_bytes(np.ascontiguousarray(s.positions, np.float32).tobytes())
    dt = np.float32(s.dt)
    r = subprocess.run([str(_bin()), "run", str(pfile), str(xfile), str(s.steps), repr(float(dt)), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:4], np.uint32)[0])
    f = np.frombuffer(raw[4:], np.float32)
    x = f[: 4 * n].reshape(n, 4)
    v = f[4 * n: 8 * n].reshape(n, 4)
    rho = f[8 * n: 9 * n]

    sim = Simulation.from_scene(s)
    for _ in range(s.steps):
        sim.update(s.dt)
    ref = sim.fluid()
    sim.close()
    assert n == s.count
    for name, a, b in (("positions", x, ref["positions"]), ("velocities", v, ref["velocities"]),
                       ("densities", rho, ref["densities"])):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
