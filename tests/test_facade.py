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


def _read_imported(path):
    raw = path.read_bytes()
    o = 0
    n = int(np.frombuffer(raw[o:o + 4], np.uint32)[0])
    o += 4
    f = np.frombuffer(raw[o:o + 36 * n], np.float32)
    o += 36 * n
    x, v, rho = f[: 4 * n].reshape(n, 4), f[4 * n: 8 * n].reshape(n, 4), f[8 * n: 9 * n]
    nb = int(np.frombuffer(raw[o:o + 4], np.uint32)[0])
    o += 4
    g = np.frombuffer(raw[o:o + 36 * nb], np.float32)
    bx, br, bc = g[: 4 * nb].reshape(nb, 4), g[4 * nb: 5 * nb], g[5 * nb: 9 * nb].reshape(nb, 4)
    return x, v, rho, bx, br, bc


@pytest.mark.gpu
@pytest.mark.parametrize("adaptive", [False, True])
def test_cpp_imported_memory_equals_c_abi(tmp_path, adaptive):
    """The reference's 8-argument createSimulation with imported particle
    memory, called as the reference application calls it: the renderer's
    buffers hold the same state, bit for bit, as the C ABI's own simulation,
    and the boundary buffers hold the walls' positions, radii and colours in
    the domain's order."""
    s = scene.dam_break((10, 12, 8), 0.01, (0.3, 0.25, 0.2), seed=31, steps=12, name="imported")
    if adaptive:
        r = abi.default_refinement()
        r.enabled = 1
        r.max_particle_count = 3 * s.count
        r.initial_cooldown = 0
        r.cooldown = 2
        r.criterion_type = abi.CRITERION["zone"]
        r.zone = abi.ZoneParameters(abi.Vec3(0.05, 0.06, 0.04), 0.04, 0.0, 0.0)
        s.refinement = r
        s.capacity = r.max_particle_count
    pfile, xfile, rfile, out = (tmp_path / n for n in ("p.bin", "x.bin", "r.bin", "out.bin"))
    pfile.write_bytes(bytes(s.params))
    xfile.write_bytes(np.ascontiguousarray(s.positions, np.float32).tobytes())
    args = [str(_bin()), "imported", str(pfile), str(xfile), str(s.steps), repr(float(np.float32(s.dt))), str(out)]
    if adaptive:
        rfile.write_bytes(bytes(s.refinement))
        args.append(str(rfile))
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    x, v, rho, bx, br, bc = _read_imported(out)

    sim = Simulation.from_scene(s)
    for _ in range(s.steps):
        sim.update(s.dt)
    ref = sim.fluid()
    sim.close()
    assert x.shape[0] == ref["positions"].shape[0]
    if adaptive:
        assert x.shape[0] != s.count  # refinement changed the count
    for name, a, b in (("positions", x, ref["positions"]), ("velocities", v, ref["velocities"]),
                       ("densities", rho, ref["densities"])):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    walls = s.boundary
    assert bx.shape[0] == len(walls)
    assert np.array_equal(bx, np.ascontiguousarray(walls["position"]))
    assert (br == np.float32(s.params.base_particle_radius)).all()
    assert (bc == np.array([0.3, 0.3, 0.3, 0.2], np.float32)).all()
