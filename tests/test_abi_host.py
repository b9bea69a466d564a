"""Host-side checks of the C ABI (no GPU needed): every symbol declared in
include/*.h is exported, struct layouts agree with a C compiler, defaults
mirror the reference, argument validation fails before any device call, and
the product's boundary generator equals the oracle's restatement."""
import ctypes as C
import re
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest

from adaptivesph_amd import abi, scene
from conftest import HAVE_GPU
from oracle import oracle

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(asph_\w+)\s*\(", h.read_text(), re.M):
            names.add(m.group(1))
    return names


def test_all_declared_symbols_exported():
    lib = abi.load_library()
    names = _declared()
    assert len(names) >= 35
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert names <= exported, sorted(names - exported)
    for n in names:
        getattr(lib, n)


def test_struct_layout_matches_c(tmp_path):
    structs = {
        "asph_parameters": abi.Parameters,
        "asph_refinement_parameters": abi.RefinementParameters,
        "asph_density_info": abi.DensityInfo,
        "asph_boundary_particle": abi.BoundaryParticle,
        "asph_fluid_view": abi.FluidView,
        "asph_batch_config": abi.BatchConfig,
        "asph_refinement_trace": abi.RefinementTrace,
        "asph_domain": abi.Domain,
    }
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "asph_debug.h"', '#include "asph_batch.h"',
           "int main(void){"]
    for cname, py in structs.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"
    # the mirrored prefix of Simulation::Parameters is 80 bytes (Simulation.cuh:64-109)
    assert abi.Parameters.tait_gamma.offset == 80


def test_defaults_mirror_reference():
    lib = abi.load_library()
    p = abi.Parameters()
    assert lib.asph_default_parameters(C.byref(p)) == 0
    assert (p.domain.min.x, p.domain.max.y, p.domain.friction, p.threads_per_block) == (-1.0, 1.0,
                                                                                        pytest.approx(0.01), 256)
    assert p.tait_gamma == 5.0
    r = abi.RefinementParameters()
    assert lib.asph_default_refinement_parameters(C.byref(r)) == 0
    want = abi.default_refinement()
    for f, _ in abi.RefinementParameters._fields_:
        a, b = getattr(r, f), getattr(want, f)
        if isinstance(a, C.Structure):
            assert bytes(a) == bytes(b), f
        else:
            assert a == pytest.approx(b), f
    assert (r.min_mass_ratio, r.max_mass_ratio, r.max_particle_count, r.initial_cooldown, r.cooldown) == (
        pytest.approx(0.3), 8.0, 1000000, 10, 10)
    assert (r.splitting.epsilon, r.splitting.alpha, r.splitting.center_mass_ratio) == (
        pytest.approx(0.65), pytest.approx(0.7), pytest.approx(0.077))


def test_validation_before_device_calls():
    lib = abi.load_library()
    s = scene.dam_break((4, 4, 4), 0.01, (0.2, 0.2, 0.2), seed=1, steps=1, boundary="")
    bad = abi.Parameters.from_buffer_copy(s.params)
    bad.base_smoothing_radius = 0.0
    h = C.c_void_p()
    rc = lib.asph_create(C.byref(bad), s.positions.ctypes.data, s.count, None, 0, None, s.count, 1, C.byref(h))
    assert rc == 1 and b"smoothing" in lib.asph_last_error()
    rc = lib.asph_create(C.byref(s.params), s.positions.ctypes.data, s.count, None, 0, None, s.count - 1, 1,
                         C.byref(h))
    assert rc == 2 and b"capacity" in lib.asph_last_error()
    rc = lib.asph_create(C.byref(s.params), s.positions.ctypes.data, s.count, None, 0, None, (1 << 28) - 63, 1,
                         C.byref(h))
    assert rc == 2 and b"2^28" in lib.asph_last_error()
    tiny = abi.Parameters.from_buffer_copy(s.params)
    tiny.base_smoothing_radius = 1e-6
    rc = lib.asph_create(C.byref(tiny), s.positions.ctypes.data, s.count, None, 0, None, s.count, 1, C.byref(h))
    assert rc == 1 and b"cells" in lib.asph_last_error()
    assert lib.asph_update(None, 0.1) == 1


def test_grid_cells_of_the_workloads():
    assert scene.cfg1().meta["n"] == (25, 50, 25)
    from adaptivesph_amd.batch import grid_cells

    assert grid_cells(scene.dam_break((2, 2, 2), 0.01, (1.0, 0.6, 0.5), boundary="").params) == 20 * 12 * 10
    assert grid_cells(scene.dam_break((2, 2, 2), 0.0039, (2.0, 1.2, 1.0), boundary="").params) == 99 * 60 * 50


@pytest.mark.parametrize("which,count", [("cfg1", 37482), ("cfg2", 851860)])
def test_boundary_generator_matches_oracle(which, count):
    s = getattr(scene, which)(boundary="")
    lib_b = scene.generate_boundary(s.params, "library")
    ora_b = oracle.boundary_generate(s.params.domain, scene.boundary_spacing(s.params), s.params.rest_density,
                                     s.params.base_smoothing_radius)
    assert len(lib_b) == count
    assert np.array_equal(lib_b.view(np.uint8), ora_b.view(np.uint8))


def test_splitmix_uniform_known_value():
    u = scene.splitmix64_uniform(0, 1)[0]
    assert u == (0xE220A8397B1DCDAF >> 40) * 2.0 ** -24
    a = scene.lattice((3, 4, 5), 0.01, (0.0, 0.0, 0.0), seed=9)
    b = scene.lattice((3, 4, 5), 0.01, (0.0, 0.0, 0.0), seed=9)
    assert np.array_equal(a, b) and a.shape == (60, 4) and (a[:, 3] == 0).all()
    # x outer, z innermost (App.cpp:326-339)
    assert a[1, 2] > a[0, 2] and abs(a[1, 0] - a[0, 0]) < 0.001


def test_scene_bytes_monotone():
    lib = abi.load_library()
    a = lib.asph_scene_bytes(1000, 0, 100, 0, 128)
    b = lib.asph_scene_bytes(2000, 0, 100, 0, 128)
    c = lib.asph_scene_bytes(2000, 0, 100, 1, 128)
    assert 0 < a < b < c


def _noalloc_call(domain, spacing, h, out, capacity):
    lib = abi.load_library()
    n = C.c_uint32(capacity)
    rc = lib.asph_boundary_generate_device_noalloc(C.byref(domain), spacing, 1000.0, h, out, C.byref(n), None)
    return rc, n.value, lib.asph_last_error()


def test_noalloc_walls_count_and_limits_without_gpu_work():
    """The allocation-free wall generator answers a count query (out = NULL)
    with the host generator's count, and refuses, before any GPU work, domains
    past its limits: more coordinate values than its scratch holds, more wall
    particles than ASPH_NOALLOC_MAX_WALLS, and a spacing that cannot advance
    the coordinate sequences."""
    lib = abi.load_library()
    for s in (scene.cfg1(boundary=""), scene.cfg2(boundary="")):
        spacing = scene.boundary_spacing(s.params)
        rc, n, _ = _noalloc_call(s.params.domain, spacing, s.params.base_smoothing_radius, None, 0)
        ref = C.c_uint32(0)
        assert lib.asph_boundary_generate(C.byref(s.params.domain), spacing, 1000.0, s.params.base_smoothing_radius,
                                          None, C.byref(ref)) == 0
        assert rc == 0 and n == ref.value
    s = scene.cfg2(boundary="")
    # 1.44M walls at spacing 3 mm: counted, but refused as output (a host
    # check: no GPU is visible here, and no device buffer is touched)
    rc, n, _ = _noalloc_call(s.params.domain, 0.003, 0.004, None, 0)
    assert rc == 0 and n > abi.NOALLOC_MAX_WALLS
    if not HAVE_GPU:
        rc, _, msg = _noalloc_call(s.params.domain, 0.003, 0.004, 1, 1 << 21)
        assert rc == abi.ASPH_ERR_CAPACITY and b"at most 1048576" in msg
    rc, _, msg = _noalloc_call(s.params.domain, 1e-5, 0.004, None, 0)
    assert rc == abi.ASPH_ERR_CAPACITY and b"65,536" in msg
    big = abi.Domain(abi.Vec3(1.0e6, 0.0, 0.0), abi.Vec3(1.0e6 + 1.0, 1.0, 1.0), 0.01)
    rc, _, msg = _noalloc_call(big, 0.01, 0.004, None, 0)
    assert rc == abi.ASPH_ERR_INVALID and b"advance" in msg


def test_grid_limits_of_the_engine():
    """makeGrid refuses grids beyond 1e9 cells or empty on an axis (the limit
    that creation and updateDomain share), for asph_grid_cells too."""
    lib = abi.load_library()
    n = C.c_uint32(0)
    d = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(2.0, 1.2, 1.0), 0.01)
    assert lib.asph_grid_cells(C.byref(d), 0.00507, C.byref(n)) == 0 and n.value == 99 * 60 * 50
    # 2^31 cells along one axis: the float quotient would not fit an int
    huge = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(1.0e9, 1.0, 1.0), 0.01)
    assert lib.asph_grid_cells(C.byref(huge), 0.1, C.byref(n)) == abi.ASPH_ERR_INVALID
    assert b"1e9" in lib.asph_last_error()
    # 1,200^3 cells: every axis fits, the product does not
    cube = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(480.0, 480.0, 480.0), 0.01)
    assert lib.asph_grid_cells(C.byref(cube), 0.1, C.byref(n)) == abi.ASPH_ERR_INVALID
    flat = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(1.0, 0.0, 1.0), 0.01)
    assert lib.asph_grid_cells(C.byref(flat), 0.1, C.byref(n)) == abi.ASPH_ERR_INVALID


def test_imported_buffers_checked_before_device_work():
    """asph_create_imported refuses imported buffers smaller than the
    capacities need (or NULL) before it touches the device: the pointers
    below are never dereferenced."""
    lib = abi.load_library()
    s = scene.dam_break((4, 4, 4), 0.01, (0.2, 0.2, 0.2), seed=1, steps=1, boundary="")
    cap, bcap = s.count, 8
    fake = 0x1000

    def buf(n):
        return abi.DeviceBuffer(fake, n)

    good_f = abi.FluidBuffers(buf(16 * cap), buf(16 * cap), buf(4 * cap), buf(4 * cap))
    good_b = abi.BoundaryBuffers(buf(16 * bcap), buf(4 * bcap), buf(16 * bcap))
    h = C.c_void_p()
    for f, b, what in ((abi.FluidBuffers(buf(16 * cap - 1), buf(16 * cap), buf(4 * cap), buf(4 * cap)), good_b,
                        b"fluid positions"),
                       (abi.FluidBuffers(buf(16 * cap), buf(16 * cap), buf(4 * cap), abi.DeviceBuffer(None, 4 * cap)),
                        good_b, b"fluid radii"),
                       (good_f, abi.BoundaryBuffers(buf(16 * bcap), buf(4 * bcap), buf(16 * bcap - 16)),
                        b"boundary colours")):
        rc = lib.asph_create_imported(C.byref(s.params), s.positions.ctypes.data, s.count, C.byref(f), C.byref(b),
                                      None, 0, None, cap, bcap, C.byref(h))
        assert rc == abi.ASPH_ERR_CAPACITY and what in lib.asph_last_error(), lib.asph_last_error()
    assert lib.asph_create_imported(C.byref(s.params), s.positions.ctypes.data, s.count, None, C.byref(good_b), None,
                                    0, None, cap, bcap, C.byref(h)) == abi.ASPH_ERR_INVALID
