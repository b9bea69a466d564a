"""asph_create_imported: the simulation's fluid arrays are the caller's
device buffers (positions, velocities, densities, radii) and its boundary
render buffers (positions, radii, colours) are written at creation and on
every updateDomain, as the reference writes into imported memory
(SphSimulation.cu:29-99, 287-333). Checked against asph_create's own
simulation, bit for bit, through steps and domain changes (host walls and
device walls)."""
import ctypes as C

import numpy as np
import pytest

from adaptivesph_amd import abi, scene
from adaptivesph_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


class Buffers:
    def __init__(self, lib, sizes):
        self.lib = lib
        self.b = {}
        for k, n in sizes.items():
            buf = abi.DeviceBuffer()
            abi.check(lib, lib.asph_buffer_allocate(n, C.byref(buf)), "allocate")
            self.b[k] = buf

    def get(self, k, dtype, shape):
        a = np.zeros(shape, dtype)
        abi.check(self.lib, self.lib.asph_buffer_download(C.byref(self.b[k]), 0, a.ctypes.data, a.nbytes), "download")
        return a

    def close(self):
        for buf in self.b.values():
            self.lib.asph_buffer_release(C.byref(buf))


def test_imported_buffers_follow_updates_and_domains():
    lib = abi.load_library()
    s = scene.dam_break((9, 10, 8), 0.01, (0.3, 0.25, 0.2), seed=8, steps=6, name="imported_abi")
    cap = s.count + 5
    bcap = 4 * len(s.boundary)
    fluid = Buffers(lib, {"positions": 16 * cap, "velocities": 16 * cap, "densities": 4 * cap, "radii": 4 * cap})
    walls = Buffers(lib, {"positions": 16 * bcap, "radii": 4 * bcap, "colors": 16 * bcap})
    fb = abi.FluidBuffers(fluid.b["positions"], fluid.b["velocities"], fluid.b["densities"], fluid.b["radii"])
    wb = abi.BoundaryBuffers(walls.b["positions"], walls.b["radii"], walls.b["colors"])
    h = C.c_void_p()
    abi.check(lib, lib.asph_create_imported(C.byref(s.params), s.positions.ctypes.data, s.count, C.byref(fb),
                                            C.byref(wb), s.boundary.ctypes.data, len(s.boundary), None, cap, bcap,
                                            C.byref(h)), "create_imported")
    plain = Simulation(s.params, s.positions, s.boundary, None, cap, bcap)
    try:
        def compare(where):
            n = s.count
            ref = plain.fluid()
            for k, dt, shape in (("positions", np.float32, (n, 4)), ("velocities", np.float32, (n, 4)),
                                 ("densities", np.float32, (n,)), ("radii", np.float32, (n,))):
                got = fluid.get(k, dt, shape)
                assert np.array_equal(got.view(np.uint32), ref[k].view(np.uint32)), (where, k)

        def compare_walls(records, where):
            v = abi.BoundaryView()
            abi.check(lib, lib.asph_get_boundary_view(h, C.byref(v)), "boundary view")
            nb = len(records)
            assert v.particle_count == nb
            assert v.positions == walls.b["positions"].data
            pos = walls.get("positions", np.float32, (nb, 4))
            assert np.array_equal(pos, np.ascontiguousarray(records["position"])), where
            assert (walls.get("radii", np.float32, (nb,)) == np.float32(s.params.base_particle_radius)).all()
            assert (walls.get("colors", np.float32, (nb, 4)) == np.array([0.3, 0.3, 0.3, 0.2], np.float32)).all()
            psi = np.zeros(nb, np.float32)
            abi.check(lib, lib.asph_buffer_download(C.byref(abi.DeviceBuffer(v.psi_values, 4 * nb)), 0,
                                                    psi.ctypes.data, psi.nbytes), "psi")
            assert np.array_equal(psi, records["psi"]), where

        compare("created")
        compare_walls(s.boundary, "created")
        for _ in range(s.steps):
            abi.check(lib, lib.asph_update(h, s.dt), "update")
            plain.update(s.dt)
        compare("stepped")
        # a moved domain with host walls, then a grown one with device walls
        d1 = abi.Domain(abi.Vec3(-0.01, 0.0, 0.0), abi.Vec3(0.29, 0.25, 0.2), 0.02)
        p1 = abi.Parameters.from_buffer_copy(s.params)
        p1.domain = d1
        b1 = scene.generate_boundary(p1)
        abi.check(lib, lib.asph_update_domain(h, C.byref(d1), b1.ctypes.data, len(b1)), "update_domain")
        plain.update_domain(d1, b1)
        compare_walls(b1, "host walls")
        d2 = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(0.32, 0.27, 0.21), 0.03)
        p2 = abi.Parameters.from_buffer_copy(s.params)
        p2.domain = d2
        spacing = scene.boundary_spacing(s.params)
        abi.check(lib, lib.asph_update_domain_walls(h, C.byref(d2), spacing), "update_domain_walls")
        plain.update_domain_walls(d2, spacing)
        compare_walls(scene.generate_boundary(p2), "device walls")
        for _ in range(s.steps):
            abi.check(lib, lib.asph_update(h, s.dt), "update")
            plain.update(s.dt)
        compare("after the domain changes")
        # a domain whose walls exceed the boundary buffers is refused
        big = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(0.8, 0.6, 0.6), 0.01)
        assert lib.asph_update_domain_walls(h, C.byref(big), spacing) == abi.ASPH_ERR_CAPACITY
        assert b"capacity" in lib.asph_last_error()
    finally:
        lib.asph_destroy(h)
        plain.close()
        fluid.close()
        walls.close()
