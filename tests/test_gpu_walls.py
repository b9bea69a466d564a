"""Wall psi on the GPU (asph_boundary_generate_device) against the host
generator (asph_boundary_generate, itself pinned against the oracle's
restatement of StaticBoundaryDomain::generate): identical records, bit for
bit, for the dam-break walls of both sizes and for a fine lattice whose
particles have more psi neighbours than the kernel's LDS table holds."""
import ctypes as C

import numpy as np
import pytest

from adaptivesph_amd import abi, scene

pytestmark = pytest.mark.gpu


def _generate(fn, domain, spacing, rho0, h):
    lib = abi.load_library()
    n = C.c_uint32(0)
    abi.check(lib, getattr(lib, fn)(C.byref(domain), spacing, rho0, h, None, C.byref(n)), "count")
    out = np.zeros(n.value, dtype=abi.BOUNDARY_DTYPE)
    abi.check(lib, getattr(lib, fn)(C.byref(domain), spacing, rho0, h, out.ctypes.data, C.byref(n)), fn)
    return out


@pytest.mark.parametrize("make", [scene.cfg1, scene.cfg2])
def test_device_psi_equals_host(make):
    s = make()
    p = s.params
    args = (p.domain, scene.boundary_spacing(p), p.rest_density, p.base_smoothing_radius)
    host = _generate("asph_boundary_generate", *args)
    dev = _generate("asph_boundary_generate_device", *args)
    assert len(host) == len(dev) == len(s.boundary)
    assert host.tobytes() == dev.tobytes()
    assert np.isfinite(dev["psi"]).all() and (dev["psi"] > 0).all()


def test_device_psi_many_neighbours():
    """h = 3 spacings: about 110 hits per wall particle, past the 64 held in
    LDS, so the kernel's selection fallback runs."""
    p = abi.default_parameters()
    p.domain.min.x, p.domain.min.y, p.domain.min.z = 0.0, 0.0, 0.0
    p.domain.max.x, p.domain.max.y, p.domain.max.z = 0.1, 0.08, 0.06
    spacing = 0.004
    h = 3.0 * spacing
    host = _generate("asph_boundary_generate", p.domain, spacing, 1000.0, h)
    dev = _generate("asph_boundary_generate_device", p.domain, spacing, 1000.0, h)
    assert host.tobytes() == dev.tobytes()


def _with_domain(p, d):
    q = abi.Parameters.from_buffer_copy(p)
    q.domain = d
    return q


@pytest.mark.parametrize("fix_cell_start", [False, True])
def test_update_domain_walls_equals_host_walls(fix_cell_start):
    """A walls-active scene (2h > the 0.05 m wall offset) whose box moves and
    shrinks twice: asph_update_domain_walls (walls, psi and binning on the
    GPU) leaves the simulation in the state asph_update_domain with the host
    generator's records does, step after step."""
    from adaptivesph_amd.simulation import Simulation
    from helpers import assert_bitexact, boundary_scene

    s = boundary_scene(seed=12)
    if fix_cell_start:
        s.params.flags |= abi.ASPH_FLAG_FIX_CELL_START
    a = Simulation.from_scene(s)
    b = Simulation.from_scene(s)
    spacing = scene.boundary_spacing(s.params)
    for k, (lo, hi) in enumerate((((0.01, 0.0, -0.01), (0.47, 0.44, 0.3)), ((0.02, 0.0, 0.0), (0.4, 0.42, 0.28)))):
        for _ in range(4):
            a.update(s.dt)
            b.update(s.dt)
        d = abi.Domain(abi.Vec3(*lo), abi.Vec3(*hi), s.params.domain.friction)
        walls = scene.generate_boundary(_with_domain(s.params, d))
        a.update_domain(d, walls)
        b.update_domain_walls(d, spacing)
        assert a.get_boundary_particles_count() == b.get_boundary_particles_count() == len(walls)
        for step in range(6):
            a.update(s.dt)
            b.update(s.dt)
            assert_bitexact(b.fluid(), a.fluid(), f"domain {k} step {step}")
        ga, gb = a.grid(), b.grid()
        for key in ga:
            assert np.array_equal(ga[key], gb[key]), key
    a.close()
    b.close()


def test_batch_scene_moves_its_tank():
    """In a graph-replayed batch, one scene's box moves twice (walls rebuilt
    on the GPU between steps); that scene evolves exactly as alone with the
    same domain updates, and the other scene is untouched."""
    from adaptivesph_amd.batch import Batch
    from adaptivesph_amd.simulation import Simulation
    from helpers import assert_bitexact, boundary_scene

    a = boundary_scene(seed=13)
    b = scene.dam_break((8, 8, 8), 0.01, (0.2, 0.2, 0.2), seed=14, steps=0, name="other")
    batch = Batch.for_scenes([a, b], graph=True, slack=1.6)
    ia, ib = batch.add(a), batch.add(b)
    alone = Simulation.from_scene(a)
    other = Simulation.from_scene(b)
    spacing = scene.boundary_spacing(a.params)
    for lo, hi in (((0.01, 0.0, -0.01), (0.47, 0.44, 0.3)), ((0.02, 0.0, 0.0), (0.4, 0.42, 0.28))):
        batch.step(3)
        for _ in range(3):
            alone.update(a.dt)
            other.update(b.dt)
        d = abi.Domain(abi.Vec3(*lo), abi.Vec3(*hi), a.params.domain.friction)
        batch.update_domain_walls(ia, d, spacing)
        alone.update_domain_walls(d, spacing)
    batch.step(5)
    for _ in range(5):
        alone.update(a.dt)
        other.update(b.dt)
    assert_bitexact(batch.fluid(ia), alone.fluid(), "moved scene")
    assert_bitexact(batch.fluid(ib), other.fluid(), "other scene")
    batch.close()
    alone.close()
    other.close()


def test_oversized_domain_refused_and_scene_unchanged():
    """updateDomain with a domain whose grid would exceed 1e9 cells (or is
    empty on an axis) is refused with ASPH_ERR_INVALID before anything
    changes: the simulation keeps its grid and walls and steps on exactly as
    a simulation that never saw the call."""
    from adaptivesph_amd.simulation import Simulation
    from helpers import assert_bitexact

    s = scene.dam_break((8, 10, 6), 0.01, (0.3, 0.2, 0.2), seed=4, steps=4)
    a = Simulation.from_scene(s)
    b = Simulation.from_scene(s)
    try:
        a.update(s.dt)
        b.update(s.dt)
        huge = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(480.0, 480.0, 480.0), 0.01)
        for call in (lambda: a.update_domain(huge, s.boundary),
                     lambda: a.update_domain_walls(huge, scene.boundary_spacing(s.params)),
                     lambda: a.update_domain(abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(0.3, 0.0, 0.2), 0.01),
                                             s.boundary)):
            with pytest.raises(abi.AsphError) as e:
                call()
            assert e.value.status == abi.ASPH_ERR_INVALID
        assert a.get_boundary_particles_count() == b.get_boundary_particles_count() == len(s.boundary)
        assert a.grid_dims() == b.grid_dims()
        for _ in range(s.steps):
            a.update(s.dt)
            b.update(s.dt)
        assert_bitexact(a.fluid(), b.fluid(), "after the refused updateDomain")
    finally:
        a.close()
        b.close()
