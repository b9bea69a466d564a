"""GPU parity of the plain WCSPH step against the FP32 oracle (bit-exact).

Cell keys, sorted indices, cell start/end (with the reference's slot-0
omission), neighbour sets, densities, accelerations, velocities and positions
must equal the oracle's bits, on the first step and over whole trajectories.
"""
import dataclasses

import numpy as np
import pytest

from adaptivesph_amd import abi, scene
from adaptivesph_amd.simulation import Simulation
from helpers import assert_bitexact, boundary_scene
from oracle import oracle

pytestmark = pytest.mark.gpu


def _pair(s):
    return Simulation.from_scene(s), oracle.OracleSim.from_scene(s)


def test_cfg1_first_step_grid_and_neighbours():
    s = scene.cfg1()
    sim, ref = _pair(s)
    sim.update(s.dt)
    ref.update(s.dt)
    g, go = sim.grid(), ref.grid()
    for k in ("keys", "indices", "cell_start", "cell_end"):
        assert np.array_equal(g[k], go[k]), k
    c, fp = sim.neighbor_sets()
    co, fpo = ref.neighbor_sets()
    assert np.array_equal(c, co)
    assert np.array_equal(fp, fpo)
    # interior particles of the lattice have 80 neighbours (81 with self)
    assert c.max() == 80
    assert_bitexact(sim.fluid(), ref.fluid(), "cfg1 step 1")


def test_cfg1_200_steps_bitexact():
    s = scene.cfg1()
    sim, ref = _pair(s)
    for step in range(1, s.steps + 1):
        sim.update(s.dt)
        ref.update(s.dt)
        if step in (1, 10, 50, 100, 200):
            assert_bitexact(sim.fluid(), ref.fluid(), f"cfg1 step {step}")
    assert sim.calculate_average_neighbor_count() == ref.average_neighbor_count()
    a = sim.get_density_info(0.1)
    b = ref.density_info(0.1)
    for f, _ in abi.DensityInfo._fields_:
        assert getattr(a, f) == getattr(b, f), f


def test_exact_lattice_bitexact():
    """An unjittered lattice (as the reference application's particle grid
    is): pairs on one lattice line have offset components of exactly 0, which
    the force loop takes through the shared reciprocal."""
    n = (16, 20, 12)
    s = scene.dam_break(n, 0.01, (0.5, 0.4, 0.3), seed=1, steps=60, name="exact_lattice")
    s = dataclasses.replace(s, positions=scene.lattice(n, 0.01, (0.0, 0.0, 0.0), 1, jitter=0.0))
    assert [len(np.unique(s.positions[:, k])) for k in range(3)] == list(n)
    sim, ref = _pair(s)
    for step in range(1, s.steps + 1):
        sim.update(s.dt)
        ref.update(s.dt)
        if step in (1, 10, 60):
            assert_bitexact(sim.fluid(), ref.fluid(), f"exact lattice step {step}")
    # the lattice lines persist: particles still share exact coordinates
    assert len(np.unique(sim.fluid()["positions"][:, 0])) < s.count


def test_boundary_terms_active():
    s = boundary_scene()
    sim, ref = _pair(s)
    for step in range(1, s.steps + 1):
        sim.update(s.dt)
        ref.update(s.dt)
    assert_bitexact(sim.fluid(), ref.fluid(), "boundary scene")
    # the wall terms must actually contribute here
    plain = Simulation(s.params, s.positions, np.zeros(0, abi.BOUNDARY_DTYPE))
    for _ in range(s.steps):
        plain.update(s.dt)
    assert not np.array_equal(plain.fluid()["densities"], sim.fluid()["densities"])


def test_viscosity_surface_tension_and_sloshing():
    s = scene.dam_break((10, 12, 8), 0.01, (0.3, 0.3, 0.2), seed=5, steps=40)
    s.params.viscosity_constant = 0.05
    s.params.surface_tension_constant = 0.02
    s.params.slosh_amplitude = 0.5
    s.params.slosh_period = 0.01
    s.params.tait_gamma = 5.0
    sim, ref = _pair(s)
    for _ in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
    assert_bitexact(sim.fluid(), ref.fluid(), "visc+st+slosh")


def test_velocity_clamp_and_wall_clamp():
    s = scene.dam_break((6, 6, 6), 0.01, (0.2, 0.2, 0.2), seed=9, steps=5)
    s.params.max_velocity = 0.05
    sim, ref = _pair(s)
    st = {"positions": s.positions.copy(), "velocities": np.zeros((s.count, 4), np.float32)}
    st["velocities"][:, 0] = -3.0  # drives particles through the x = 0 wall
    sim.set_fluid(dict(st))
    ref.set_fluid(dict(st))
    for _ in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
    a = sim.fluid()
    assert_bitexact(a, ref.fluid(), "clamps")
    assert (a["positions"][:, 0] >= 0.0).all()


@pytest.mark.parametrize("n", [1, 2, 3])
def test_slot0_omission_tiny(n):
    """N = 1 and N = 2 in different cells lose every neighbour (reference quirk)."""
    s = scene.dam_break((1, 1, 1), 0.01, (0.2, 0.2, 0.2), seed=1, steps=1, boundary="")
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = [0.01, 0.11, 0.15][:n]
    pos[:, 1] = 0.01
    pos[:, 2] = 0.01
    sim = Simulation(s.params, pos, s.boundary)
    ref = oracle.OracleSim(s.params, pos, s.boundary)
    sim.update(s.dt)
    ref.update(s.dt)
    g, go = sim.grid(), ref.grid()
    for k in ("keys", "indices", "cell_start", "cell_end"):
        assert np.array_equal(g[k], go[k]), k
    a, b = sim.fluid(), ref.fluid()
    assert_bitexact(a, b, f"tiny n={n}")


def test_slot0_omission_and_fix_flag():
    """Lowest-key cell with exactly one particle: that cell and the next are
    invisible; ASPH_FLAG_FIX_CELL_START restores them."""
    s = scene.dam_break((6, 6, 6), 0.01, (0.3, 0.3, 0.3), seed=3, steps=3, boundary="")
    pos = s.positions.copy()
    pos[0, :3] = [0.001, 0.001, 0.001]
    pos[1:, 0] += 0.055  # every other particle leaves the first cell column
    for flags in (0, abi.ASPH_FLAG_FIX_CELL_START):
        s.params.flags = flags
        sim = Simulation(s.params, pos, s.boundary)
        ref = oracle.OracleSim(s.params, pos, s.boundary)
        for _ in range(s.steps):
            sim.update(s.dt)
            ref.update(s.dt)
        g, go = sim.grid(), ref.grid()
        for k in ("keys", "indices", "cell_start", "cell_end"):
            assert np.array_equal(g[k], go[k]), (flags, k)
        assert_bitexact(sim.fluid(), ref.fluid(), f"quirk flags={flags}")
    s.params.flags = 0


def test_update_domain_and_diagnostics():
    s = scene.dam_break((8, 8, 8), 0.02, (0.5, 0.5, 0.5), seed=4, steps=5)
    sim, ref = _pair(s)
    for _ in range(3):
        sim.update(s.dt)
        ref.update(s.dt)
    assert sim.get_boundary_particles_count() == len(s.boundary)
    d = abi.Domain(abi.Vec3(0.0, 0.0, 0.0), abi.Vec3(0.6, 0.5, 0.4), 0.01)
    b = scene.generate_boundary(_with_domain(s.params, d))
    sim.update_domain(d, b)
    assert sim.get_boundary_particles_count() == len(b)
    sim.update(s.dt)
    ref2 = oracle.OracleSim(_with_domain(s.params, d), ref.fluid()["positions"], b)
    ref2.set_fluid(ref.fluid())
    ref2.update(s.dt)
    assert_bitexact(sim.fluid(), ref2.fluid(), "after updateDomain")


def _with_domain(p, d):
    q = abi.Parameters.from_buffer_copy(p)
    q.domain = d
    return q


def test_invalid_arguments_raise():
    s = scene.dam_break((4, 4, 4), 0.01, (0.2, 0.2, 0.2), seed=1, steps=1, boundary="")
    bad = abi.Parameters.from_buffer_copy(s.params)
    bad.base_smoothing_radius = 0.0
    with pytest.raises(RuntimeError, match="smoothing"):
        Simulation(bad, s.positions, s.boundary)
    with pytest.raises(RuntimeError, match="capacity"):
        Simulation(s.params, s.positions, s.boundary, None, s.count - 1)
    pos = s.positions.copy()
    pos[0, 3] = 1.0
    with pytest.raises(RuntimeError, match="w == 0"):
        Simulation(s.params, pos, s.boundary)


def test_cfg2_first_step_bitexact_and_fp64_tolerance():
    """The 1,048,576-particle dam break: keys, sorted order, cell ranges,
    neighbour sets and state after one step equal the FP32 oracle's bits;
    density and acceleration agree with the FP64 oracle to 1e-4."""
    s = scene.cfg2()
    sim, ref = _pair(s)
    sim.update(s.dt)
    ref.update(s.dt)
    g, go = sim.grid(), ref.grid()
    for k in ("keys", "indices", "cell_start", "cell_end"):
        assert np.array_equal(g[k], go[k]), k
    c, fp = sim.neighbor_sets()
    co, fpo = ref.neighbor_sets()
    assert np.array_equal(c, co) and np.array_equal(fp, fpo)
    assert c.max() == 80
    got = sim.fluid()
    assert_bitexact(got, ref.fluid(), "cfg2 step 1")
    ref.close()
    r64 = oracle.OracleSim.from_scene(s, precision=64)
    r64.update(s.dt)
    f64 = r64.fluid64()
    rho_rel = np.abs(got["densities"].astype(np.float64) - f64["densities"]) / f64["densities"]
    assert rho_rel.max() <= 1e-4
    a64 = f64["accelerations"][:, :3]
    da = np.linalg.norm(got["accelerations"][:, :3].astype(np.float64) - a64, axis=1)
    assert da.max() <= 1e-4 * np.linalg.norm(a64, axis=1).mean()
    # size-independent invariants of the step: unit masses unchanged, every
    # particle inside the box, densities positive
    assert (got["masses"] == np.float32(s.params.base_particle_mass)).all()
    lo = np.array([s.params.domain.min.x, s.params.domain.min.y, s.params.domain.min.z], np.float32)
    hi = np.array([s.params.domain.max.x, s.params.domain.max.y, s.params.domain.max.z], np.float32)
    assert ((got["positions"][:, :3] >= lo) & (got["positions"][:, :3] <= hi)).all()
    assert (got["densities"] > 0).all()


def test_cfg1_fp64_drift_within_stated_bounds():
    """Trajectory drift against the FP64 oracle (the GPU equals the FP32
    oracle bit for bit, so this is FP32's own rounding noise): first step
    within 1e-4, and the drift bounds quoted with the results."""
    s = scene.cfg1()
    sim = Simulation.from_scene(s)
    r64 = oracle.OracleSim.from_scene(s, precision=64)
    dx = s.meta["dx"]
    bounds = {1: (1e-5, 1e-5), 10: (1e-4, 1e-4), 50: (2e-3, 1e-3), 100: (2e-2, 1e-2)}
    for step in range(1, 101):
        sim.update(s.dt)
        r64.update(s.dt)
        if step in bounds:
            a, b = sim.fluid(), r64.fluid64()
            d = np.linalg.norm(a["positions"][:, :3].astype(np.float64) - b["positions"][:, :3], axis=1) / dx
            rho = np.abs(a["densities"].astype(np.float64) - b["densities"]) / b["densities"]
            assert d.max() <= bounds[step][0], (step, d.max())
            assert rho.max() <= bounds[step][1], (step, rho.max())


def test_plain_scene_with_varied_h_and_m():
    """A plain scene whose smoothing lengths and masses are set per particle
    (not all h0 / m0) takes the general density loop; set back to the base
    values it takes the uniform one. Both match the oracle bit for bit."""
    s = scene.dam_break((10, 12, 8), 0.01, (0.3, 0.3, 0.3), seed=4, steps=4)
    sim, ref = _pair(s)
    rng = np.random.default_rng(7)
    st = {"positions": s.positions.copy(),
          "smoothing_radii": (s.params.base_smoothing_radius * rng.uniform(0.8, 1.2, s.count)).astype(np.float32),
          "masses": (s.params.base_particle_mass * rng.uniform(0.5, 1.5, s.count)).astype(np.float32)}
    sim.set_fluid(dict(st))
    ref.set_fluid(dict(st))
    for _ in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
    assert_bitexact(sim.fluid(), ref.fluid(), "varied h, m")
    f = ref.fluid()
    back = {"positions": f["positions"], "velocities": f["velocities"],
            "smoothing_radii": np.full(s.count, s.params.base_smoothing_radius, np.float32),
            "masses": np.full(s.count, s.params.base_particle_mass, np.float32)}
    sim.set_fluid(dict(back))
    ref.set_fluid(dict(back))
    for _ in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
    assert_bitexact(sim.fluid(), ref.fluid(), "base h, m again")


def test_box_filled_with_active_walls():
    """Fluid filling the whole box with 2h above the wall offset: every wall,
    edge and corner of the domain contributes boundary terms, so the culling
    of boundary walks (per-cell wall distance bounds, boxes of each cell's wall
    slots) is exercised everywhere; bit-exact against the oracle."""
    s = scene.dam_break((8, 8, 6), 0.03, (0.24, 0.24, 0.18), seed=13, steps=10, name="box_filled")
    sim, ref = _pair(s)
    for _ in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
    a = sim.fluid()
    assert_bitexact(a, ref.fluid(), "box filled")
    rho0 = s.params.rest_density
    assert np.isfinite(a["densities"]).all() and a["densities"].max() > 0.5 * rho0


def test_shared_reciprocal_division_equals_division():
    """The density and force loops divide by a shared reciprocal and take
    square roots without range scaling (the steps of the correctly rounded
    operations without their scalings); on 2^30 pseudo-random operand pairs
    with magnitudes in [2^-40, 2^40], plus +0 and the d / h range, every
    quotient has the bits of a / b, and on 2^30 values in [2^-96, 2^127) and
    +0 every root has the bits of sqrtf."""
    import ctypes as C

    lib = abi.load_library()
    bad = C.c_uint64(123)
    assert lib.asph_debug_division_check(7, 1 << 30, C.byref(bad)) == 0, abi.last_error()
    assert bad.value == 0


def test_morris_laminar_viscosity_bitexact():
    """The Morris laminar term (extension, laminar_viscosity > 0) beside the
    artificial viscosity: 40 steps of a dam break equal the FP32 oracle bit
    for bit, and the term changes the trajectory."""
    s = scene.dam_break((10, 14, 8), 0.01, (0.3, 0.3, 0.2), seed=31, steps=40)
    s.params.laminar_viscosity = 5e-3
    sim, ref = _pair(s)
    for k in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
        if k in (0, 9, 39):
            assert_bitexact(sim.fluid(), ref.fluid(), f"Morris step {k + 1}")
    s.params.laminar_viscosity = 0.0
    plain = Simulation.from_scene(s)
    for _ in range(s.steps):
        plain.update(s.dt)
    assert not np.array_equal(plain.fluid()["velocities"], sim.fluid()["velocities"])


def _union_cells(keys: np.ndarray, dims: tuple[int, int, int]) -> np.ndarray:
    """Cells in the union of the +-1 cell boxes of each run of 64 consecutive
    sorted slots (the list builder's waves while the scene keeps the sorted
    query order), clamped to the grid."""
    gx, gy, gz = dims
    k = keys.astype(np.int64)
    x, y, z = k % gx, (k // gx) % gy, k // (gx * gy)
    out = []
    for w in range(0, len(k), 64):
        sl = slice(w, w + 64)
        n = 1
        for c, g in ((x[sl], gx), (y[sl], gy), (z[sl], gz)):
            n *= min(int(c.max()) + 1, g - 1) - max(int(c.min()) - 1, 0) + 1
        out.append(n)
    return np.array(out)


def test_sparse_scene_waves_over_many_cells():
    """A sparse cloud (about one particle per two cells): 64 consecutive
    sorted particles span several rows of cells, so the list builder's waves
    walk union boxes of more than 64 cells, whose slot ranges arrive in
    groups of 64 cells, one per lane (and the widest waves walk lane by lane).
    Bit-exact against the oracle over several steps."""
    s = scene.dam_break((16, 16, 16), 0.01, (0.48, 0.48, 0.48), seed=31, steps=5, name="sparse_cloud")
    rng = np.random.default_rng(31)
    st = {"positions": np.zeros((s.count, 4), np.float32),
          "velocities": np.zeros((s.count, 4), np.float32)}
    st["positions"][:, :3] = rng.uniform(0.04, 0.44, (s.count, 3))
    st["velocities"][:, :3] = rng.uniform(-0.05, 0.05, (s.count, 3))
    sim, ref = _pair(s)
    sim.set_fluid(dict(st))
    ref.set_fluid(dict(st))
    for i in range(s.steps):
        sim.update(s.dt)
        ref.update(s.dt)
        if i == 0:
            cells = _union_cells(ref.grid()["keys"], ref.grid_dims())
            wide = int(np.count_nonzero((cells > 64) & (cells <= 384)))
            print(f"waves {len(cells)}, 65-384 union cells: {wide}, over 384: {int(np.count_nonzero(cells > 384))}")
            assert wide >= 8, cells
    assert_bitexact(sim.fluid(), ref.fluid(), "sparse cloud")
