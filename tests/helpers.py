"""Shared helpers of the parity tests."""
from __future__ import annotations

import numpy as np

from adaptivesph_amd import abi, scene

FIELDS = ("positions", "velocities", "accelerations", "densities", "radii", "smoothing_radii", "masses")


def assert_bitexact(a: dict, b: dict, where: str, fields=FIELDS) -> None:
    for k in fields:
        x = np.ascontiguousarray(a[k])
        y = np.ascontiguousarray(b[k])
        assert x.shape == y.shape, f"{where}: {k} shape {x.shape} vs {y.shape}"
        if not np.array_equal(x.view(np.uint32), y.view(np.uint32)):
            bad = np.flatnonzero((x != y).reshape(x.shape[0], -1).any(axis=1))
            d = np.abs(x.astype(np.float64) - y.astype(np.float64)).max()
            raise AssertionError(f"{where}: {k} differs at {bad.size} particles (first {bad[:8]}), max |diff| {d:.3e}")


def boundary_scene(seed: int = 11, steps: int = 30) -> scene.Scene:
    """A coarse dam break whose support 2h exceeds the 0.05 m wall offset, so
    the boundary density, pressure and friction terms are non-zero."""
    s = scene.dam_break((8, 10, 6), 0.03, (0.5, 0.45, 0.3), seed=seed, steps=steps, name="boundary_active")
    return s


def floor_block_scene(n=(10, 12, 8), dx: float = 0.01, box=(0.3, 0.3, 0.2), seed: int = 21,
                      steps: int = 8) -> scene.Scene:
    """A lattice block standing on the floor in the middle of the box, so some
    particles are near a wall and others far from every wall."""
    s = scene.dam_break(n, dx, box, seed=seed, steps=steps, name="floor_block")
    origin = (0.5 * (box[0] - n[0] * dx), 0.0, 0.5 * (box[2] - n[2] * dx))
    s.positions = scene.lattice(n, dx, origin, seed)
    return s


def random_state(s: scene.Scene, seed: int, vmax: float = 0.1) -> dict:
    """Positions of the scene with random velocities (|v| < vmax, w = 0)."""
    rng = np.random.default_rng(seed)
    n = s.count
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.uniform(-vmax, vmax, size=(n, 3)).astype(np.float32)
    return {"positions": s.positions.copy(), "velocities": v}


def refinement(criterion: str, max_count: int, **kw) -> abi.RefinementParameters:
    r = abi.default_refinement()
    r.enabled = 1
    r.criterion_type = abi.CRITERION[criterion]
    r.max_particle_count = int(max_count)
    r.initial_cooldown = kw.pop("initial_cooldown", 0)
    r.cooldown = kw.pop("cooldown", 1)
    for k, v in kw.items():
        if k == "velocity":
            r.velocity = abi.ThresholdPair(*v)
        elif k == "vorticity":
            r.vorticity = abi.ThresholdPair(*v)
        elif k == "interface":
            r.interface_ratio = abi.ThresholdPair(*v)
        elif k == "zone":
            c, rad = v
            r.zone = abi.ZoneParameters(abi.Vec3(*c), rad, 0.02, 0.05)
        elif k == "batch_ratio":
            r.max_batch_ratio = v
        elif k == "max_mass_ratio":
            r.max_mass_ratio = v
        else:
            raise KeyError(k)
    return r
