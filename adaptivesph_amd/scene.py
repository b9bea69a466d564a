"""Headless scene generators (the reference's App builds its scene in the GUI,
app/App.cpp:312-341; these are the workloads the benchmarks and tests run).

Positions are a lattice in the min corner of the box, x outer, then y, then z
innermost (the loop order of App.cpp:326-339), at min + (i + 1/2) dx per axis,
each coordinate jittered by (2u - 1) * jitter * dx with u drawn from
SplitMix64(seed) as u = (z >> 40) * 2^-24, three draws per particle in x, y, z
order. Coordinates are formed in FP64 and rounded to FP32.

Physics defaults: rho0 = 1000, m = rho0 dx^3, particle radius dx / 2 (so the
wall spacing 2r equals dx, App.cpp:143), h = 1.3 dx, Tait gamma = 7,
c0 = 10 sqrt(2 g 0.5), g = (0, -9.81, 0), sigma = 0, wall friction 0.01,
max speed 100 m/s (inactive), and the Morris kinematic viscosity nu = 1e-6
mapped onto the reference's artificial-viscosity alpha by the 3-D relation
alpha = 2 (d + 2) nu / (h c0) = 10 nu / (h c0). The time step is the CFL
choice dt = 0.25 h / c0, fixed per run.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """count uniforms in [0, 1) from SplitMix64(seed): (z >> 40) * 2^-24."""
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


@dataclass
class Scene:
    """One simulation's inputs: parameters, initial positions, static boundary."""

    name: str
    params: abi.Parameters
    positions: np.ndarray  # (N, 4) float32, w = 0
    boundary: np.ndarray  # BOUNDARY_DTYPE records
    refinement: abi.RefinementParameters | None
    capacity: int
    boundary_capacity: int
    dt: float
    steps: int
    meta: dict = field(default_factory=dict)

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])

    @property
    def adaptive(self) -> bool:
        return self.refinement is not None and bool(self.refinement.enabled)


def lattice(n: tuple[int, int, int], dx: float, origin: tuple[float, float, float], seed: int,
            jitter: float = 0.01) -> np.ndarray:
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    u = splitmix64_uniform(seed, 3 * base.shape[0]).reshape(-1, 3)
    pos = np.asarray(origin, dtype=np.float64) + (base + 0.5) * dx + (2.0 * u - 1.0) * jitter * dx
    out = np.zeros((base.shape[0], 4), dtype=np.float32)
    out[:, :3] = pos.astype(np.float32)
    return out


def application_grid(count: tuple[int, int, int], domain: abi.Domain) -> np.ndarray:
    """The reference application's default particle grid (App.cpp:312-339):
    count particles per axis filling the domain minus a 0.15 m margin, in FP32
    as the application computes it, x outer and z innermost."""
    f = np.float32
    margin = f(0.15)
    lo = np.array([domain.min.x, domain.min.y, domain.min.z], f)
    hi = np.array([domain.max.x, domain.max.y, domain.max.z], f)
    spacing = ((hi - lo) - f(2.0) * margin) / (np.asarray(count, f) - f(1.0))
    base = lo + margin
    i, j, k = np.meshgrid(np.arange(count[0]), np.arange(count[1]), np.arange(count[2]), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(f)
    out = np.zeros((idx.shape[0], 4), f)
    out[:, :3] = base + spacing * idx
    return out


def physics(box_min, box_max, dx: float, gamma: float = 7.0, nu: float = 1e-6, h_ratio: float = 1.3,
            g: float = 9.81, c_height: float = 0.5) -> tuple[abi.Parameters, float]:
    """Parameters of a scene with lattice spacing dx, and its time step."""
    p = abi.default_parameters()
    p.domain.min = abi.Vec3(*map(float, box_min))
    p.domain.max = abi.Vec3(*map(float, box_max))
    p.domain.friction = 0.01
    p.gravity = abi.Vec3(0.0, -g, 0.0)
    c0 = 10.0 * math.sqrt(2.0 * g * c_height)
    h = h_ratio * dx
    p.speed_of_sound = c0
    p.rest_density = 1000.0
    p.restitution = 0.0
    p.viscosity_constant = 10.0 * nu / (h * c0)
    p.surface_tension_constant = 0.0
    p.max_velocity = 100.0
    p.base_smoothing_radius = h
    p.base_particle_radius = 0.5 * dx
    p.base_particle_mass = 1000.0 * dx ** 3
    p.threads_per_block = 256
    p.tait_gamma = gamma
    dt = float(np.float32(0.25 * float(np.float32(h)) / float(np.float32(c0))))
    return p, dt


def boundary_spacing(params: abi.Parameters) -> float:
    """Wall particle spacing 2 r (App.cpp:142) in FP32."""
    return float(np.float32(params.base_particle_radius) * np.float32(2.0))


def generate_boundary(params: abi.Parameters, generator: str = "library") -> np.ndarray:
    """StaticBoundaryDomain::generate(domain, 2 r, rho0, h) (App.cpp:143-148)."""
    if generator != "library":
        raise ValueError(f"unknown boundary generator {generator!r}")
    spacing = boundary_spacing(params)
    lib = abi.load_library()
    n = C.c_uint32(0)
    abi.check(lib, lib.asph_boundary_generate(C.byref(params.domain), spacing, params.rest_density,
                                              params.base_smoothing_radius, None, C.byref(n)), "boundary count")
    out = np.zeros(n.value, dtype=abi.BOUNDARY_DTYPE)
    abi.check(lib, lib.asph_boundary_generate(C.byref(params.domain), spacing, params.rest_density,
                                              params.base_smoothing_radius, out.ctypes.data, C.byref(n)),
              "boundary generate")
    return out


def dam_break(n=(25, 50, 25), dx: float = 0.01, box=(1.0, 0.6, 0.5), seed: int = 1, steps: int = 200,
              gamma: float = 7.0, boundary: str = "library", name: str = "dam_break") -> Scene:
    """Plain WCSPH dam break: an n lattice in the min corner of a closed box."""
    pos = lattice(n, dx, (0.0, 0.0, 0.0), seed)
    p, dt = physics((0.0, 0.0, 0.0), box, dx, gamma=gamma)
    b = generate_boundary(p, boundary) if boundary else np.zeros(0, dtype=abi.BOUNDARY_DTYPE)
    return Scene(name, p, pos, b, None, pos.shape[0], max(len(b), 1), dt, steps,
                 {"n": tuple(n), "dx": dx, "box": tuple(box), "seed": seed})


def cfg1(boundary: str = "library") -> Scene:
    """31,250-particle dam break, dx = 0.01 m in a 1.0 x 0.6 x 0.5 m box, 200 steps."""
    return dam_break((25, 50, 25), 0.01, (1.0, 0.6, 0.5), seed=1, steps=200, boundary=boundary, name="cfg1")


def cfg2(boundary: str = "library") -> Scene:
    """1,048,576-particle dam break (128 x 64 x 128, dx = 0.0039 m) in a
    2.0 x 1.2 x 1.0 m box (the 1.0 x 0.6 x 0.5 box scaled uniformly by 2), 1000 steps."""
    return dam_break((128, 64, 128), 0.0039, (2.0, 1.2, 1.0), seed=2, steps=1000, boundary=boundary, name="cfg2")


def slosh_refinement(max_count: int, center, radius: float = 0.1, path_radius: float = 0.2,
                     path_period: float = 2.0) -> abi.RefinementParameters:
    r = abi.default_refinement()
    r.enabled = 1
    r.max_particle_count = int(max_count)
    r.initial_cooldown = 0
    r.cooldown = 1
    r.criterion_type = abi.CRITERION["zone"]
    r.zone = abi.ZoneParameters(abi.Vec3(*map(float, center)), radius, path_radius, path_period)
    return r


def sloshing_tank(n=(125, 50, 80), dx: float = 0.008, box=(1.0, 0.6, 0.64), seed: int = 3, steps: int = 2000,
                  adaptive: bool = True, max_count: int = 4_000_000, boundary: str = "library",
                  name: str = "slosh") -> Scene:
    """Tank with the bottom filled; horizontal wall acceleration 0.5 m/s^2,
    period 1 s, as a body force; moving spherical refinement zone."""
    pos = lattice(n, dx, (0.0, 0.0, 0.0), seed)
    p, dt = physics((0.0, 0.0, 0.0), box, dx)
    p.slosh_amplitude = 0.5
    p.slosh_period = 1.0
    fill = n[1] * dx
    center = (0.5 * box[0], 0.5 * fill, 0.5 * box[2])
    ref = slosh_refinement(max_count, center) if adaptive else None
    cap = max_count if adaptive else pos.shape[0]
    b = generate_boundary(p, boundary) if boundary else np.zeros(0, dtype=abi.BOUNDARY_DTYPE)
    return Scene(name, p, pos, b, ref, cap, max(len(b), 1), dt, steps,
                 {"n": tuple(n), "dx": dx, "box": tuple(box), "seed": seed})


def kernel_merged_h(s: Scene) -> Scene:
    """Switches the scene to the kernel-normalised merged smoothing length
    (abi.FLAG_KERNEL_MERGED_H). With the reference's coefficient every merge
    multiplies h by about 1.8-2, and split/merge cycles under frequent
    refinement grow h without bound (over 100 h0 within 100 steps of a
    sloshing tank refining every step), until neighbour searches span the
    whole tank."""
    s.params.flags |= abi.FLAG_KERNEL_MERGED_H
    s.meta["merged_h"] = "kernel"
    return s


def cfg3(boundary: str = "library", merged_h: str = "kernel") -> Scene:
    """Adaptive sloshing tank: 500,000 particles at dx = 0.008 m (125 x 50 x 80 in a
    1.0 x 0.6 x 0.64 m tank), refinement zone r = 0.1 m on a circle of radius
    0.2 m (period 2 s), refinement every step, cap 4M, 2000 steps."""
    s = sloshing_tank(boundary=boundary, name="cfg3")
    return kernel_merged_h(s) if merged_h == "kernel" else s


def random_scene(seed: int, n_min: int = 5_000, n_max: int = 200_000, box=(1.0, 0.6, 0.5),
                 adaptive: bool | None = None, kind: str | None = None, boundary: str = "library",
                 steps: int = 500, merged_h: str = "reference") -> Scene:
    """One scene of the batched / served workloads: log-uniform size, dam break
    (fill height U[0.2, 0.6] of the box) or sloshing tank, refinement on or off."""
    rng = np.random.default_rng(seed)
    target = int(math.exp(rng.uniform(math.log(n_min), math.log(n_max))))
    fill = rng.uniform(0.2, 0.6)
    if kind is None:
        kind = "dam" if rng.uniform() < 0.5 else "slosh"
    if adaptive is None:
        adaptive = bool(rng.uniform() < 0.5)
    height = fill * box[1]
    width_x = 0.4 * box[0] if kind == "dam" else box[0]
    vol = width_x * height * box[2]
    dx = (vol / target) ** (1.0 / 3.0)
    n = (max(2, round(width_x / dx)), max(2, round(height / dx)), max(2, round(box[2] / dx)))
    if kind == "dam":
        sc = dam_break(n, dx, box, seed=seed, steps=steps, boundary=boundary, name=f"dam{seed}")
        if adaptive:
            r = abi.default_refinement()
            r.enabled = 1
            r.max_particle_count = 2 * sc.count
            sc.refinement = r
            sc.capacity = 2 * sc.count
    else:
        sc = sloshing_tank(n, dx, box, seed=seed, steps=steps, adaptive=adaptive, max_count=2 * n[0] * n[1] * n[2],
                           boundary=boundary, name=f"slosh{seed}")
        if adaptive:
            sc.refinement.zone.radius = 0.1 * box[0]
    sc.meta.update({"kind": kind, "fill": fill, "target": target})
    if merged_h == "kernel" and adaptive:
        kernel_merged_h(sc)
    return sc


def cfg4_scenes(count: int = 256, boundary: str = "library", steps: int = 500,
                merged_h: str = "kernel") -> list[Scene]:
    """256 independent scenes, per-scene seed 1000 + i (about 13-14M particles)."""
    return [random_scene(1000 + i, boundary=boundary, steps=steps, merged_h=merged_h) for i in range(count)]
