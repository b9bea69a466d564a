"""ctypes mirror of include/asph.h and loader of the HIP library.

The library is built in-tree into adaptivesph_amd/lib/libasph.so by
``__graft_entry__.build()``. Loading fails loudly when it is missing: there is
no CPU fallback for the simulation entry points.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libasph.so"

ASPH_OK = 0
ASPH_ERR_INVALID = 1
ASPH_ERR_CAPACITY = 2
ASPH_ERR_DEVICE = 3
ASPH_ERR_STATE = 4
ASPH_FLAG_FIX_CELL_START = 1
MAX_PARTICLE_CAPACITY = (1 << 28) - 64  # ASPH_MAX_PARTICLE_CAPACITY
NOALLOC_MAX_WALLS = 1 << 20  # ASPH_NOALLOC_MAX_WALLS
NOALLOC_MAX_BUCKETS = 1 << 22  # ASPH_NOALLOC_MAX_BUCKETS
FLAG_KERNEL_MERGED_H = 2
FLAG_TOLERANCE_MATH = 4

CRITERION = {"velocity": 0, "interface": 1, "vorticity": 2, "zone": 3}

FIELD_POSITIONS = 1
FIELD_VELOCITIES = 2
FIELD_ACCELERATIONS = 4
FIELD_DENSITIES = 8
FIELD_RADII = 16
FIELD_SMOOTHING_RADII = 32
FIELD_MASSES = 64
FIELD_ALL = 127


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Vec4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Domain(C.Structure):
    _fields_ = [("min", Vec3), ("max", Vec3), ("friction", C.c_float)]


class Parameters(C.Structure):
    """Simulation::Parameters (Simulation.cuh:64-109) plus appended extensions."""

    _fields_ = [
        ("domain", Domain),
        ("gravity", Vec3),
        ("speed_of_sound", C.c_float),
        ("rest_density", C.c_float),
        ("restitution", C.c_float),
        ("viscosity_constant", C.c_float),
        ("surface_tension_constant", C.c_float),
        ("max_velocity", C.c_float),
        ("base_smoothing_radius", C.c_float),
        ("base_particle_radius", C.c_float),
        ("base_particle_mass", C.c_float),
        ("threads_per_block", C.c_uint32),
        ("tait_gamma", C.c_float),
        ("slosh_amplitude", C.c_float),
        ("slosh_period", C.c_float),
        ("flags", C.c_uint32),
        ("laminar_viscosity", C.c_float),
    ]


class SplittingParameters(C.Structure):
    _fields_ = [
        ("epsilon", C.c_float),
        ("alpha", C.c_float),
        ("center_mass_ratio", C.c_float),
        ("vertex_mass_ratio", C.c_float),
    ]


class ThresholdPair(C.Structure):
    _fields_ = [("split", C.c_float), ("merge", C.c_float)]


class ZoneParameters(C.Structure):
    _fields_ = [
        ("center", Vec3),
        ("radius", C.c_float),
        ("path_radius", C.c_float),
        ("path_period", C.c_float),
    ]


class RefinementParameters(C.Structure):
    """RefinementParameters (RefinementParameters.cuh:80-104) plus the zone extension."""

    _fields_ = [
        ("enabled", C.c_uint32),
        ("min_mass_ratio", C.c_float),
        ("max_mass_ratio", C.c_float),
        ("max_particle_count", C.c_uint32),
        ("max_batch_ratio", C.c_float),
        ("initial_cooldown", C.c_uint32),
        ("cooldown", C.c_uint32),
        ("criterion_type", C.c_uint32),
        ("splitting", SplittingParameters),
        ("velocity", ThresholdPair),
        ("vorticity", ThresholdPair),
        ("curvature", ThresholdPair),
        ("interface_ratio", ThresholdPair),
        ("zone", ZoneParameters),
    ]


class DensityInfo(C.Structure):
    _fields_ = [
        ("rest_density", C.c_float),
        ("min_density", C.c_float),
        ("max_density", C.c_float),
        ("average_density", C.c_float),
        ("under_density_count", C.c_uint32),
        ("normal_density_count", C.c_uint32),
        ("over_density_count", C.c_uint32),
    ]


class BoundaryParticle(C.Structure):
    _fields_ = [("position", Vec4), ("psi", C.c_float)]


BOUNDARY_DTYPE = np.dtype([("position", np.float32, (4,)), ("psi", np.float32)])
assert BOUNDARY_DTYPE.itemsize == C.sizeof(BoundaryParticle) == 20


class FluidView(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("velocities", C.c_void_p),
        ("accelerations", C.c_void_p),
        ("densities", C.c_void_p),
        ("radii", C.c_void_p),
        ("smoothing_radii", C.c_void_p),
        ("masses", C.c_void_p),
        ("particle_count", C.c_uint32),
    ]


class RefinementTrace(C.Structure):
    _fields_ = [
        ("refined", C.c_uint32),
        ("count_before", C.c_uint32),
        ("merge_eligible", C.c_uint32),
        ("merge_pairs", C.c_uint32),
        ("count_after_merge", C.c_uint32),
        ("split_selected", C.c_uint32),
        ("split_performed", C.c_uint32),
        ("count_after", C.c_uint32),
    ]


def default_parameters() -> Parameters:
    """Reference struct defaults (Simulation.cuh:68-70,108) and extension defaults."""
    p = Parameters()
    p.domain.min = Vec3(-1.0, -1.0, -1.0)
    p.domain.max = Vec3(1.0, 1.0, 1.0)
    p.domain.friction = 0.01
    p.threads_per_block = 256
    p.tait_gamma = 5.0
    p.slosh_amplitude = 0.0
    p.slosh_period = 0.0
    p.flags = 0
    p.laminar_viscosity = 0.0
    return p


def default_refinement() -> RefinementParameters:
    """RefinementParameters defaults (RefinementParameters.cuh:8-104)."""
    r = RefinementParameters()
    r.enabled = 0
    r.min_mass_ratio = 0.3
    r.max_mass_ratio = 8.0
    r.max_particle_count = 1000000
    r.max_batch_ratio = 0.2
    r.initial_cooldown = 10
    r.cooldown = 10
    r.criterion_type = 0
    r.splitting = SplittingParameters(0.65, 0.70, 0.077, 0.077)
    r.velocity = ThresholdPair(2.0, 0.1)
    r.vorticity = ThresholdPair(1.0, 0.3)
    r.curvature = ThresholdPair(2.0, 0.6)
    r.interface_ratio = ThresholdPair(0.05, 0.15)
    r.zone = ZoneParameters(Vec3(0.0, 0.0, 0.0), 0.0, 0.0, 0.0)
    return r


def as_ptr(a: np.ndarray | None, ctype=C.c_float):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


_LIB = None


def load_library() -> C.CDLL:
    """Load libasph.so (the HIP product library); raise if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()'). "
            "There is no CPU fallback."
        )
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    _declare(lib)
    _LIB = lib
    return lib


def _declare(lib: C.CDLL) -> None:
    vp = C.c_void_p
    u32 = C.c_uint32
    f32 = C.c_float
    P = C.POINTER
    sigs = {
        "asph_last_error": (C.c_char_p, []),
        "asph_default_parameters": (C.c_int, [P(Parameters)]),
        "asph_default_refinement_parameters": (C.c_int, [P(RefinementParameters)]),
        "asph_boundary_generate": (C.c_int, [P(Domain), f32, f32, f32, vp, P(u32)]),
        "asph_boundary_generate_device": (C.c_int, [P(Domain), f32, f32, f32, vp, P(u32)]),
        "asph_boundary_generate_device_noalloc": (C.c_int, [P(Domain), f32, f32, f32, vp, P(u32), vp]),
        "asph_create": (
            C.c_int,
            [P(Parameters), vp, u32, vp, u32, P(RefinementParameters), u32, u32, P(vp)],
        ),
        "asph_create_imported": (
            C.c_int,
            [P(Parameters), vp, u32, P(FluidBuffers), P(BoundaryBuffers), vp, u32, P(RefinementParameters), u32, u32,
             P(vp)],
        ),
        "asph_buffer_allocate": (C.c_int, [C.c_size_t, P(DeviceBuffer)]),
        "asph_buffer_release": (C.c_int, [P(DeviceBuffer)]),
        "asph_buffer_download": (C.c_int, [P(DeviceBuffer), C.c_size_t, vp, C.c_size_t]),
        "asph_buffer_upload": (C.c_int, [P(DeviceBuffer), C.c_size_t, vp, C.c_size_t]),
        "asph_get_boundary_view": (C.c_int, [vp, P(BoundaryView)]),
        "asph_destroy": (C.c_int, [vp]),
        "asph_update": (C.c_int, [vp, f32]),
        "asph_update_domain": (C.c_int, [vp, P(Domain), vp, u32]),
        "asph_update_domain_walls": (C.c_int, [vp, P(Domain), f32]),
        "asph_fluid_count": (C.c_int, [vp, P(u32)]),
        "asph_boundary_count": (C.c_int, [vp, P(u32)]),
        "asph_average_neighbor_count": (C.c_int, [vp, P(f32)]),
        "asph_get_density_info": (C.c_int, [vp, f32, P(DensityInfo)]),
        "asph_enable_adaptive_refinement": (C.c_int, [vp]),
        "asph_get_fluid_view": (C.c_int, [vp, P(FluidView)]),
        "asph_copy_fluid": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp, vp]),
        "asph_debug_set_fluid": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp, vp]),
        "asph_debug_set_frame": (C.c_int, [vp, u32]),
        "asph_debug_grid_dims": (C.c_int, [vp, P(C.c_int32)]),
        "asph_debug_grid": (C.c_int, [vp, vp, vp, vp, vp]),
        "asph_debug_neighbor_sets": (C.c_int, [vp, vp, vp]),
        "asph_debug_refinement": (C.c_int, [vp, P(RefinementTrace), vp, vp, vp, vp]),
        "asph_debug_division_check": (C.c_int, [C.c_uint64, C.c_uint64, P(C.c_uint64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in _extra_sigs().items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def _extra_sigs():
    """Batch / serving / profiling entry points (include/asph_batch.h)."""
    vp = C.c_void_p
    u32 = C.c_uint32
    u64 = C.c_uint64
    f32 = C.c_float
    P = C.POINTER
    return {
        "asph_batch_create": (C.c_int, [P(BatchConfig), P(vp)]),
        "asph_batch_destroy": (C.c_int, [vp]),
        "asph_batch_add_scene": (
            C.c_int,
            [vp, P(Parameters), vp, u32, vp, u32, P(RefinementParameters), u32, f32, P(u32)],
        ),
        "asph_batch_remove_scene": (C.c_int, [vp, u32]),
        "asph_batch_update_domain_walls": (C.c_int, [vp, u32, P(Domain), f32]),
        "asph_batch_step": (C.c_int, [vp, u32]),
        "asph_batch_synchronize": (C.c_int, [vp]),
        "asph_batch_step_timed": (C.c_int, [vp, P(f32)]),
        "asph_batch_scene_count": (C.c_int, [vp, u32, P(u32)]),
        "asph_batch_average_neighbor_count": (C.c_int, [vp, u32, P(f32)]),
        "asph_batch_copy_scene": (C.c_int, [vp, u32, u32, vp, vp, vp, vp, vp, vp, vp]),
        "asph_batch_set_launch": (C.c_int, [vp, u32, u32, u32, u32]),
        "asph_batch_set_schedule": (C.c_int, [vp, C.c_int32, u32]),
        "asph_batch_memory": (C.c_int, [vp, P(u64), P(u64)]),
        "asph_batch_step_particles": (C.c_int, [vp, P(u32)]),
        "asph_batch_profile_step": (C.c_int, [vp, P(f32), u32, P(u32)]),
        "asph_batch_profile_step_async": (C.c_int, [vp]),
        "asph_batch_profile_collect": (C.c_int, [vp, P(f32), u32, P(u32), P(u32)]),
        "asph_profile_stage_name": (C.c_char_p, [u32]),
        "asph_scene_bytes": (u64, [u32, u32, u32, u32, u32]),
        "asph_grid_cells": (C.c_int, [P(Domain), f32, P(u32)]),
        "asph_device_info": (C.c_int, [C.c_char_p, u32]),
    }


class DeviceBuffer(C.Structure):
    """asph_device_buffer: a device pointer and its size in bytes."""

    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t)]


class FluidBuffers(C.Structure):
    """asph_fluid_buffers (FluidParticlesDataImportedBuffer)."""

    _fields_ = [("positions", DeviceBuffer), ("velocities", DeviceBuffer), ("densities", DeviceBuffer),
                ("radii", DeviceBuffer)]


class BoundaryBuffers(C.Structure):
    """asph_boundary_buffers (BoundaryParticlesDataImportedBuffer)."""

    _fields_ = [("positions", DeviceBuffer), ("radii", DeviceBuffer), ("colors", DeviceBuffer)]


class BoundaryView(C.Structure):
    """asph_boundary_view (BoundaryParticlesData)."""

    _fields_ = [("positions", C.c_void_p), ("colors", C.c_void_p), ("psi_values", C.c_void_p),
                ("viscosity_coefficients", C.c_void_p), ("radii", C.c_void_p), ("particle_count", C.c_uint32)]


class BatchConfig(C.Structure):
    _fields_ = [
        ("max_particles", C.c_uint32),
        ("max_boundary", C.c_uint32),
        ("max_cells", C.c_uint32),
        ("max_scenes", C.c_uint32),
        ("list_capacity", C.c_uint32),
        ("flags", C.c_uint32),
    ]


BATCH_GRAPH = 1
BATCH_ADAPTIVE = 2


class AsphError(RuntimeError):
    """A non-zero status of the C ABI; `status` is the ASPH_ERR_* code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def check(lib: C.CDLL, rc: int, what: str) -> None:
    if rc != ASPH_OK:
        msg = lib.asph_last_error().decode(errors="replace")
        raise AsphError(rc, f"{what} failed (status {rc}): {msg}")
