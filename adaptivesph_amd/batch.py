"""Batched stepping of many independent scenes on one GPU (include/asph_batch.h).

All scenes share one set of device pools; every kernel of a step covers all
scenes through a scene table, and a step can be captured once as a hipGraph
and replayed. Scenes can be added and removed between steps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .scene import Scene


def grid_cells(params: abi.Parameters) -> int:
    lib = abi.load_library()
    n = C.c_uint32()
    abi.check(lib, lib.asph_grid_cells(C.byref(params.domain), params.base_smoothing_radius, C.byref(n)), "grid cells")
    return int(n.value)


def scene_bytes(s: Scene, list_capacity: int = 0) -> int:
    """Device bytes of scene `s` inside a batch with `list_capacity` list
    entries per particle (0: the default of a batch of this scene alone)."""
    lib = abi.load_library()
    return int(lib.asph_scene_bytes(s.capacity, len(s.boundary), grid_cells(s.params), int(s.adaptive),
                                    list_capacity))


class Batch:
    def __init__(self, max_particles: int, max_boundary: int, max_cells: int, max_scenes: int,
                 adaptive: bool = False, graph: bool = True, list_capacity: int = 0):
        # list_capacity 0: the library default (128, or 640 for adaptive batches)
        self._lib = abi.load_library()
        cfg = abi.BatchConfig()
        cfg.max_particles = int(max_particles)
        cfg.max_boundary = int(max_boundary)
        cfg.max_cells = int(max_cells)
        cfg.max_scenes = int(max_scenes)
        cfg.list_capacity = int(list_capacity)
        cfg.flags = (abi.BATCH_GRAPH if graph else 0) | (abi.BATCH_ADAPTIVE if adaptive else 0)
        h = C.c_void_p()
        abi.check(self._lib, self._lib.asph_batch_create(C.byref(cfg), C.byref(h)), "batch create")
        self._h = h
        self.list_capacity = list_capacity

    @classmethod
    def for_scenes(cls, scenes: list[Scene], graph: bool = True, slack: float = 1.0, **kw) -> "Batch":
        cap = int(sum(s.capacity for s in scenes) * slack) + 1
        bcap = int(sum(len(s.boundary) for s in scenes) * slack) + 1
        cells = int(sum(grid_cells(s.params) + 1 for s in scenes) * slack) + 1
        adaptive = any(s.adaptive for s in scenes)
        return cls(cap, bcap, cells, max(len(scenes), 1), adaptive=adaptive, graph=graph, **kw)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.asph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        abi.check(self._lib, rc, what)

    def add(self, s: Scene) -> int:
        sid = C.c_uint32()
        self._check(self._lib.asph_batch_add_scene(
            self._h, C.byref(s.params), s.positions.ctypes.data, s.count,
            s.boundary.ctypes.data if len(s.boundary) else None, len(s.boundary),
            C.byref(s.refinement) if s.refinement is not None else None, s.capacity, s.dt, C.byref(sid)),
            f"add scene {s.name}")
        return int(sid.value)

    def remove(self, sid: int) -> None:
        self._check(self._lib.asph_batch_remove_scene(self._h, int(sid)), "remove scene")

    def update_domain_walls(self, sid: int, domain: abi.Domain, spacing: float) -> None:
        """Moves scene `sid`'s box; its walls are rebuilt on the GPU."""
        self._check(self._lib.asph_batch_update_domain_walls(self._h, int(sid), C.byref(domain), float(spacing)),
                    "update domain walls")

    def step(self, steps: int = 1) -> None:
        self._check(self._lib.asph_batch_step(self._h, int(steps)), "step")

    def step_timed(self) -> float:
        ms = C.c_float()
        self._check(self._lib.asph_batch_step_timed(self._h, C.byref(ms)), "timed step")
        return float(ms.value)

    def synchronize(self) -> None:
        self._check(self._lib.asph_batch_synchronize(self._h), "synchronize")

    def step_particles(self) -> int:
        n = C.c_uint32()
        self._check(self._lib.asph_batch_step_particles(self._h, C.byref(n)), "step particles")
        return int(n.value)

    def count(self, sid: int) -> int:
        n = C.c_uint32()
        self._check(self._lib.asph_batch_scene_count(self._h, int(sid), C.byref(n)), "count")
        return int(n.value)

    def average_neighbor_count(self, sid: int) -> float:
        """calculateAverageNeighborCount of scene `sid` (grid of the last step)."""
        v = C.c_float()
        self._check(self._lib.asph_batch_average_neighbor_count(self._h, int(sid), C.byref(v)),
                    "average neighbour count")
        return float(v.value)

    def fluid(self, sid: int, mask: int = abi.FIELD_ALL) -> dict[str, np.ndarray]:
        n = self.count(sid)
        out = {
            "positions": np.zeros((n, 4), np.float32),
            "velocities": np.zeros((n, 4), np.float32),
            "accelerations": np.zeros((n, 4), np.float32),
            "densities": np.zeros(n, np.float32),
            "radii": np.zeros(n, np.float32),
            "smoothing_radii": np.zeros(n, np.float32),
            "masses": np.zeros(n, np.float32),
        }
        self._check(self._lib.asph_batch_copy_scene(self._h, int(sid), mask, *(a.ctypes.data for a in out.values())),
                    "copy scene")
        return out

    def set_launch(self, block_stream: int = 256, radix_bits: int = 0, graph: bool = True,
                   block_neighbor: int = 256) -> None:
        self._check(self._lib.asph_batch_set_launch(self._h, block_stream, block_neighbor, radix_bits, int(graph)),
                    "set launch")

    def set_schedule(self, query_tile: int = -1, force_blocks_per_cu: int = 0) -> None:
        """Query order of the neighbour kernels (0 cell order; 1, 2: 2x2x2 or
        4x4x4 sub-cell tiles; -1 the default) and the force kernels' grid (0
        one thread per particle; k: k persistent blocks per compute unit)."""
        self._check(self._lib.asph_batch_set_schedule(self._h, int(query_tile), int(force_blocks_per_cu)),
                    "set schedule")

    def memory(self) -> tuple[int, int]:
        a = C.c_uint64()
        u = C.c_uint64()
        self._check(self._lib.asph_batch_memory(self._h, C.byref(a), C.byref(u)), "memory")
        return int(a.value), int(u.value)

    def profile_step(self) -> dict[str, float]:
        ms = (C.c_float * 32)()
        n = C.c_uint32()
        self._check(self._lib.asph_batch_profile_step(self._h, ms, 32, C.byref(n)), "profile step")
        return {self._lib.asph_profile_stage_name(i).decode(): float(ms[i]) for i in range(n.value)}

    def profile_step_async(self) -> None:
        """An eager step with HIP events between its stages, not waited for."""
        self._check(self._lib.asph_batch_profile_step_async(self._h), "profile step")

    def profile_collect(self) -> tuple[dict[str, float], int]:
        """Stage times (ms) averaged over the steps of profile_step_async
        since the last collection, and their number (waits for the stream)."""
        ms = (C.c_float * 32)()
        n = C.c_uint32()
        k = C.c_uint32()
        self._check(self._lib.asph_batch_profile_collect(self._h, ms, 32, C.byref(n), C.byref(k)), "profile collect")
        return {self._lib.asph_profile_stage_name(i).decode(): float(ms[i]) for i in range(n.value)}, int(k.value)
