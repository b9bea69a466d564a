"""Python mirror of the reference's simulation interface
(cuda/include/cuda/simulation/Simulation.cuh:61-145) over the C ABI.

``Simulation`` keeps the reference's names and argument meaning:
create from parameters + positions + boundary domain (+ optional refinement
parameters and capacities), ``update(dt)`` (synchronous), the count getters,
``calculate_average_neighbor_count``, ``get_density_info(threshold)``,
``enable_adaptive_refinement``, and particle-buffer access (device pointers in
FluidParticlesData layout via ``fluid_view`` and a host download via
``fluid``). Errors raise RuntimeError, as the reference raises C++ exceptions.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .scene import Scene


class Simulation:
    def __init__(self, params: abi.Parameters, positions: np.ndarray, boundary: np.ndarray,
                 refinement: abi.RefinementParameters | None = None, max_fluid_capacity: int | None = None,
                 max_boundary_capacity: int | None = None):
        self._lib = abi.load_library()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 4)
        b = np.ascontiguousarray(boundary, dtype=abi.BOUNDARY_DTYPE)
        cap = pos.shape[0] if max_fluid_capacity is None else int(max_fluid_capacity)
        bcap = max(len(b), 1) if max_boundary_capacity is None else int(max_boundary_capacity)
        self._params = params
        handle = C.c_void_p()
        rc = self._lib.asph_create(C.byref(params), pos.ctypes.data, pos.shape[0],
                                   b.ctypes.data if len(b) else None, len(b),
                                   C.byref(refinement) if refinement is not None else None, cap, bcap,
                                   C.byref(handle))
        abi.check(self._lib, rc, "createSimulation")
        self._h = handle

    @classmethod
    def from_scene(cls, scene: Scene) -> "Simulation":
        return cls(scene.params, scene.positions, scene.boundary, scene.refinement, scene.capacity,
                   scene.boundary_capacity)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.asph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str) -> None:
        abi.check(self._lib, rc, what)

    # ----------------------------------------------------- reference API --
    def update(self, dt: float) -> None:
        self._check(self._lib.asph_update(self._h, float(dt)), "update")

    def update_domain(self, domain: abi.Domain, boundary: np.ndarray) -> None:
        b = np.ascontiguousarray(boundary, dtype=abi.BOUNDARY_DTYPE)
        self._check(self._lib.asph_update_domain(self._h, C.byref(domain), b.ctypes.data if len(b) else None,
                                                 len(b)), "updateDomain")

    def update_domain_walls(self, domain: abi.Domain, spacing: float) -> None:
        """updateDomain with StaticBoundaryDomain::generate's walls for `domain`
        (spacing, the rest density and h0) built on the GPU."""
        self._check(self._lib.asph_update_domain_walls(self._h, C.byref(domain), float(spacing)),
                    "updateDomain (device walls)")

    def get_fluid_particles_count(self) -> int:
        n = C.c_uint32()
        self._check(self._lib.asph_fluid_count(self._h, C.byref(n)), "getFluidParticlesCount")
        return n.value

    def get_boundary_particles_count(self) -> int:
        n = C.c_uint32()
        self._check(self._lib.asph_boundary_count(self._h, C.byref(n)), "getBoundaryParticlesCount")
        return n.value

    def calculate_average_neighbor_count(self) -> float:
        v = C.c_float()
        self._check(self._lib.asph_average_neighbor_count(self._h, C.byref(v)), "calculateAverageNeighborCount")
        return v.value

    def get_density_info(self, threshold: float) -> abi.DensityInfo:
        info = abi.DensityInfo()
        self._check(self._lib.asph_get_density_info(self._h, float(threshold), C.byref(info)), "getDensityInfo")
        return info

    def enable_adaptive_refinement(self) -> None:
        self._check(self._lib.asph_enable_adaptive_refinement(self._h), "enableAdaptiveRefinement")

    def fluid_view(self) -> abi.FluidView:
        v = abi.FluidView()
        self._check(self._lib.asph_get_fluid_view(self._h, C.byref(v)), "fluid view")
        return v

    # ----------------------------------------------------------- access --
    def fluid(self, mask: int = abi.FIELD_ALL) -> dict[str, np.ndarray]:
        n = self.get_fluid_particles_count()
        out = {
            "positions": np.zeros((n, 4), np.float32),
            "velocities": np.zeros((n, 4), np.float32),
            "accelerations": np.zeros((n, 4), np.float32),
            "densities": np.zeros(n, np.float32),
            "radii": np.zeros(n, np.float32),
            "smoothing_radii": np.zeros(n, np.float32),
            "masses": np.zeros(n, np.float32),
        }
        self._check(self._lib.asph_copy_fluid(self._h, mask, *(a.ctypes.data for a in out.values())), "copy fluid")
        return out

    # ------------------------------------------------------------ debug --
    def set_fluid(self, state: dict[str, np.ndarray]) -> None:
        n = int(state["positions"].shape[0])

        def ptr(key, dtype):
            a = state.get(key)
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            state[key] = a
            return a.ctypes.data

        self._check(self._lib.asph_debug_set_fluid(
            self._h, n, ptr("positions", np.float32), ptr("velocities", np.float32),
            ptr("accelerations", np.float32), ptr("densities", np.float32), ptr("radii", np.float32),
            ptr("smoothing_radii", np.float32), ptr("masses", np.float32)), "set fluid")

    def set_frame(self, frame: int) -> None:
        self._check(self._lib.asph_debug_set_frame(self._h, int(frame)), "set frame")

    def grid_dims(self) -> tuple[int, int, int]:
        d = (C.c_int32 * 3)()
        self._check(self._lib.asph_debug_grid_dims(self._h, d), "grid dims")
        return int(d[0]), int(d[1]), int(d[2])

    def grid(self) -> dict[str, np.ndarray]:
        n = self.get_fluid_particles_count()
        gx, gy, gz = self.grid_dims()
        out = {
            "keys": np.zeros(n, np.int32),
            "indices": np.zeros(n, np.int32),
            "cell_start": np.zeros(gx * gy * gz, np.int32),
            "cell_end": np.zeros(gx * gy * gz, np.int32),
        }
        self._check(self._lib.asph_debug_grid(self._h, *(a.ctypes.data for a in out.values())), "grid")
        return out

    def neighbor_sets(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.get_fluid_particles_count()
        counts = np.zeros(n, np.uint32)
        prints = np.zeros(n, np.uint64)
        self._check(self._lib.asph_debug_neighbor_sets(self._h, counts.ctypes.data, prints.ctypes.data),
                    "neighbor sets")
        return counts, prints

    def refinement_trace(self) -> dict:
        t = abi.RefinementTrace()
        self._check(self._lib.asph_debug_refinement(self._h, C.byref(t), None, None, None, None), "trace")
        ml = np.zeros(max(t.merge_eligible, 1), np.uint32)
        cand = np.zeros(max(t.count_before, 1), np.uint32)
        keep = np.zeros(max(t.merge_pairs, 1), np.uint32)
        split = np.zeros(max(t.split_performed, 1), np.uint32)
        self._check(self._lib.asph_debug_refinement(self._h, C.byref(t), ml.ctypes.data, cand.ctypes.data,
                                                    keep.ctypes.data, split.ctypes.data), "trace")
        out = {f: int(getattr(t, f)) for f, _ in abi.RefinementTrace._fields_}
        out["merge_list"] = ml[: t.merge_eligible]
        out["candidates"] = cand[: t.count_before] if t.count_before > 1 else np.zeros(t.count_before, np.uint32)
        out["pair_keep"] = keep[: t.merge_pairs]
        out["split_parents"] = split[: t.split_performed]
        return out
