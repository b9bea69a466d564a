"""Continuous-batching server of SPH scenes on one GPU.

Requests (a seeded scene and a step count) arrive over time. A request is
admitted between two steps as soon as the device-memory budget, the batch's
particle/boundary/cell pools and a scene slot allow it; from then on it is
stepped together with every other resident scene (each kernel launch of a step
covers all of them) and its state stays in device memory until it has run its
steps. Then its final state (positions, velocities, densities) is read back,
and its slot and pool ranges are released for the next request. Scene
construction (lattice and wall generation) runs on host threads while the
device steps.

Admission is first-come first-served: the oldest waiting request is admitted
first, and one that does not fit blocks younger ones, so large scenes are not
starved. A request that could never fit (larger than the whole pool) is
rejected.

Reported: completed scenes/s, particle-updates/s, p50/p99 of the per-step
device time (HIP events around each step), of the host loop iteration, and of
the completion latency (arrival to result on the host).
"""
from __future__ import annotations

import math
import time
from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import abi
from . import scene as scenes
from .batch import Batch, grid_cells

RESULT_FIELDS = abi.FIELD_POSITIONS | abi.FIELD_VELOCITIES | abi.FIELD_DENSITIES


@dataclass
class Request:
    rid: int
    seed: int
    steps: int
    arrival: float  # seconds (wall clock) or steps (virtual clock) since the start
    scene: scenes.Scene | None = None
    bytes: int = 0
    slot: int | None = None
    admitted: float | None = None
    finished: float | None = None
    done: int = 0
    rejected: bool = False
    result: dict | None = None
    future: Future | None = field(default=None, repr=False)


def percentile(xs, q):
    if not xs:
        return None
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def unit_bytes(list_capacity: int = 0, adaptive: bool = True) -> tuple[int, int, int, int]:
    """Device bytes of one fluid slot, one boundary slot and one grid cell of a
    scene inside a batch, and the fixed per-scene part (from the library)."""
    lib = abi.load_library()
    a = int(adaptive)
    base = int(lib.asph_scene_bytes(0, 0, 0, a, list_capacity))
    fp = int(lib.asph_scene_bytes(1, 0, 0, a, list_capacity)) - base
    fb = int(lib.asph_scene_bytes(0, 1, 0, a, list_capacity)) - base
    fc = int(lib.asph_scene_bytes(0, 0, 1, a, list_capacity)) - base
    return fp, fb, fc, base


def pools_for_budget(budget_bytes: int, list_capacity: int = 0, adaptive: bool = True,
                     boundary_ratio: float = 0.5, cell_ratio: float = 0.25) -> tuple[int, int, int]:
    """Pool sizes (fluid slots, boundary slots, cells) whose device footprint
    fits `budget_bytes`, with boundary and cell pools sized per fluid slot;
    the fluid pool is capped at the library's particle capacity limit."""
    fp, fb, fc, _ = unit_bytes(list_capacity, adaptive)
    per = fp + boundary_ratio * fb + cell_ratio * fc
    # 2 cells beyond the ratio; the library refuses fluid pools above its capacity limit
    p = int(min((budget_bytes - 2 * fc) // per, abi.MAX_PARTICLE_CAPACITY))
    return p, int(p * boundary_ratio), int(p * cell_ratio) + 2


class WallClock:
    virtual = False

    def __init__(self):
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def idle(self, until: float | None) -> None:
        time.sleep(0.0005 if until is None else max(0.0, min(until - self.now(), 0.002)))

    def stepped(self) -> None:
        pass


class StepClock:
    """Virtual time measured in steps: arrivals are given in steps, and the
    clock jumps to the next arrival when nothing is resident."""

    virtual = True

    def __init__(self):
        self.t = 0.0

    def now(self) -> float:
        return self.t

    def idle(self, until: float | None) -> None:
        if until is not None:
            self.t = max(self.t, until)

    def stepped(self) -> None:
        self.t += 1.0


class Server:
    """Continuous batching of scenes in one Batch against a device-memory budget."""

    def __init__(self, budget_bytes: int, max_scenes: int = 256, list_capacity: int = 0,
                 boundary_ratio: float = 0.5, cell_ratio: float = 0.25, keep_results: bool = False,
                 workers: int = 4, graph: bool = True, pools: tuple[int, int, int] | None = None,
                 batch=None):
        self.budget = int(budget_bytes)
        self.list_capacity = list_capacity
        p, b, c = pools or pools_for_budget(self.budget, list_capacity, True, boundary_ratio, cell_ratio)
        self.pools = (p, b, c)
        # `batch` replaces the device batch (an object with the Batch methods
        # add/remove/step_timed/step_particles/fluid/close), for host tests.
        self.batch = batch if batch is not None else Batch(p, b, c, max_scenes, adaptive=True, graph=graph,
                                                           list_capacity=list_capacity)
        self.max_scenes = max_scenes
        self.used_bytes = 0
        self.resident: dict[int, Request] = {}
        self.waiting: deque[Request] = deque()
        self.completed: list[Request] = []
        self.rejected: list[Request] = []
        self.step_ms: list[float] = []
        self.loop_ms: list[float] = []
        self.resident_trace: list[int] = []
        self.particle_updates = 0
        self.keep_results = keep_results
        self.pool = ThreadPoolExecutor(max_workers=workers)

    def close(self):
        self.pool.shutdown(wait=True, cancel_futures=True)
        self.batch.close()

    # ------------------------------------------------------------ intake --
    def submit(self, req: Request, make_scene) -> None:
        req.future = self.pool.submit(make_scene, req)
        self.waiting.append(req)

    def scene_bytes(self, s: scenes.Scene) -> int:
        lib = abi.load_library()
        return int(lib.asph_scene_bytes(s.capacity, len(s.boundary), grid_cells(s.params), 1, self.list_capacity))

    def _never_fits(self, s: scenes.Scene, nbytes: int) -> bool:
        p, b, c = self.pools
        return (nbytes > self.budget or s.capacity > p or len(s.boundary) > b
                or grid_cells(s.params) + 1 > c)

    def admit(self, now: float, virtual: bool = False) -> int:
        n = 0
        while self.waiting:
            req = self.waiting[0]
            if not req.future.done() and not virtual:
                break
            s = req.future.result()
            req.scene = s
            req.bytes = self.scene_bytes(s)
            if self._never_fits(s, req.bytes):
                self.waiting.popleft()
                req.rejected = True
                req.scene = None
                self.rejected.append(req)
                continue
            if len(self.resident) >= self.max_scenes or self.used_bytes + req.bytes > self.budget:
                break
            try:
                req.slot = self.batch.add(s)
            except abi.AsphError as e:
                if e.status == abi.ASPH_ERR_CAPACITY:
                    break  # pool ranges free up as resident scenes complete
                raise
            self.waiting.popleft()
            req.admitted = now
            self.used_bytes += req.bytes
            self.resident[req.slot] = req
            n += 1
        return n

    # -------------------------------------------------------------- step --
    def step(self, clock) -> None:
        t = time.perf_counter()
        self.step_ms.append(self.batch.step_timed())
        self.particle_updates += self.batch.step_particles()
        self.resident_trace.append(len(self.resident))
        clock.stepped()
        finished = []
        for slot, req in self.resident.items():
            req.done += 1
            if req.done >= req.steps:
                finished.append(slot)
        for slot in finished:
            req = self.resident.pop(slot)
            result = self.batch.fluid(slot, RESULT_FIELDS)
            req.result = result if self.keep_results else None
            self.batch.remove(slot)
            req.finished = clock.now()
            self.used_bytes -= req.bytes
            if not self.keep_results:
                req.scene = None
            self.completed.append(req)
        self.loop_ms.append((time.perf_counter() - t) * 1e3)

    # --------------------------------------------------------------- run --
    def run(self, requests: list[Request], make_scene, clock=None, deadline: float | None = None,
            progress=None, progress_every: float = 10.0) -> dict:
        """Submits each request at its arrival time and steps until all are
        done (or, with a deadline on the clock, until then: requests still
        resident or waiting are left unfinished). `progress(text)` is called
        every `progress_every` wall-clock seconds."""
        clock = clock or WallClock()
        wall0 = time.perf_counter()
        next_report = wall0 + progress_every
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        while pending or self.waiting or self.resident:
            t = clock.now()
            if deadline is not None and t >= deadline:
                break
            if progress is not None and time.perf_counter() >= next_report:
                next_report += progress_every
                progress(f"t={t:.1f}: {len(self.completed)} done, {len(self.resident)} resident, "
                         f"{len(self.waiting)} waiting, {len(pending)} not arrived, "
                         f"{self.used_bytes / 1e9:.2f} GB in use, {len(self.step_ms)} steps")
            while pending and pending[0].arrival <= t:
                self.submit(pending.popleft(), make_scene)
            self.admit(t, clock.virtual)
            if self.resident:
                self.step(clock)
            else:
                clock.idle(pending[0].arrival if pending and not self.waiting else None)
        wall = time.perf_counter() - wall0
        done = self.completed
        lat = [r.finished - r.arrival for r in done]
        return {
            "wall_s": wall,
            "requests": len(requests),
            "completed": len(done),
            "rejected": len(self.rejected),
            "unfinished": len(requests) - len(done) - len(self.rejected),
            "scenes_per_s": len(done) / wall if wall > 0 else 0.0,
            "particle_updates_per_s": self.particle_updates / wall if wall > 0 else 0.0,
            "step_ms_p50": percentile(self.step_ms, 50),
            "step_ms_p99": percentile(self.step_ms, 99),
            "loop_ms_p50": percentile(self.loop_ms, 50),
            "loop_ms_p99": percentile(self.loop_ms, 99),
            "completion_p50": percentile(lat, 50),
            "completion_p99": percentile(lat, 99),
            "queue_p50": percentile([r.admitted - r.arrival for r in done], 50),
            "resident_mean": float(np.mean(self.resident_trace)) if self.resident_trace else 0.0,
            "resident_max": max(self.resident_trace) if self.resident_trace else 0,
            "steps": len(self.step_ms),
        }


def poisson_requests(rate: float, duration: float, seed: int = 5, steps=(100, 2000),
                     base_seed: int = 50_000) -> list[Request]:
    """Open-loop Poisson arrivals at `rate` per unit time over `duration`;
    each request runs uniform[steps] steps (inclusive)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        i = len(out)
        out.append(Request(rid=i, seed=base_seed + i, steps=int(rng.integers(steps[0], steps[1] + 1)),
                           arrival=t))
    return out


def serving_scene(req: Request, n_min: int = 10_000, n_max: int = 500_000,
                  merged_h: str = "kernel") -> scenes.Scene:
    """The scene of a request: log-uniform size, dam break or sloshing tank,
    refinement on or off 50/50 (all decided by the request seed)."""
    return scenes.random_scene(req.seed, n_min=n_min, n_max=n_max, steps=req.steps, merged_h=merged_h)


def mean_log_uniform(n_min: float, n_max: float) -> float:
    """Mean of a log-uniform distribution on [n_min, n_max]."""
    return (n_max - n_min) / math.log(n_max / n_min)
