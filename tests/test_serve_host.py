"""Admission policy of the continuous-batching server, on the host: a stand-in
batch with first-fit pools replaces the device batch (the library is loaded
only for its host functions: scene footprints and grid sizes)."""
import numpy as np
import pytest

from adaptivesph_amd import abi, scene, serve


class FakeBatch:
    """Pools of fluid slots with first-fit ranges, like the engine's."""

    def __init__(self, particles: int, slots: int = 64):
        self.free = [(0, particles)]
        self.slots = [None] * slots
        self.steps = 0
        self.log = []

    def add(self, s):
        n = s.capacity
        for k, (off, cnt) in enumerate(self.free):
            if cnt >= n:
                self.free[k] = (off + n, cnt - n)
                break
        else:
            raise abi.AsphError(abi.ASPH_ERR_CAPACITY, "particle pool exhausted")
        sid = self.slots.index(None)
        self.slots[sid] = (s, off, n)
        self.log.append(("add", s.name, self.steps))
        return sid

    def remove(self, sid):
        s, off, n = self.slots[sid]
        self.slots[sid] = None
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, c in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + c)
            else:
                merged.append((o, c))
        self.free = merged
        self.log.append(("remove", s.name, self.steps))

    def step_timed(self):
        self.steps += 1
        return 1.0

    def step_particles(self):
        return sum(x[0].count for x in self.slots if x is not None)

    def fluid(self, sid, mask=abi.FIELD_ALL):
        s = self.slots[sid][0]
        return {"positions": s.positions.copy()}

    def close(self):
        pass


def _scene(n, name, seed=1):
    s = scene.dam_break(n, 0.01, (0.3, 0.3, 0.2), seed=seed, steps=0, name=name)
    return s


@pytest.fixture(scope="module")
def sizes():
    return {"small": _scene((6, 6, 6), "small"), "big": _scene((12, 12, 8), "big")}


def _server(sizes, particles, budget=None, slots=64):
    fb = FakeBatch(particles, slots)
    budget = budget or 1 << 40
    srv = serve.Server(budget, max_scenes=slots, pools=(particles, 1 << 30, 1 << 30), batch=fb, workers=2)
    return srv, fb


def _make(sizes, kinds):
    def make(req):
        base = sizes[kinds[req.rid]]
        s = scene.Scene(**{**base.__dict__, "name": f"r{req.rid}"})
        return s
    return make


def test_fcfs_admission_and_release(sizes):
    cap = sizes["big"].capacity + sizes["small"].capacity
    srv, fb = _server(sizes, cap)
    kinds = ["big", "small", "big", "small"]
    reqs = [serve.Request(rid=i, seed=i, steps=[5, 3, 4, 2][i], arrival=0.0) for i in range(4)]
    stats = srv.run(reqs, _make(sizes, kinds), clock=serve.StepClock())
    srv.close()
    assert stats["completed"] == 4 and stats["rejected"] == 0
    adds = [(n, t) for op, n, t in fb.log if op == "add"]
    # r0 (big) and r1 (small) fill the pool; r2 (big) waits for r0 to leave even
    # after r1 completes (head of line), r3 is admitted after r2
    assert adds[:2] == [("r0", 0), ("r1", 0)]
    assert adds[2] == ("r2", 5)
    assert [n for n, _ in adds] == ["r0", "r1", "r2", "r3"]
    done = {r.rid: r for r in srv.completed}
    assert done[0].finished == 5 and done[1].finished == 3
    assert done[2].admitted == 5 and done[2].finished == 9
    # nothing left resident, pool whole again
    assert fb.free == [(0, cap)]
    assert all(x is None for x in fb.slots)
    assert stats["steps"] == max(r.finished for r in srv.completed)


def test_budget_bounds_resident_bytes(sizes):
    small = sizes["small"]
    srv, fb = _server(sizes, 1 << 30)
    one = srv.scene_bytes(small)
    srv.budget = 3 * one  # room for three small scenes
    reqs = [serve.Request(rid=i, seed=i, steps=4, arrival=0.0) for i in range(7)]
    peak = []
    orig = srv.step

    def step(clock):
        peak.append(srv.used_bytes)
        orig(clock)
    srv.step = step
    stats = srv.run(reqs, _make(sizes, ["small"] * 7), clock=serve.StepClock())
    srv.close()
    assert stats["completed"] == 7
    assert max(peak) == 3 * one
    assert stats["resident_max"] == 3


def test_scene_larger_than_pool_is_rejected(sizes):
    srv, fb = _server(sizes, sizes["small"].capacity)
    reqs = [serve.Request(rid=0, seed=0, steps=2, arrival=0.0), serve.Request(rid=1, seed=1, steps=2, arrival=1.0)]
    stats = srv.run(reqs, _make(sizes, ["big", "small"]), clock=serve.StepClock())
    srv.close()
    assert stats["rejected"] == 1 and stats["completed"] == 1
    assert srv.rejected[0].rid == 0


def test_late_arrivals_join_between_steps(sizes):
    srv, fb = _server(sizes, 1 << 30)
    reqs = [serve.Request(rid=0, seed=0, steps=10, arrival=0.0),
            serve.Request(rid=1, seed=1, steps=3, arrival=4.0),
            serve.Request(rid=2, seed=2, steps=2, arrival=20.0)]
    stats = srv.run(reqs, _make(sizes, ["small"] * 3), clock=serve.StepClock())
    srv.close()
    done = {r.rid: r for r in srv.completed}
    assert done[1].admitted == 4 and done[1].finished == 7  # overlapped with r0
    assert done[0].finished == 10
    assert done[2].admitted == 20 and done[2].finished == 22  # idle gap skipped
    assert stats["steps"] == 12


def test_poisson_requests_statistics():
    reqs = serve.poisson_requests(rate=50.0, duration=200.0, seed=3)
    arr = np.array([r.arrival for r in reqs])
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 1 / 50) < 0.05 / 50
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05  # exponential: cv = 1
    steps = np.array([r.steps for r in reqs])
    assert steps.min() >= 100 and steps.max() <= 2000
    assert abs(steps.mean() - 1050) < 20
    again = serve.poisson_requests(rate=50.0, duration=200.0, seed=3)
    assert [(r.arrival, r.steps, r.seed) for r in again] == [(r.arrival, r.steps, r.seed) for r in reqs]


def test_pools_fit_budget():
    budget = 200 * 10**9
    p, b, c = serve.pools_for_budget(budget)
    fp, fb, fc, base = serve.unit_bytes()
    assert p * fp + b * fb + c * fc <= budget
    assert p * fp + b * fb + c * fc > 0.99 * budget
    assert serve.mean_log_uniform(10_000, 500_000) == pytest.approx(125_253, rel=1e-4)


def test_pools_capped_at_the_library_capacity_limit():
    """A budget larger than the library's fluid pool limit (2^28 - 64 slots,
    about 230 GB at list capacity 128) sizes the pool to the limit instead of
    asking for a pool the library refuses."""
    for lc in (128, 640):
        p, b, c = serve.pools_for_budget(10**12, list_capacity=lc)
        assert p <= abi.MAX_PARTICLE_CAPACITY
    p, _, _ = serve.pools_for_budget(10**12, list_capacity=128)
    assert p == abi.MAX_PARTICLE_CAPACITY
