"""Continuous-batching server on the GPU: requests that arrive while others
are stepping are admitted between steps as pool space frees up, and each
request's final state is bit-identical to the same scene stepped alone."""
import pytest

from adaptivesph_amd import scene, serve
from adaptivesph_amd.simulation import Simulation
from helpers import assert_bitexact

pytestmark = pytest.mark.gpu

RESULT = ("positions", "velocities", "densities")


def _request_scene(req):
    # small scenes of the served mix: dam break or sloshing tank, refinement
    # on or off, decided by the seed
    s = scene.random_scene(req.seed, n_min=2_000, n_max=9_000, box=(0.3, 0.2, 0.2), steps=req.steps)
    return s


def _alone(s, steps):
    sim = Simulation.from_scene(s)
    for _ in range(steps):
        sim.update(s.dt)
    out = sim.fluid()
    sim.close()
    return out


def test_served_results_equal_alone():
    reqs = [serve.Request(rid=i, seed=70_000 + i, steps=st, arrival=arr)
            for i, (st, arr) in enumerate([(30, 0), (12, 0), (25, 3), (8, 5), (20, 6), (15, 30), (10, 31)])]
    made = [_request_scene(r) for r in reqs]
    assert any(s.adaptive for s in made) and any(not s.adaptive for s in made)
    # pool for about three of these scenes at once, so later requests queue
    cap = sorted(s.capacity for s in made)[-3:]
    pools = (sum(cap) + 16, sum(len(s.boundary) for s in made), 1 << 20)
    srv = serve.Server(1 << 40, max_scenes=8, pools=pools, keep_results=True)
    stats = srv.run(reqs, _request_scene, clock=serve.StepClock())
    srv.close()
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    # the mix actually overlapped and queued
    assert stats["resident_max"] >= 2
    assert any(r.admitted > r.arrival for r in srv.completed)
    for r in srv.completed:
        alone = _alone(_request_scene(r), r.steps)
        assert_bitexact(r.result, alone, f"request {r.rid} ({r.scene.name}, {r.steps} steps, "
                                         f"admitted at step {r.admitted})", fields=RESULT)


def test_serving_wall_clock_run_completes():
    reqs = serve.poisson_requests(rate=40.0, duration=1.0, seed=9, steps=(5, 40), base_seed=71_000)
    srv = serve.Server(512 << 20, max_scenes=32)
    stats = srv.run(reqs, lambda r: scene.random_scene(r.seed, n_min=2_000, n_max=20_000, steps=r.steps))
    srv.close()
    assert stats["completed"] + stats["rejected"] == len(reqs)
    assert stats["step_ms_p99"] >= stats["step_ms_p50"] > 0
    assert stats["completion_p99"] >= stats["completion_p50"] > 0


def test_served_cfg5_sizes_against_alone_and_oracle():
    """Requests drawn as the serving workload draws them (seeded scenes of
    log-uniform 10k-500k particles, dam break or sloshing tank, refinement
    on or off), admitted against the 200 GB pool budget with overlapping
    arrivals: each result equals the scene stepped alone, and the smallest
    one equals the FP32 oracle."""
    from oracle import oracle

    reqs = [serve.Request(rid=i, seed=90_000 + i, steps=st, arrival=arr)
            for i, (st, arr) in enumerate([(12, 0), (8, 2), (10, 4), (6, 5)])]

    def make(r):
        return serve.serving_scene(r, merged_h="reference")

    made = [make(r) for r in reqs]
    assert min(s.count for s in made) >= 10_000 and max(s.count for s in made) <= 500_000
    srv = serve.Server(200 * 10**9, max_scenes=8, keep_results=True)
    stats = srv.run(reqs, make, clock=serve.StepClock())
    srv.close()
    assert stats["completed"] == len(reqs) and stats["resident_max"] >= 2
    for r in srv.completed:
        assert_bitexact(r.result, _alone(make(r), r.steps), f"request {r.rid} ({r.scene.name})", fields=RESULT)
    small = min(srv.completed, key=lambda r: r.scene.count)
    ref = oracle.OracleSim.from_scene(make(small))
    for _ in range(small.steps):
        ref.update(small.scene.dt)
    assert_bitexact(small.result, ref.fluid(), f"request {small.rid} against the oracle", fields=RESULT)
