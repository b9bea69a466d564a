"""Batched stepping: every scene of a batch evolves exactly as it does alone
(bit-exact), with graph replay or eager launches, and scenes added or removed
between steps (continuous batching) do not perturb the others."""
import numpy as np
import pytest

from adaptivesph_amd import abi, scene
from adaptivesph_amd.batch import Batch
from adaptivesph_amd.simulation import Simulation
from helpers import assert_bitexact, boundary_scene, refinement

pytestmark = pytest.mark.gpu


def _mix():
    a = scene.dam_break((10, 14, 8), 0.01, (0.3, 0.3, 0.2), seed=41, steps=0, name="a")
    b = scene.sloshing_tank((12, 6, 8), 0.012, (0.2, 0.15, 0.12), seed=42, steps=0, adaptive=False, name="b")
    c = boundary_scene(seed=43)
    d = scene.dam_break((8, 8, 8), 0.01, (0.2, 0.2, 0.2), seed=44, steps=0, name="d")
    d.refinement = refinement("velocity", 6000, velocity=(0.03, 0.2), initial_cooldown=2, cooldown=3)
    d.capacity = 6000
    e = scene.sloshing_tank((12, 6, 8), 0.01, (0.2, 0.15, 0.12), seed=45, steps=0, max_count=9000, name="e")
    e.refinement.zone.radius = 0.04
    e.refinement.zone.path_radius = 0.03
    e.refinement.zone.path_period = 0.02
    return [a, b, c, d, e]


def _alone(s, steps):
    sim = Simulation.from_scene(s)
    for _ in range(steps):
        sim.update(s.dt)
    out = sim.fluid()
    sim.close()
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_alone(graph):
    scenes = _mix()
    batch = Batch.for_scenes(scenes, graph=graph)
    ids = [batch.add(s) for s in scenes]
    steps = 25
    batch.step(steps)
    batch.synchronize()
    for s, sid in zip(scenes, ids):
        assert_bitexact(batch.fluid(sid), _alone(s, steps), f"{s.name} graph={graph}")


def test_graph_replays_match_eager_every_step():
    scenes = _mix()
    g = Batch.for_scenes(scenes, graph=True)
    e = Batch.for_scenes(scenes, graph=False)
    gi = [g.add(s) for s in scenes]
    ei = [e.add(s) for s in scenes]
    for step in range(12):
        g.step(1)
        e.step(1)
        for s, a, b in zip(scenes, gi, ei):
            assert_bitexact(g.fluid(a), e.fluid(b), f"{s.name} replay {step}")


def test_profiled_steps_between_replays_keep_results():
    """Eager steps with stage events enqueued between graph replays without
    waiting (the benchmark's timed region) advance the scenes exactly as
    plain steps do; the collected stage times cover every such step."""
    scenes = _mix()
    g = Batch.for_scenes(scenes, graph=True)
    p = Batch.for_scenes(scenes, graph=True)
    gi = [g.add(s) for s in scenes]
    pi = [p.add(s) for s in scenes]
    g.step(9)
    for _ in range(3):
        p.step(2)
        p.profile_step_async()
    stages, samples = p.profile_collect()
    assert samples == 3
    assert {"neighbors", "force"} <= set(stages) and all(v >= 0.0 for v in stages.values())
    for s, a, b in zip(scenes, gi, pi):
        assert_bitexact(g.fluid(a), p.fluid(b), f"{s.name} after profiled steps")
    assert p.profile_collect()[1] == 0
    g.close()
    p.close()


def test_continuous_admission_and_removal():
    scenes = _mix()
    batch = Batch.for_scenes(scenes, graph=True, slack=1.5)
    first = [batch.add(s) for s in scenes[:3]]
    batch.step(10)
    late = [batch.add(s) for s in scenes[3:]]
    batch.step(7)
    batch.remove(first[1])
    batch.step(8)
    for s, sid in zip(scenes[3:], late):
        assert_bitexact(batch.fluid(sid), _alone(s, 15), f"late {s.name}")
    for s, sid, n in ((scenes[0], first[0], 25), (scenes[2], first[2], 25)):
        assert_bitexact(batch.fluid(sid), _alone(s, n), f"early {s.name}")
    # the freed region is reused
    again = batch.add(scenes[1])
    batch.step(5)
    assert_bitexact(batch.fluid(again), _alone(scenes[1], 5), "readmitted b")


def test_launch_shapes_do_not_change_results():
    scenes = _mix()[:3]
    ref = None
    for block, radix, nb in ((256, 8, 256), (128, 6, 64), (512, 4, 128), (64, 7, 256), (1024, 10, 64),
                             (256, 0, 128)):
        b = Batch.for_scenes(scenes, graph=True)
        ids = [b.add(s) for s in scenes]
        b.set_launch(block, radix, True, block_neighbor=nb)
        b.step(6)
        out = [b.fluid(i) for i in ids]
        if ref is None:
            ref = out
        else:
            for s, x, y in zip(scenes, out, ref):
                assert_bitexact(x, y, f"{s.name} block={block} radix={radix} neighbour block={nb}")
        b.close()


@pytest.mark.parametrize("tolerance", [False, True])
def test_schedules_do_not_change_results(tolerance):
    """Query tiles (cell order, 2x2x2 and 4x4x4 Morton sub-cell tiles, 4x4x4
    x-major slabs, automatic) and force grids (one thread per particle, persistent grids)
    give bit-identical results, for plain and refining scenes (whose cells
    exceed 64 particles once refined) in both arithmetic modes."""
    scenes = _mix()
    for s in scenes:
        if tolerance:
            s.params.flags |= abi.FLAG_TOLERANCE_MATH
    ref = None
    for tile, blocks in ((-1, 0), (0, 0), (1, 4), (2, 16), (0, 1), (1, 0), (3, 0), (3, 8), (4, 0)):
        b = Batch.for_scenes(scenes, graph=True)
        ids = [b.add(s) for s in scenes]
        b.set_schedule(tile, blocks)
        b.step(8)
        out = [b.fluid(i) for i in ids]
        if ref is None:
            ref = out
        else:
            for s, x, y in zip(scenes, out, ref):
                assert_bitexact(x, y, f"{s.name} query tile={tile} force blocks per CU={blocks}")
        b.close()


def test_schedule_arguments_are_checked():
    scenes = _mix()[:1]
    b = Batch.for_scenes(scenes, graph=False)
    b.add(scenes[0])
    for tile, blocks in ((5, 0), (-2, 0), (0, 65)):
        with pytest.raises(abi.AsphError):
            b.set_schedule(tile, blocks)
    b.close()


def test_neighbour_grid_covers_more_slots_than_capped_grids():
    """More particles than 16,384 blocks of 64 threads hold: the list builder
    (one particle per thread) still reaches every slot."""
    big = scene.cfg2()
    small = scene.dam_break((8, 10, 8), 0.01, (0.2, 0.2, 0.2), seed=46, steps=0, name="small")
    scenes = [big, small]
    b = Batch.for_scenes(scenes, graph=True)
    ids = [b.add(s) for s in scenes]
    b.set_launch(256, 0, True, block_neighbor=64)
    b.step(2)
    assert b.step_particles() > 16384 * 64
    for s, sid in zip(scenes, ids):
        assert_bitexact(b.fluid(sid), _alone(s, 2), f"{s.name} with a neighbour block of 64")
    b.close()


@pytest.mark.parametrize("list_capacity", [8, 64])
def test_list_overflow_paths_match_alone(list_capacity):
    """A list capacity far below the neighbour counts sends particles to the
    overflow kernels (the walk again); results stay those of the scenes run
    alone with the default capacity."""
    scenes = _mix()
    b = Batch.for_scenes(scenes, graph=True, list_capacity=list_capacity)
    ids = [b.add(s) for s in scenes]
    steps = 12
    b.step(steps)
    for s, sid in zip(scenes, ids):
        assert_bitexact(b.fluid(sid), _alone(s, steps), f"{s.name} list capacity {list_capacity}")
    b.close()


def test_benchmark_scene_mix_equals_alone():
    """Twelve scenes of the batched benchmark's generator (dam breaks and
    sloshing tanks, half refining, kernel-normalised merged h; sizes capped
    at 30k particles) stepped 15 times in one graph-replayed batch: every
    scene equals itself run alone."""
    scenes = [scene.random_scene(1000 + i, n_max=30_000, steps=15, merged_h="kernel") for i in range(12)]
    assert {s.meta["kind"] for s in scenes} == {"dam", "slosh"} and len({s.adaptive for s in scenes}) == 2
    b = Batch.for_scenes(scenes, graph=True)
    ids = [b.add(s) for s in scenes]
    b.step(15)
    for s, sid in zip(scenes, ids):
        assert_bitexact(b.fluid(sid), _alone(s, 15), f"{s.name}")
    b.close()


def _oracle_run(s, steps):
    from oracle import oracle

    ref = oracle.OracleSim.from_scene(s)
    for _ in range(steps):
        ref.update(s.dt)
    out = ref.fluid()
    ref.close()
    return out


def test_batch_scenes_equal_oracle():
    """Six scenes of different kinds in one graph-replayed batch (plain dam
    break, plain sloshing tank, walls-active dam break, refining dam break
    with the velocity criterion, refining sloshing tank with the moving zone,
    refining dam break with the reference's merged h): each equals the FP32
    oracle stepping that scene alone (SphSimulation::update,
    AdaptiveSphSimulation::update), bit for bit, after 14 steps."""
    scenes = _mix()
    f = scene.dam_break((8, 9, 7), 0.01, (0.2, 0.2, 0.2), seed=47, steps=0, name="f")
    f.refinement = refinement("velocity", 5000, velocity=(0.02, 0.3), initial_cooldown=1, cooldown=2)
    f.capacity = 5000
    scenes.append(f)
    assert sum(s.adaptive for s in scenes) == 3
    b = Batch.for_scenes(scenes, graph=True)
    ids = [b.add(s) for s in scenes]
    steps = 14
    b.step(steps)
    for s, sid in zip(scenes, ids):
        assert_bitexact(b.fluid(sid), _oracle_run(s, steps), f"{s.name} against the oracle")
    b.close()


def test_cfg4_generator_full_size():
    """The batched workload's full generator: 256 scenes of 5k-200k particles
    (about 14M; dam breaks and sloshing tanks, half refining), 3 graph-replayed
    steps in one batch. Every scene keeps a finite state and a plausible
    count; 8 sampled scenes equal themselves stepped alone, and two of them
    (one plain, one refining) equal the FP32 oracle."""
    scenes = scene.cfg4_scenes(256, steps=3)
    total = sum(s.count for s in scenes)
    assert 12_000_000 < total < 16_000_000 and sum(s.adaptive for s in scenes) > 100
    b = Batch.for_scenes(scenes, graph=True)
    ids = [b.add(s) for s in scenes]
    steps = 3
    b.step(steps)
    b.synchronize()
    counts = [b.count(i) for i in ids]
    for s, n in zip(scenes, counts):
        assert (n == s.count) if not s.adaptive else (0 < n <= s.capacity), s.name
    rng = np.random.default_rng(3)
    sample = sorted(rng.choice(len(scenes), 8, replace=False).tolist())
    for k in sample:
        s = scenes[k]
        got = b.fluid(ids[k])
        assert all(np.isfinite(got[f]).all() for f in ("positions", "velocities", "densities")), s.name
        assert_bitexact(got, _alone(s, steps), f"{s.name} alone")
    plain = next(k for k in sample if not scenes[k].adaptive)
    refining = next(k for k in sample if scenes[k].adaptive)
    for k in (plain, refining):
        assert_bitexact(b.fluid(ids[k]), _oracle_run(scenes[k], steps), f"{scenes[k].name} against the oracle")
    b.close()
