#!/usr/bin/env python3
"""Benchmark of the per-step SPH pipeline on one MI355X.

Default workload: the 1,048,576-particle WCSPH dam break (128 x 64 x 128
lattice, dx = 0.0039 m, 2.0 x 1.2 x 1.0 m box, FP32, no refinement), stepped
as one scene through the batched engine with the step captured as a hipGraph.
A step is one full pass of the pipeline (kick, drift, cell keys, sort, cell
ranges, gather, density, forces, kick) over all particles. The inputs are
resident in device memory before the timed region. Prints one JSON line.

Other workloads (--workload):
  cfg1      31,250-particle dam break (1.0 x 0.6 x 0.5 m box)
  adaptive  500k-particle sloshing tank refining every step (cap 4M)
  batch     256 scenes of 5k-200k particles, half adaptive, stepped together
            (graph-captured step, autotuned launch shape), beside the same
            scenes stepped one scene after another
  serve     Poisson request stream admitted against a 200 GB device budget
            (continuous batching), 10k-500k particles, 100-2000 steps each
  walls     updateDomain of the 1M dam break with its walls (851,860 wall
            particles) regenerated for a moved box: walls, psi and binning on
            the GPU, beside the host generator and host ingestion
  copy      1 GiB device copy (calibration of hardware byte counters)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle and launch (minimum compulsory traffic):
#   density: reads position 16 B, h 4 B, mass 4 B; writes density 4 B
#   force (+ fused second half kick): reads position, velocity (32 B),
#     density, h, mass (12 B); writes acceleration 16 B and velocity 16 B
ALGO_BYTES = {"density": 28, "force": 76}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str, merged_h: str = "kernel"):
    from adaptivesph_amd import scene

    if name == "cfg2":
        return [scene.cfg2()], "dam_break_1M"
    if name == "cfg2_lattice":
        # the same dam break on an unjittered lattice (as the reference
        # application lays out its particles): many pairs share a coordinate
        import dataclasses

        s = scene.cfg2()
        pos = scene.lattice(s.meta["n"], s.meta["dx"], (0.0, 0.0, 0.0), s.meta["seed"], jitter=0.0)
        return [dataclasses.replace(s, name="dam_break_1M_lattice", positions=pos)], "dam_break_1M_lattice"
    if name == "cfg1":
        return [scene.cfg1()], "dam_break_31k"
    if name == "adaptive":
        return [scene.cfg3(merged_h=merged_h)], "adaptive_slosh_500k"
    if name == "batch":
        with ThreadPoolExecutor(8) as ex:
            scenes = list(ex.map(lambda i: scene.random_scene(1000 + i, steps=500, merged_h=merged_h), range(256)))
        return scenes, "batch_256_scenes"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(scene_obj, steps: int) -> dict:
    """The oracle (FP32 host restatement, OpenMP) on a bounded sample."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
    o = oracle.OracleSim.from_scene(scene_obj)
    o.update(scene_obj.dt)  # warm-up step
    t0 = time.perf_counter()
    for _ in range(steps):
        o.update(scene_obj.dt)
    el = time.perf_counter() - t0
    return {
        "value": scene_obj.count * steps / el,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{scene_obj.name}: {scene_obj.count} particles, {steps} steps after 1 warm-up step, "
                  f"{el:.1f} s on {threads} OpenMP threads",
    }


def stage_profile(batch, steps: int) -> dict:
    """Per-stage device time (HIP events between the stages of eager steps on
    the engine's stream), averaged over `steps` steps."""
    stages = {}
    for _ in range(max(steps, 1)):
        for k, v in batch.profile_step().items():
            stages.setdefault(k, []).append(v)
    return {k: sum(v) / len(v) for k, v in stages.items()}


def roofline_of(stage_ms: dict, particles: int, workload: str) -> dict:
    # the density stage is the neighbour-list build plus the density kernel
    stage_ms = dict(stage_ms, density=stage_ms.get("neighbors", 0.0) + stage_ms.get("density", 0.0))
    dom = max(("density", "force"), key=lambda k: stage_ms.get(k, 0.0))
    dur = stage_ms[dom] * 1e-3
    achieved = ALGO_BYTES[dom] * particles / dur / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            t = json.loads(tfile.read_text()).get(workload, {}).get(dom)
            traffic = t.get("bytes_per_launch") if t else None
        except (ValueError, AttributeError):
            traffic = None
    return {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": ALGO_BYTES[dom] * particles,
        "launch_ms": round(stage_ms[dom], 4),
    }


def base_line(args, metric: str, value: float, unit: str, ms_per_step: float, config: dict) -> dict:
    return {
        "metric": metric,
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
    }


def run_single(args) -> dict:
    from adaptivesph_amd.batch import Batch

    scenes, workload = make_workload(args.workload, args.merged_h)
    s = scenes[0]
    batch = Batch.for_scenes(scenes, graph=not args.no_graph, list_capacity=args.list_capacity)
    sid = batch.add(s)
    log(f"workload {workload}: {s.count} particles")
    batch.step(args.warmup)
    batch.synchronize()
    stage_ms = stage_profile(batch, args.profile_steps)
    log("stage ms:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))
    batch.synchronize()
    particles = 0
    t0 = time.perf_counter()
    if s.adaptive:
        # the particle count changes every step: count each step's particles
        for _ in range(args.steps):
            batch.step(1)
            particles += batch.step_particles()
    else:
        batch.step(args.steps)
        batch.synchronize()
        particles = s.count * args.steps
    elapsed = time.perf_counter() - t0
    count = batch.count(sid)
    value = particles / elapsed
    metric = "particle-updates/s at 1M WCSPH" if workload == "dam_break_1M" else f"particle-updates/s ({workload})"
    out = base_line(args, metric, value, "particle-updates/s", elapsed / args.steps * 1e3,
                    {"workload": workload, "scenes": 1, "particles": s.count, "final_particles": count,
                     "graph": not args.no_graph, "dt": s.dt,
                     **({"merged_h": args.merged_h} if s.adaptive else {})})
    out["roofline"] = roofline_of(stage_ms, batch.step_particles(), workload)
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(s, args.cpu_steps)
    out["stage_ms"] = {k: round(v, 4) for k, v in stage_ms.items()}
    batch.close()
    return out


def run_batch(args) -> dict:
    """256 scenes stepped together, and the same scenes one after another
    (each in its own engine, graph-captured, synchronous steps as in the
    reference's update()). Both start from the same initial state and time the
    same steps (warmup .. warmup + steps) of every scene. The launch shape is
    chosen by the autotuner on a separate copy of the batch."""
    import numpy as np

    from adaptivesph_amd.autotune import autotune
    from adaptivesph_amd.batch import Batch

    t0 = time.perf_counter()
    scenes, workload = make_workload("batch", args.merged_h)
    total0 = sum(s.count for s in scenes)
    log(f"workload {workload}: {len(scenes)} scenes, {total0} particles, "
        f"{sum(s.adaptive for s in scenes)} adaptive ({time.perf_counter() - t0:.1f} s to build)")
    tuning = Batch.for_scenes(scenes, graph=True, list_capacity=args.list_capacity)
    for s in scenes:
        tuning.add(s)
    tune = autotune(tuning)
    tuning.close()
    log("autotune:", json.dumps(tune))

    batch = Batch.for_scenes(scenes, graph=True, list_capacity=args.list_capacity)
    for s in scenes:
        batch.add(s)
    batch.set_launch(tune["block_stream"], tune["radix_bits"], True, block_neighbor=tune["block_neighbor"])
    batch.step(args.warmup)
    batch.synchronize()
    particles = 0
    t1 = time.perf_counter()
    for _ in range(args.steps):
        batch.step(1)
        particles += batch.step_particles()
    elapsed = time.perf_counter() - t1
    log(f"batched: {elapsed / args.steps * 1e3:.2f} ms/step")
    lat = [batch.step_timed() for _ in range(args.latency_steps)]
    stage_ms = stage_profile(batch, args.profile_steps)
    log("stage ms:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))
    stage_particles = batch.step_particles()
    batch.close()
    value = particles / elapsed
    scene_steps = len(scenes) * args.steps / elapsed

    # the same scenes, one engine per scene, stepped one scene after another
    seq_particles = 0
    seq_elapsed = 0.0
    for k, s in enumerate(scenes):
        b = Batch.for_scenes([s], graph=True, list_capacity=args.list_capacity)
        b.add(s)
        b.set_launch(tune["block_stream"], tune["radix_bits"], True, block_neighbor=tune["block_neighbor"])
        b.step(args.warmup)
        b.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            b.step(1)
            seq_particles += b.step_particles()
        seq_elapsed += time.perf_counter() - t2
        b.close()
        if k % 64 == 63:
            log(f"sequential: {k + 1} scenes, {seq_elapsed:.1f} s")
    seq_value = seq_particles / seq_elapsed

    out = base_line(args, "batched scenes/s + p99 step latency", scene_steps, "scene-steps/s",
                    elapsed / args.steps * 1e3,
                    {"workload": workload, "scenes": len(scenes), "particles": total0,
                     "adaptive_scenes": sum(s.adaptive for s in scenes), "merged_h": args.merged_h, "graph": True,
                     "launch": {k: tune[k] for k in ("block_stream", "radix_bits", "block_neighbor")}})
    out["particle_updates_per_s"] = value
    out["step_ms_p50"] = float(np.percentile(lat, 50))
    out["step_ms_p99"] = float(np.percentile(lat, 99))
    out["sequential"] = {"particle_updates_per_s": seq_value, "scene_steps_per_s": len(scenes) * args.steps /
                         seq_elapsed, "elapsed_s": seq_elapsed, "speedup": value / seq_value}
    out["autotune"] = tune["ns_per_particle"]
    out["roofline"] = roofline_of(stage_ms, stage_particles, workload)
    out["cpu_baseline"] = None
    out["stage_ms"] = {k: round(v, 4) for k, v in stage_ms.items()}
    return out


def run_serve(args) -> dict:
    from adaptivesph_amd import serve

    budget = int(args.budget_gb * 1e9)
    reqs = serve.poisson_requests(args.rate, args.duration, seed=5)
    log(f"serving: {len(reqs)} requests at {args.rate}/s over {args.duration} s, budget {args.budget_gb} GB")
    srv = serve.Server(budget, max_scenes=256, list_capacity=args.list_capacity)
    # one small request first, so the first measured request does not pay
    # for library and kernel loading
    warm = [serve.Request(rid=-1, seed=1, steps=3, arrival=0.0)]
    srv.run(warm, lambda r: serve.serving_scene(r, 10_000, 20_000), clock=serve.StepClock())
    srv.completed.clear()
    srv.step_ms.clear()
    srv.loop_ms.clear()
    srv.resident_trace.clear()
    srv.particle_updates = 0
    stats = srv.run(reqs, lambda r: serve.serving_scene(r, merged_h=args.merged_h), progress=log)
    srv.close()
    log("serving:", json.dumps(stats))
    out = base_line(args, "served scenes/s + p99 step latency", stats["scenes_per_s"], "scenes/s",
                    stats["step_ms_p50"],
                    {"workload": "serve_poisson", "rate_per_s": args.rate, "duration_s": args.duration,
                     "budget_gb": args.budget_gb, "particles": "log-uniform 10k-500k", "merged_h": args.merged_h,
                     "steps": "uniform 100-2000", "pools": list(srv.pools)})
    out["steps"] = stats["steps"]
    out["warmup"] = 0
    out.update({k: stats[k] for k in ("particle_updates_per_s", "step_ms_p50", "step_ms_p99", "loop_ms_p50",
                                      "loop_ms_p99", "completion_p50", "completion_p99", "queue_p50",
                                      "completed", "rejected", "unfinished", "resident_mean", "resident_max",
                                      "wall_s")})
    out["cpu_baseline"] = None
    return out


def run_walls(args) -> dict:
    """Latency of updateDomain with regenerated walls (a moving tank): the
    device path (asph_update_domain_walls) against the host path
    (asph_boundary_generate + asph_update_domain), alternating between two
    domains so every call rebuilds the walls."""
    from adaptivesph_amd import abi, scene
    from adaptivesph_amd.simulation import Simulation

    s = scene.cfg2()
    sim = Simulation.from_scene(s)
    spacing = scene.boundary_spacing(s.params)
    d0 = s.params.domain
    domains = [abi.Domain(abi.Vec3(d0.min.x + 0.001 * k, d0.min.y, d0.min.z),
                          abi.Vec3(d0.max.x + 0.001 * k, d0.max.y, d0.max.z), d0.friction) for k in (1, 2)]
    for k in range(max(args.warmup, 1)):
        sim.update_domain_walls(domains[k % 2], spacing)
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.update_domain_walls(domains[k % 2], spacing)
    dev_ms = (time.perf_counter() - t0) / args.steps * 1e3
    nb = sim.get_boundary_particles_count()
    log(f"walls: {nb} wall particles, device updateDomain {dev_ms:.2f} ms")
    out = base_line(args, "updateDomain with regenerated walls (latency)", dev_ms, "ms", dev_ms,
                    {"workload": "walls_dam_break_1M", "wall_particles": nb, "fluid_particles": s.count})
    out["higher_is_better"] = False
    if not args.no_cpu_baseline:
        reps = 3
        t1 = time.perf_counter()
        for k in range(reps):
            p = abi.Parameters.from_buffer_copy(s.params)
            p.domain = domains[k % 2]
            sim.update_domain(domains[k % 2], scene.generate_boundary(p))
        host_ms = (time.perf_counter() - t1) / reps * 1e3
        out["cpu_baseline"] = {"value": host_ms, "unit": "ms", "cores": min(16, os.cpu_count() or 1), "kind": "port",
                               "sample": f"{reps} host updateDomain calls: asph_boundary_generate (psi on host "
                                         "threads) + asph_update_domain (host binning, upload)"}
    else:
        out["cpu_baseline"] = None
    sim.close()
    return out


def run_copy(args) -> dict:
    """A plain device-to-device copy of 1 GiB (torch's elementwise copy
    kernel): the known-size reference that scales hardware byte counters."""
    import torch

    n = 1 << 28  # 2^28 float32 = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    gbs = 2 * n * 4 / (ms * 1e-3) / 1e9
    out = base_line(args, "copy GB/s", gbs, "GB/s", ms, {"workload": "copy_1GiB", "bytes_read": n * 4,
                                                         "bytes_written": n * 4})
    out["cpu_baseline"] = None
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--latency-steps", type=int, default=50)
    ap.add_argument("--rate", type=float, default=2.8, help="serve: requests per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: device memory budget")
    ap.add_argument("--list-capacity", type=int, default=0,
                    help="neighbour-list entries per particle (0: the library default, 128 for plain "
                         "workloads, 640 with refinement)")
    ap.add_argument("--merged-h", choices=("kernel", "reference"), default="kernel",
                    help="merged smoothing length of adaptive scenes: kernel-normalised or the reference's")
    args = ap.parse_args()
    if args.list_capacity == 0:
        # the library's defaults: refined regions hold up to ~13x denser
        # daughters, and particles there have several hundred neighbours
        args.list_capacity = 128 if args.workload in ("cfg1", "cfg2", "cfg2_lattice", "copy", "walls") else 640
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build steps scenes on one "
              "GPU and leaves multi-GPU sharding out", file=sys.stderr)
        return 2

    import ctypes as C

    from adaptivesph_amd import abi

    lib = abi.load_library()
    buf = C.create_string_buffer(256)
    lib.asph_device_info(buf, 256)
    log("device:", buf.value.decode())

    if args.workload == "copy":
        out = run_copy(args)
    elif args.workload == "batch":
        out = run_batch(args)
    elif args.workload == "walls":
        out = run_walls(args)
    elif args.workload == "serve":
        out = run_serve(args)
    else:
        out = run_single(args)
    if args.workload != "copy":
        out["config"]["list_capacity"] = args.list_capacity
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
