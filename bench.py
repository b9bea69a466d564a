#!/usr/bin/env python3
"""Benchmark of the per-step SPH pipeline on one MI355X.

Default workload: the 1,048,576-particle WCSPH dam break (128 x 64 x 128
lattice, dx = 0.0039 m, 2.0 x 1.2 x 1.0 m box, FP32, no refinement), stepped
as one scene through the batched engine with the step captured as a hipGraph.
A step is one full pass of the pipeline (kick, drift, cell keys, sort, cell
ranges, gather, density, forces, kick) over all particles. The inputs are
resident in device memory before the timed region. Prints one JSON line whose
value is the 1M dam break's particle-updates/s; the same line carries the
second half of the headline metric under "batched" (256 scenes stepped
together: scene-steps/s, p50 / p99 step latency, speed-up over the same scenes
stepped one after another), per-kernel figures of the neighbour stages
("kernels": algorithmic bytes and flops against the HBM and FP32 peaks, and
the stored counter figures) and the CPU baseline (the oracle, median of 3 runs
of 50 steps on the host's cores).

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
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak
HEADLINE_METRIC = "particle-updates/s at 1M WCSPH; batched scenes/s + p99 step latency"
# Algorithmic bytes per particle and launch (minimum compulsory traffic):
#   density: reads position 16 B, h 4 B, mass 4 B; writes density 4 B
#   force (+ fused second half kick): reads position, velocity (32 B),
#     density, h, mass (12 B); writes acceleration 16 B and velocity 16 B
ALGO_BYTES = {"density": 28, "force": 76}
# Algorithmic FP32 operations per neighbour pair (add, mul, div, sqrt and
# compare count one each; per-particle factors such as p/rho^2, 2 alpha h c
# and the kernel coefficients counted once per particle, not per pair):
#   density (Algorithm.cu:17-54), per hit including self: offset 3, d^2 5,
#     range test 1, sqrt 1, W(d, h_j) 9, rho += m W 2 -> 21
#   force (Algorithm.cu:56-174), per hit j != i: pressure term 29 (offset 3,
#     d^2 5, test 1, sqrt 1, direction 3, W'(d, h_j) 7, p term 1, update 8)
#     + artificial viscosity 23 (velocity difference 3, compression 5, test 1,
#     nu 2, Pi 4, m Pi W' 2, update 6) -> 52
FLOPS_PER_PAIR = {"density": 21, "force": 52}
COUNTERS = ROOT / "profiles" / "counters.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str, merged_h: str = "kernel", math: str = "exact"):
    scenes, workload = _make_workload(name, merged_h)
    if math == "tolerance":
        from adaptivesph_amd import abi

        for s in scenes:
            s.params.flags |= abi.FLAG_TOLERANCE_MATH
    return scenes, workload


def _make_workload(name: str, merged_h: str):
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


def host_facts() -> dict:
    """CPU model, logical CPUs of the machine and CPUs this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(scene_obj, steps: int, runs: int) -> dict:
    """The oracle (FP32 host restatement of the reference's kernels, OpenMP
    over particles) on a bounded sample of the workload: one warm-up step, then
    `runs` consecutive runs of `steps` steps each; the median run is reported."""
    # OpenMP reads OMP_NUM_THREADS when the oracle library loads; default to
    # every CPU this process may use
    os.environ.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    from oracle import oracle

    threads = oracle.threads()
    o = oracle.OracleSim.from_scene(scene_obj)
    o.update(scene_obj.dt)  # warm-up step
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(steps):
            o.update(scene_obj.dt)
        times.append(time.perf_counter() - t0)
    o.close()
    med = sorted(times)[len(times) // 2]
    facts = host_facts()
    return {
        "value": scene_obj.count * steps / med,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{scene_obj.name}: {scene_obj.count} particles, median of {runs} runs of {steps} steps "
                  f"(after 1 warm-up step; runs {', '.join(f'{t:.1f}' for t in times)} s) on {threads} OpenMP "
                  f"threads of {facts['nproc']} CPUs ({facts['affinity_cpus']} usable), {facts['cpu_model']}",
        **facts,
        "run_s": [round(t, 3) for t in times],
    }


def timed_steps(batch, steps: int, samples: int, count_particles: bool) -> tuple[float, int, dict]:
    """The timed region: `steps` steps, of which `samples` (spread evenly over
    the region) run eagerly with HIP events recorded on the engine's stream
    between their stages, the others as graph replays; nothing waits for the
    device inside the region (the events are read after it). Returns the
    elapsed wall time (device synchronised on both sides), the particles
    updated (each step's count when count_particles, else 0) and each stage's
    device time averaged over the sampled steps. The per-stage times therefore
    cover the same steps as the throughput, as a step's cost drifts with the
    flow."""
    samples = max(1, min(samples, steps))
    every = steps // samples  # step k is sampled when k % every == every - 1
    taken = 0
    particles = 0
    batch.synchronize()
    t0 = time.perf_counter()
    k = 0
    while k < steps:
        if k % every == every - 1 and taken < samples:
            batch.profile_step_async()
            taken += 1
            run = 1
        else:
            # graph replays up to the next sampled step
            run = 1 if count_particles else max(1, min(every - 1 - k % every, steps - k))
            batch.step(run)
        if count_particles:
            particles += batch.step_particles()
        k += run
    batch.synchronize()
    elapsed = time.perf_counter() - t0
    stages, _ = batch.profile_collect()
    return elapsed, particles, stages


def counters_key(workload: str, math: str) -> str:
    """Key of a workload's entry in profiles/counters.json."""
    return workload if math == "exact" else f"{workload}/{math}"


def stored_counters(workload: str) -> dict:
    """Hardware-counter figures of the workload's neighbour stages
    (rocprofv3 --pmc runs summarised in profiles/counters.json)."""
    try:
        data = json.loads(COUNTERS.read_text())
    except (OSError, ValueError):
        return {}
    out = data.get(workload, {})
    return {k: dict(v, source=f"profiles/counters.json ({data.get('_build', '?')})") for k, v in out.items()}


def stage_figures(stage_ms: dict, particles: int, pairs: float | None, workload: str) -> dict:
    """Per neighbour stage: time, algorithmic bytes and flops against the HBM
    and FP32 peaks, and the stored counter figures (measured DRAM traffic,
    VALU instructions and issue share)."""
    ms = {"density": stage_ms.get("neighbors", 0.0) + stage_ms.get("density", 0.0),
          "force": stage_ms.get("force", 0.0)}
    counters = stored_counters(workload)
    out = {}
    for k, t in ms.items():
        if t <= 0.0:
            continue
        sec = t * 1e-3
        f = {"launch_ms": round(t, 4), "algorithmic_bytes": ALGO_BYTES[k] * particles,
             "algorithmic_gbs": round(ALGO_BYTES[k] * particles / sec / 1e9, 2)}
        f["hbm_frac"] = round(f["algorithmic_gbs"] / HBM_PEAK_GBS, 5)
        if pairs is not None:
            n_pairs = pairs + (particles if k == "density" else 0)  # the density sums the self term too
            flops = FLOPS_PER_PAIR[k] * n_pairs
            f["pairs"] = int(n_pairs)
            f["flops"] = int(flops)
            f["tflops"] = round(flops / sec / 1e12, 3)
            f["fp32_frac"] = round(f["tflops"] / FP32_PEAK_TFLOPS, 4)
        c = counters.get(k)
        if c:
            f["counters"] = c
        out[k] = f
    return out


def binding_resource(stage: str, valu_frac: float | None) -> str:
    """What limits the stage's main kernel, read from its counters."""
    if (valu_frac or 0.0) < 0.8:
        return "vector memory path (per-lane 16-byte gathers through L1; see profiles/valu_*.md)"
    if stage == "density":
        return ("instruction issue: the wave filter's VALU, LDS, scalar and texture-path work together "
                "(VALU issue share near 1; see profiles/valu_*.md)")
    return "VALU issue (instruction-bound pair loop; see valu_issue_frac)"


def roofline_of(stages: dict) -> dict:
    """The bench line's roofline object for the dominant kernel: its
    algorithmic bytes against HBM (the contract's figure), beside the binding
    resource its counters show (VALU issue) and its algorithmic FP32 rate."""
    dom = max(stages, key=lambda k: stages[k]["launch_ms"])
    f = stages[dom]
    c = f.get("counters", {})
    return {
        "bound": "hbm",
        "kernel": {"density": "kNeighbors (neighbour lists + density)",
                   "force": "kForce (kForceTol in the tolerance mode)"}[dom],
        "achieved": f["algorithmic_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": f["hbm_frac"],
        "traffic": c.get("dram_bytes_per_launch"),
        "algorithmic_bytes_per_launch": f["algorithmic_bytes"],
        "launch_ms": f["launch_ms"],
        "binding_resource": binding_resource(dom, c.get("valu_issue_frac")),
        "valu_issue_frac": c.get("valu_issue_frac"),
        "fp32_tflops": f.get("tflops"),
        "fp32_peak_tflops": FP32_PEAK_TFLOPS,
        "fp32_frac": f.get("fp32_frac"),
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

    scenes, workload = make_workload(args.workload, args.merged_h, args.math)
    s = scenes[0]
    batch = Batch.for_scenes(scenes, graph=not args.no_graph, list_capacity=args.list_capacity)
    sid = batch.add(s)
    log(f"workload {workload}: {s.count} particles")
    batch.step(args.warmup)
    batch.synchronize()
    # the particle count of an adaptive scene changes every step: count each
    # step's particles
    elapsed, particles, stage_ms = timed_steps(batch, args.steps, args.profile_steps, s.adaptive)
    if not s.adaptive:
        particles = s.count * args.steps
    log("stage ms:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))
    stage_particles = batch.step_particles()
    pairs = batch.average_neighbor_count(sid) * batch.count(sid)
    count = batch.count(sid)
    batch.close()
    value = particles / elapsed
    headline = workload == "dam_break_1M"
    metric = HEADLINE_METRIC if headline else f"particle-updates/s ({workload})"
    out = base_line(args, metric, value, "particle-updates/s", elapsed / args.steps * 1e3,
                    {"workload": workload, "scenes": 1, "particles": s.count, "final_particles": count,
                     "graph": not args.no_graph, "dt": s.dt, "math": args.math,
                     **({"merged_h": args.merged_h} if s.adaptive else {})})
    stages = stage_figures(stage_ms, stage_particles, pairs, counters_key(workload, args.math))
    out["roofline"] = roofline_of(stages)
    out["kernels"] = stages
    out["stage_ms"] = {k: round(v, 4) for k, v in stage_ms.items()}
    if headline and args.math == "exact" and not args.no_tolerance_leg:
        # the same workload in the tolerance mode (hardware reciprocals in the
        # density and force sums; within 1e-5 of the FP32 oracle), beside the
        # bit-exact value
        out["tolerance_mode"] = measure_tolerance(args)
    if headline and not args.no_batch:
        # the second half of the headline metric: 256 scenes stepped together
        out["batched"] = measure_batch(args, with_sequential=True)
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(s, args.cpu_steps, args.cpu_runs)
    return out


def measure_tolerance(args) -> dict:
    """The default workload stepped the same way in the tolerance mode."""
    from adaptivesph_amd.batch import Batch

    scenes, workload = make_workload(args.workload, args.merged_h, "tolerance")
    s = scenes[0]
    batch = Batch.for_scenes(scenes, graph=not args.no_graph, list_capacity=args.list_capacity)
    batch.add(s)
    batch.step(args.warmup)
    batch.synchronize()
    elapsed, particles, stage_ms = timed_steps(batch, args.steps, args.profile_steps, s.adaptive)
    batch.close()
    particles = s.count * args.steps
    log(f"tolerance mode: {elapsed / args.steps * 1e3:.4f} ms/step")
    return {"math": "tolerance", "value": particles / elapsed, "unit": "particle-updates/s",
            "ms_per_step": elapsed / args.steps * 1e3, "steps": args.steps, "workload": workload,
            "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
            "accuracy": "density, pressure and acceleration within 1e-5 of the FP32 oracle; keys, order, "
                        "neighbour sets and refinement decisions bit-exact (tests/test_gpu_tolerance.py)"}


def measure_batch(args, with_sequential: bool) -> dict:
    """256 scenes stepped together (graph-captured step, launch shape chosen
    by the autotuner on a separate copy of the batch), timed over args.steps
    steps after args.warmup; then, if asked, the same scenes one after another
    (each in its own engine, graph-captured, synchronous steps as in the
    reference's update()), from the same initial state over the same steps."""
    import numpy as np

    from adaptivesph_amd.autotune import autotune
    from adaptivesph_amd.batch import Batch

    t0 = time.perf_counter()
    scenes, workload = make_workload("batch", args.merged_h, args.math)
    total0 = sum(s.count for s in scenes)
    log(f"workload {workload}: {len(scenes)} scenes, {total0} particles, "
        f"{sum(s.adaptive for s in scenes)} adaptive ({time.perf_counter() - t0:.1f} s to build)")
    lc = args.batch_list_capacity
    tuning = Batch.for_scenes(scenes, graph=True, list_capacity=lc)
    for s in scenes:
        tuning.add(s)
    tune = autotune(tuning)
    tuning.close()
    log("autotune:", json.dumps(tune))

    batch = Batch.for_scenes(scenes, graph=True, list_capacity=lc)
    sids = [batch.add(s) for s in scenes]
    batch.set_launch(tune["block_stream"], tune["radix_bits"], True, block_neighbor=tune["block_neighbor"])
    batch.set_schedule(tune["query_tile"], tune["force_blocks"])
    batch.step(args.warmup)
    batch.synchronize()
    elapsed, particles, stage_ms = timed_steps(batch, args.steps, args.profile_steps, True)
    log(f"batched: {elapsed / args.steps * 1e3:.2f} ms/step")
    lat = [batch.step_timed() for _ in range(args.latency_steps)]
    log("stage ms:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))
    stage_particles = batch.step_particles()
    pairs = sum(batch.average_neighbor_count(i) * batch.count(i) for i in sids)
    batch.close()
    out = {
        "metric": "batched scenes/s + p99 step latency",
        "value": len(scenes) * args.steps / elapsed,
        "unit": "scene-steps/s",
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "particle_updates_per_s": particles / elapsed,
        "step_ms_p50": float(np.percentile(lat, 50)),
        "step_ms_p99": float(np.percentile(lat, 99)),
        "latency_steps": len(lat),
        "config": {"workload": workload, "scenes": len(scenes), "particles": total0,
                   "adaptive_scenes": sum(s.adaptive for s in scenes), "merged_h": args.merged_h, "graph": True,
                   "math": args.math,
                   "list_capacity": lc,
                   "launch": {k: tune[k] for k in ("block_stream", "radix_bits", "block_neighbor", "query_tile",
                                                   "force_blocks")}},
        "autotune": tune["ns_per_particle"],
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "kernels": stage_figures(stage_ms, stage_particles, pairs, counters_key(workload, args.math)),
    }
    if with_sequential:
        seq_particles = 0
        seq_elapsed = 0.0
        for k, s in enumerate(scenes):
            b = Batch.for_scenes([s], graph=True, list_capacity=lc)
            b.add(s)
            b.set_launch(tune["block_stream"], tune["radix_bits"], True, block_neighbor=tune["block_neighbor"])
            b.set_schedule(tune["query_tile"], tune["force_blocks"])
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
        out["sequential"] = {"particle_updates_per_s": seq_particles / seq_elapsed,
                             "scene_steps_per_s": len(scenes) * args.steps / seq_elapsed,
                             "elapsed_s": seq_elapsed,
                             "speedup": (particles / elapsed) / (seq_particles / seq_elapsed)}
    return out


def run_batch(args) -> dict:
    m = measure_batch(args, with_sequential=not args.no_sequential)
    out = base_line(args, m["metric"], m["value"], m["unit"], m["ms_per_step"], m["config"])
    for k in ("particle_updates_per_s", "step_ms_p50", "step_ms_p99", "sequential", "autotune", "kernels",
              "stage_ms"):
        out[k] = m.get(k)
    out["roofline"] = roofline_of(m["kernels"])
    out["cpu_baseline"] = None
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
    ap.add_argument("--cpu-steps", type=int, default=50, help="CPU baseline: steps per run")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline: runs (the median is reported)")
    ap.add_argument("--no-batch", action="store_true", help="default workload: leave out the batched leg")
    ap.add_argument("--no-tolerance-leg", action="store_true",
                    help="default workload: leave out the tolerance-mode measurement beside the exact one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch workload: leave out the per-scene sequential comparison")
    ap.add_argument("--profile-steps", type=int, default=5,
                    help="steps of the timed region (spread evenly) whose stages are timed with HIP events")
    ap.add_argument("--latency-steps", type=int, default=50)
    ap.add_argument("--rate", type=float, default=2.8, help="serve: requests per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: device memory budget")
    ap.add_argument("--list-capacity", type=int, default=0,
                    help="neighbour-list entries per particle (0: the library default, 128 for plain "
                         "workloads, 640 with refinement)")
    ap.add_argument("--math", choices=("exact", "tolerance"), default="exact",
                    help="density and force sums: bit-exact with the reference's FP32 arithmetic, or the tolerance "
                         "mode (hardware reciprocal square root and reciprocal; within 1e-5)")
    ap.add_argument("--merged-h", choices=("kernel", "reference"), default="kernel",
                    help="merged smoothing length of adaptive scenes: kernel-normalised or the reference's")
    args = ap.parse_args()
    args.batch_list_capacity = args.list_capacity or 640
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
