#!/usr/bin/env python3
"""Benchmark of the per-step SPH pipeline on one MI355X.

Default workload: the 1,048,576-particle WCSPH dam break (128 x 64 x 128
lattice, dx = 0.0039 m, 2.0 x 1.2 x 1.0 m box, FP32, no refinement), stepped
as one scene through the batched engine with the step captured as a hipGraph.
A step is one full pass of the pipeline (kick, drift, cell keys, sort, cell
ranges, gather, density, forces, kick) over all particles. The inputs are
resident in device memory before the timed region. Prints one JSON line.

Other workloads (--workload): cfg1 (31,250-particle dam break), batch (256
scenes of 5k-200k particles, half adaptive, one launch per stage), adaptive
(500k-particle sloshing tank refining every step).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
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


def make_workload(name: str):
    from adaptivesph_amd import scene

    if name == "cfg2":
        return [scene.cfg2()], "dam_break_1M"
    if name == "cfg1":
        return [scene.cfg1()], "dam_break_31k"
    if name == "adaptive":
        return [scene.cfg3()], "adaptive_slosh_500k"
    if name == "batch":
        return scene.cfg4_scenes(256), "batch_256_scenes"
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build steps scenes on one "
              "GPU and leaves multi-GPU sharding out", file=sys.stderr)
        return 2

    from adaptivesph_amd import abi
    from adaptivesph_amd.batch import Batch

    lib = abi.load_library()
    import ctypes as C

    buf = C.create_string_buffer(256)
    lib.asph_device_info(buf, 256)
    log("device:", buf.value.decode())

    scenes, workload = make_workload(args.workload)
    total = sum(s.count for s in scenes)
    batch = Batch.for_scenes(scenes, graph=not args.no_graph)
    ids = [batch.add(s) for s in scenes]
    log(f"workload {workload}: {len(scenes)} scene(s), {total} particles")

    batch.step(args.warmup)
    batch.synchronize()
    # live per-stage timing with HIP events on the engine's stream
    stages = {}
    for _ in range(max(args.profile_steps, 1)):
        for k, v in batch.profile_step().items():
            stages.setdefault(k, []).append(v)
    stage_ms = {k: sum(v) / len(v) for k, v in stages.items()}
    log("stage ms:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))

    batch.synchronize()
    t0 = time.perf_counter()
    batch.step(args.steps)
    batch.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    counts = [batch.count(i) for i in ids]
    value = sum(counts) * args.steps / elapsed

    dom = max(("density", "force"), key=lambda k: stage_ms.get(k, 0.0))
    dur = stage_ms[dom] * 1e-3
    achieved = ALGO_BYTES[dom] * sum(counts) / dur / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            t = json.loads(tfile.read_text()).get(workload, {}).get(dom)
            traffic = t.get("bytes_per_launch") if t else None
        except (ValueError, AttributeError):
            traffic = None
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": ALGO_BYTES[dom] * sum(counts),
        "launch_ms": round(stage_ms[dom], 4),
    }
    cpu = None
    if not args.no_cpu_baseline and len(scenes) == 1:
        cpu = cpu_baseline(scenes[0], args.cpu_steps)
    out = {
        "metric": "particle-updates/s at 1M WCSPH" if workload == "dam_break_1M" else f"particle-updates/s ({workload})",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "scenes": len(scenes),
            "particles": sum(counts),
            "graph": not args.no_graph,
            "dt": scenes[0].dt if len(scenes) == 1 else "per scene",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
    }
    print(json.dumps(out), flush=True)
    batch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
