"""Launch-shape autotuner of a batch.

Five launch parameters shape a step, each acting on its own stages:

* block_stream   threads per block of the streaming kernels (kick/drift/key,
                 gather);
* radix_bits     digit width of the cell-key radix sort (cell sort; 0 = the
                 width with the fewest passes);
* block_neighbor threads per block of the neighbour kernels (density, forces);
* query_tile     the cell-tile shape of the neighbour kernels' waves: the
                 sorted (cell) order, 2x2x2 / 4x4x4 sub-cell tiles in Morton
                 order, 4x4x4 sub-cells in x-major order (slabs) within each
                 cell, or slabs per scene once its waves loosen (the
                 reordering runs in the gather stage);
* force_blocks   the force kernels' grid: one thread per particle, or a
                 persistent grid of 4, 8 or 16 blocks per compute unit.

The tuner measures, in a fixed order, each candidate value on the stages it
acts on (HIP events between the stages of eager steps of the batch's own
scenes), normalised by the particles of that step, and keeps the fastest
(coordinate search: one parameter at a time, the others at their current
choice). The scenes evolve while the tuner steps them (adaptive scenes change
their particle counts), so every candidate is measured in `rounds` interleaved
rounds and judged by its median. Every launch shape gives bit-identical
results, so the tuning steps are ordinary steps of the simulation: they
advance the scenes, and count as warm-up where a benchmark follows.
"""
from __future__ import annotations

import numpy as np

SPACE = {
    "block_stream": ((128, 256, 512, 1024), ("kick_drift_key", "gather")),
    "radix_bits": ((0, 4, 6, 7, 8, 10), ("cell_sort",)),
    "block_neighbor": ((64, 128, 256), ("neighbors", "density", "force")),
    "query_tile": ((0, 1, 2, 3, 4), ("gather", "neighbors", "density", "force")),
    "force_blocks": ((0, 4, 8, 16), ("force",)),
}
# query_tile -1: the engine's own default (the automatic query order)
DEFAULT = {"block_stream": 256, "radix_bits": 0, "block_neighbor": 256, "query_tile": -1, "force_blocks": 0}


def _apply(batch, shape: dict, graph: bool) -> None:
    batch.set_launch(shape["block_stream"], shape["radix_bits"], graph, block_neighbor=shape["block_neighbor"])
    batch.set_schedule(shape["query_tile"], shape["force_blocks"])


def autotune(batch, rounds: int = 3, graph: bool = True, space: dict | None = None) -> dict:
    """Chooses and sets the launch shape; returns it with the measured
    per-particle stage times (ns per particle) of every candidate."""
    space = space or SPACE
    shape = dict(DEFAULT)
    table = {}
    steps = 0
    for param, (values, stages) in space.items():
        samples = {v: [] for v in values}
        for _ in range(rounds):
            for v in values:
                trial = dict(shape, **{param: v})
                _apply(batch, trial, graph)
                ms = batch.profile_step()
                n = max(batch.step_particles(), 1)
                steps += 1
                samples[v].append(sum(ms.get(s, 0.0) for s in stages) * 1e6 / n)
        med = {v: float(np.median(samples[v])) for v in values}
        best = min(values, key=lambda v: (med[v], values.index(v)))
        shape[param] = best
        table[param] = {str(v): round(med[v], 4) for v in values}
    _apply(batch, shape, graph)
    return {**shape, "ns_per_particle": table, "steps_used": steps}
