"""Launch-shape autotuner of a batch.

Searches, in a fixed order, the block size of the streaming kernels and the
digit width of the cell-key radix sort, timing each candidate on the batch's
own scenes with HIP events, and keeps the fastest. Every candidate computes
bit-identical results (the launch shape changes only how work is split), so
the tuning steps are ordinary steps of the simulation: they advance the
scenes, and count as warm-up where a benchmark follows.
"""
from __future__ import annotations

import numpy as np

BLOCK_STREAM = (128, 256, 512, 1024)
RADIX_BITS = (4, 6, 7, 8)


def candidates(block_stream=BLOCK_STREAM, radix_bits=RADIX_BITS):
    return [(b, r) for r in radix_bits for b in block_stream]


def autotune(batch, steps: int = 4, graph: bool = True, cands=None) -> dict:
    """Times `steps` steps per candidate (after one untimed step that captures
    the candidate's graph) and sets the candidate with the lowest median step
    time. Ties keep the earlier candidate. Returns the choice and the table."""
    table = []
    best = None
    for bs, rb in cands or candidates():
        batch.set_launch(bs, rb, graph)
        batch.step(1)
        batch.synchronize()
        ms = float(np.median([batch.step_timed() for _ in range(steps)]))
        table.append({"block_stream": bs, "radix_bits": rb, "ms": round(ms, 4)})
        if best is None or ms < best[2]:
            best = (bs, rb, ms)
    batch.set_launch(best[0], best[1], graph)
    return {"block_stream": best[0], "radix_bits": best[1], "ms": round(best[2], 4), "table": table,
            "steps_used": len(table) * (steps + 1)}
