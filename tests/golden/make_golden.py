"""Writes the golden fixtures of the parity tests from the oracle (the FP32
host restatement of the reference pipeline). The reference ships no tests or
fixtures of its own and cannot be built here, so these vectors pin the
oracle's own outputs: CPU tests regenerate them with the oracle, GPU tests
reproduce them with the HIP library, both bit for bit.

Inputs are regenerated from the seeded scene generators; each fixture stores
outputs only (plus a digest of the inputs). Run from the repository root:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from golden_cases import CASES, run_oracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def digest(arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main() -> None:
    for name in CASES:
        res, inputs = run_oracle(name)
        res["input_digest"] = np.array(digest(inputs))
        np.savez_compressed(OUT / f"{name}.npz", **res)
        print(name, {k: v.shape for k, v in res.items()})


if __name__ == "__main__":
    main()
