"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must reproduce them bit for bit on the CPU, and the HIP library
must reproduce them bit for bit on the GPU."""
from pathlib import Path

import numpy as np
import pytest

from golden_cases import CASES, inputs_of, run, run_oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def _check(out, arrs, name):
    ref = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    from golden_cases import digest

    assert str(ref["input_digest"]) == digest(arrs), "inputs of the case changed"
    for k in ref.files:
        if k == "input_digest":
            continue
        a = np.ascontiguousarray(out[k])
        b = np.ascontiguousarray(ref[k])
        assert a.shape == b.shape, k
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{name}: {k}"


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name):
    out, arrs = run_oracle(name)
    _check(out, arrs, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_hip_reproduces_golden(name):
    from adaptivesph_amd.simulation import Simulation

    out, arrs = run(lambda s: Simulation.from_scene(s), name)
    _check(out, arrs, name)
    assert inputs_of(name)[0].count > 0
