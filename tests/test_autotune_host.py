"""Launch-shape autotuner on the host: a stand-in batch whose stage times
depend on the launch shape (and drift upward as the scenes evolve)."""
from adaptivesph_amd.autotune import SPACE, autotune


class FakeBatch:
    def __init__(self):
        self.shape = None
        self.steps = 0
        self.history = []

    def set_launch(self, block_stream, radix_bits, graph, block_neighbor=256):
        self.shape = (block_stream, radix_bits, block_neighbor)
        self.history.append(self.shape)

    def set_schedule(self, query_tile, force_blocks):
        self.schedule = (query_tile, force_blocks)

    def step_particles(self):
        return 1000 + 50 * self.steps  # the particle count grows

    def profile_step(self):
        bs, rb, nb = self.shape
        qt, fb = self.schedule
        self.steps += 1
        drift = 1.0 + 0.05 * self.steps  # slower as the scenes evolve
        n = self.step_particles() / 1000
        return {
            "kick_drift_key": n * drift * (1.0 + abs(bs - 512) / 1024),
            "gather": n * drift * (0.5 + 0.1 * (qt > 0)),
            "cell_sort": n * drift * (1.0 + abs(rb - 6) / 4),
            "neighbors": n * drift * (3.0 - 0.4 * (qt == 1)),
            "density": n * drift * (2.0 + abs(nb - 128) / 128),
            "force": n * drift * (2.0 + abs(fb - 8) / 16),
        }


def test_autotune_picks_each_parameter_on_its_own_stages():
    b = FakeBatch()
    got = autotune(b, rounds=3)
    assert (got["block_stream"], got["radix_bits"], got["block_neighbor"]) == (512, 6, 128)
    assert (got["query_tile"], got["force_blocks"]) == (1, 8)
    assert b.shape == (512, 6, 128) and b.schedule == (1, 8)
    n_cand = sum(len(v) for v, _ in SPACE.values())
    assert got["steps_used"] == 3 * n_cand
    assert set(got["ns_per_particle"]) == set(SPACE)


def test_autotune_is_deterministic_and_keeps_earlier_on_ties():
    class Flat(FakeBatch):
        def profile_step(self):
            self.steps += 1
            return {k: 1.0 for k in ("kick_drift_key", "gather", "cell_sort", "neighbors", "density", "force")}

        def step_particles(self):
            return 1000

    a = autotune(Flat())
    b = autotune(Flat())
    assert (a["block_stream"], a["radix_bits"], a["block_neighbor"]) == (128, 0, 64)
    assert (a["query_tile"], a["force_blocks"]) == (0, 0)
    assert a == b
