"""The serving loop's scheduling on the CPU, with a stand-in runner whose
blocks complete after a few polls: admission in arrival order within the
memory budget and slot count, every block recorded, jobs that can never fit
refused, and the request stream reproducible from its seed."""
import numpy as np
import pytest

from alea_amd import serving


class FakeRunner:
    log = []

    def __init__(self, job):
        self.job = job
        self.left = 0
        FakeRunner.log.append(("admit", job.job_id))

    def start_block(self):
        self.left = 2 + self.job.n_particles // 60

    def done(self):
        self.left -= 1
        return self.left <= 0

    def block_energy(self):
        return -0.1 - 1e-3 * len(self.job.block_energy)

    def close(self):
        FakeRunner.log.append(("close", self.job.job_id))


def _jobs(n=9):
    rng = np.random.default_rng(0)
    return [serving.Job(i, 0.002 * i, int(rng.choice([26, 122, 242])), 30.0, (0.1, 0.2),
                        int(rng.choice([512, 4096, 8192])), n_blocks=3, sweeps_per_block=5)
            for i in range(n)]


def test_admission_respects_budget_and_order():
    jobs = _jobs()
    sizes = [serving.footprint_bytes(j.n_particles, j.n_walkers, j.precision) for j in jobs]
    budget = 2 * max(sizes)
    FakeRunner.log = []
    srv = serving.Server(hbm_budget_bytes=budget, max_active=3, runner=FakeRunner, poll_interval=1e-5)
    st = srv.run(jobs)
    assert st["jobs"] == len(jobs) and srv.peak_bytes <= budget and srv.peak_active <= 3
    admits = [i for kind, i in FakeRunner.log if kind == "admit"]
    assert admits == sorted(admits)                       # first come, first admitted
    assert all(len(j.block_energy) == j.n_blocks == len(j.block_latency) for j in jobs)
    assert all(j.admitted >= j.arrival and j.finished >= j.admitted for j in jobs)
    for k in ("p50_block_ms", "p99_block_ms", "p50_request_s", "p99_request_s", "walker_moves_per_s"):
        assert np.isfinite(st[k])
    assert st["walker_moves"] == sum(j.walker_moves for j in jobs)


def test_job_larger_than_budget_is_refused():
    j = serving.Job(0, 0.0, 242, 30.0, (0.1, 0.2), 8192)
    with pytest.raises(ValueError):
        serving.Server(hbm_budget_bytes=1 << 20, runner=FakeRunner).run([j])


def test_request_stream_is_reproducible():
    a = serving.poisson_jobs(12, rate=5.0, seed=4)
    b = serving.poisson_jobs(12, rate=5.0, seed=4)
    assert [(j.arrival, j.n_particles, j.rs, j.twist, j.n_walkers) for j in a] == \
           [(j.arrival, j.n_particles, j.rs, j.twist, j.n_walkers) for j in b]
    assert all(j.n_particles in (26, 58, 74, 122, 242) and 512 <= j.n_walkers <= 8192 for j in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
