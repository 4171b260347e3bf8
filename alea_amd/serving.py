"""Continuous-batching serving loop for VMC jobs on one GPU.

A request is a VMC job: on the 2-D paramagnetic electron gas (N electrons at
density rs with a twist; Slater-Jastrow-Ewald), or on the 2-D bosonic dipole
gas (N dipoles at density n r0^2; McMillan Jastrow, 1/r^3); n_walkers
walkers, n_blocks blocks of sweeps_per_block sweeps.  Requests arrive over time; the server admits them,
in arrival order, as soon as their device footprint fits the HBM budget and an
active slot is free; admitted jobs live on the GPU as their own walker batch
(their own alea_system, whose work runs on its own non-blocking stream), so
the blocks of all active jobs are in flight on the device together and a new
job starts while others are mid-run.  Each block is queued asynchronously
(alea_vmc_sweeps with no acceptance read-back, replayed as one captured
hipGraph per job) and polled (alea_query); when
it completes, its local energy is measured and the job's next block is
queued.  A finished job frees its memory and slot at once.

Per job it records the energy per particle of every block, the block
latencies (queued -> completed) and the request latency (arrival ->
finished); over the stream, the p50/p99 of both and the aggregate
walker-moves/s.  Jobs are independent (own walkers, own device streams of
random numbers from the job's seed), so a job's results are the same alone
and in any mix.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .system import dipole_spec, heg_spec, square_lattice, twisted_wavevectors, uniform_positions
from .walkers import FP64, MIXED, gpu_walkers, seed as seed_walkers, set_sweep_graphs, vmc_sweeps_fused

MAX_FUSED = 8    # systems in one fused launch (kMaxFused of the C-ABI)


@dataclass
class Job:
    job_id: int
    arrival: float                 # seconds after the stream starts
    n_particles: int               # paramagnetic: N/2 up, N/2 down
    rs: float
    twist: tuple
    n_walkers: int
    n_blocks: int = 20
    sweeps_per_block: int = 50
    seed: int = 0
    precision: int = MIXED
    kind: str = "heg"              # "heg" or "dipoles" (rs is then 1/sqrt(pi n r0^2))
    # filled in by the server
    admitted: float | None = None
    finished: float | None = None
    block_energy: list = field(default_factory=list)     # E_L per particle, mean over walkers
    block_latency: list = field(default_factory=list)    # seconds, queued -> completed
    accepted: int = 0

    @property
    def tau(self) -> float:
        return (0.5 * self.rs) ** 2          # Metropolis step sigma = 0.5 rs

    @property
    def walker_moves(self) -> int:
        return self.n_walkers * self.n_particles * self.n_blocks * self.sweeps_per_block


def footprint_bytes(n_particles: int, n_walkers: int, precision: int, n_kj: int = 40,
                    complex_orbitals: bool = True, slater: bool = True) -> int:
    """Device bytes of one job's walker batch, as the C-ABI allocates them
    (alea_system::build): positions, k-space densities, log parts, the two
    determinant inverses (rows padded to 16) in the determinant precision,
    condition estimates, rebuild workspace (FP64), RNG states and the
    per-walker scalars."""
    nb = n_particles // 2 if slater else 0
    rows = (nb + 15) // 16 * 16
    e = 2 if complex_orbitals else 1
    t = 4 if precision == MIXED else 8
    inv = 2 * e * rows * nb
    inv += inv & 1
    per_walker = (16 * n_particles + 16 * 2 * n_kj + 32 + inv * t + 16 + 8
                  + 8 * max(1, 2 * e * nb * nb) + 32 + 16 + 16 + 24 + 8 + 4 + 8 + 4)
    return per_walker * n_walkers + (1 << 20)


def job_footprint(j) -> int:
    return footprint_bytes(j.n_particles, j.n_walkers, j.precision, slater=j.kind == "heg")


def poisson_jobs(n_jobs: int, rate: float, seed: int = 4, sizes=(26, 58, 74, 122, 242),
                 rs_choices=(30.0, 40.0, 60.0), walkers=(512, 8192), n_blocks: int = 20,
                 sweeps_per_block: int = 50, precision: int = MIXED, dipole_fraction: float = 0.0,
                 dipoles=(256, 1.0)) -> list:
    """A stream of requests: Poisson arrivals at `rate` per second; N, rs and
    the twist (uniform in the first Brillouin zone, redrawn while it splits a
    degenerate |k + theta| shell at the Fermi level) and a log-uniform walker
    count rounded to 64, all from `seed`.  A `dipole_fraction` of the jobs are
    dipole-gas jobs of dipoles = (N, n r0^2)."""
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    for i in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        if dipole_fraction > 0.0 and rng.random() < dipole_fraction:
            lo, hi = math.log2(walkers[0]), math.log2(walkers[1])
            nw = int(round(2.0 ** rng.uniform(lo, hi) / 64.0)) * 64
            nd, dens = dipoles
            jobs.append(Job(i, t, nd, 1.0 / math.sqrt(math.pi * dens), (0.0, 0.0), max(64, nw), n_blocks,
                            sweeps_per_block, seed=int(rng.integers(1 << 62)), precision=precision,
                            kind="dipoles"))
            continue
        n = int(rng.choice(sizes))
        rs = float(rng.choice(rs_choices))
        while True:
            twist = tuple(float(x) for x in rng.uniform(-0.5, 0.5, 2))
            try:   # a twist that splits a degenerate |k + theta| shell is drawn again
                twisted_wavevectors(square_lattice(n, rs), n // 2, twist)
                break
            except ValueError:
                continue
        # the cell's breakup tables, built once here rather than at admission
        heg_spec(n // 2, n // 2, rs, twist=twist)
        lo, hi = math.log2(walkers[0]), math.log2(walkers[1])
        nw = int(round(2.0 ** rng.uniform(lo, hi) / 64.0)) * 64
        # FP32 inverses only for blocks of up to 64 orbitals: the larger
        # low-density blocks reach condition numbers (> 1e7) at which an FP32
        # inverse has no correct digits, so those jobs run in FP64
        prec = precision if n // 2 <= 64 else FP64
        jobs.append(Job(i, t, n, rs, twist, max(64, nw), n_blocks, sweeps_per_block,
                        seed=int(rng.integers(1 << 62)), precision=prec))
    return jobs


class GpuRunner:
    """One admitted job on the GPU: its walker batch and the block in flight."""

    def __init__(self, job: Job, graphs: bool = True):
        if job.kind == "dipoles":
            spec = dipole_spec(job.n_particles, density=1.0 / (math.pi * job.rs * job.rs))
        else:
            spec = heg_spec(job.n_particles // 2, job.n_particles // 2, job.rs, twist=job.twist)
        self.job = job
        self.complex_orbitals = bool(spec.slater_twisted)
        self.w = gpu_walkers(spec, job.n_walkers, job.precision)
        self.w.set_positions(uniform_positions(spec, job.n_walkers, np.random.default_rng(job.seed)))
        self.w.evaluate()
        seed_walkers(self.w, job.seed)
        # every block has the same launches: one captured hipGraph per job
        set_sweep_graphs(self.w, graphs)

    def start_block(self):
        self.w.api("vmc_sweeps", self.w.h, self.job.sweeps_per_block, self.job.tau, None, None)

    def bucket_key(self):
        """Jobs with equal keys share fused launches: the same kind and
        particle count (hence block sizes and register tile), precision,
        orbital form and block length."""
        j = self.job
        return (j.kind, j.n_particles, j.precision, self.complex_orbitals, j.sweeps_per_block)

    @staticmethod
    def start_fused(runners):
        """One block of every runner's job as fused launches
        (alea_vmc_sweeps_fused), queued on the first runner's stream."""
        vmc_sweeps_fused([r.w for r in runners], runners[0].job.sweeps_per_block, [r.job.tau for r in runners])

    def done(self) -> bool:
        import ctypes as C
        d = C.c_int()
        self.w.api("query", self.w.h, C.byref(d))
        return d.value != 0

    def block_energy(self) -> float:
        e, _ = self.w.local_energy()
        return float(np.mean(e)) / self.job.n_particles

    def close(self):
        self.w.close()


def _pct(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else float("nan")


class Server:
    """Admission by HBM footprint and free slots, FIFO by arrival.

    With fuse (and a runner that has bucket_key and start_fused), active jobs
    of one size bucket form a group whose blocks run as fused launches: one
    sweep kernel over the walkers of every member.  A job admitted while its
    bucket's group is mid-block joins the group at the group's next block;
    a finished member leaves at once.  Groups of different buckets run on
    their own streams and overlap on the device."""

    def __init__(self, hbm_budget_bytes: int = 64 << 30, max_active: int = 8, runner=GpuRunner,
                 clock=time.perf_counter, poll_interval: float = 1e-4, on_finish=None, fuse: bool = True,
                 on_block=None):
        self.fuse = fuse and hasattr(runner, "start_fused") and hasattr(runner, "bucket_key")
        self.budget = hbm_budget_bytes
        self.max_active = max_active
        self.runner = runner
        self.clock = clock
        self.poll_interval = poll_interval
        self.peak_bytes = 0
        self.peak_active = 0
        self.peak_fused = 1     # most jobs in one fused launch
        self.on_finish = on_finish      # called with each finished job (progress reporting)
        self.on_block = on_block        # called with a job after each of its blocks

    def run(self, jobs: list) -> dict:
        if self.fuse:
            return self._run_fused(jobs)
        pending = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        for j in pending:
            if job_footprint(j) > self.budget:
                raise ValueError(f"job {j.job_id} needs more than the whole HBM budget")
        active = []          # (runner, bytes, block_start)
        used = 0
        t0 = self.clock()
        done_jobs = []
        while pending or active:
            now = self.clock() - t0
            # admit in arrival order while memory and slots allow
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                j = pending[0]
                need = job_footprint(j)
                if used + need > self.budget:
                    break
                pending.pop(0)
                r = self.runner(j)
                j.admitted = self.clock() - t0
                r.start_block()
                active.append([r, need, self.clock() - t0])
                used += need
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_active = max(self.peak_active, len(active))
            progressed = False
            for entry in list(active):
                r, need, start = entry
                if not r.done():
                    continue
                progressed = True
                j = r.job
                j.block_latency.append(self.clock() - t0 - start)
                j.block_energy.append(r.block_energy())
                if self.on_block:
                    self.on_block(j)
                if len(j.block_energy) < j.n_blocks:
                    entry[2] = self.clock() - t0
                    r.start_block()
                else:
                    j.finished = self.clock() - t0
                    r.close()
                    active.remove(entry)
                    used -= need
                    done_jobs.append(j)
                    if self.on_finish:
                        self.on_finish(j)
            if not progressed:
                if not active and pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                else:
                    time.sleep(self.poll_interval)
        return self._stats(done_jobs, self.clock() - t0)

    def _run_fused(self, jobs: list) -> dict:
        pending = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        for j in pending:
            if job_footprint(j) > self.budget:
                raise ValueError(f"job {j.job_id} needs more than the whole HBM budget")
        groups = []          # {"key", "members": [[runner, bytes, block_start]], "joining": [...]}
        used = 0
        n_active = 0
        t0 = self.clock()
        done_jobs = []

        def launch(g):
            g["members"] += g["joining"]
            g["joining"] = []
            now = self.clock() - t0
            for e in g["members"]:
                e[2] = now
            self.peak_fused = max(self.peak_fused, len(g["members"]))
            self.runner.start_fused([e[0] for e in g["members"]])

        while pending or groups:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now and n_active < self.max_active:
                j = pending[0]
                need = job_footprint(j)
                if used + need > self.budget:
                    break
                pending.pop(0)
                r = self.runner(j)
                j.admitted = self.clock() - t0
                entry = [r, need, None]
                key = r.bucket_key()
                g = next((g for g in groups if g["key"] == key
                          and len(g["members"]) + len(g["joining"]) < MAX_FUSED), None)
                if g is None:
                    g = {"key": key, "members": [], "joining": [entry]}
                    groups.append(g)
                    launch(g)
                else:
                    g["joining"].append(entry)      # joins at the group's next block
                used += need
                n_active += 1
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_active = max(self.peak_active, n_active)
            progressed = False
            for g in list(groups):
                if not g["members"][0][0].done():     # the block runs on the first member's stream
                    continue
                progressed = True
                end = self.clock() - t0
                for e in list(g["members"]):
                    r, need, start = e
                    j = r.job
                    j.block_latency.append(end - start)
                    j.block_energy.append(r.block_energy())
                    if self.on_block:
                        self.on_block(j)
                    if len(j.block_energy) >= j.n_blocks:
                        j.finished = self.clock() - t0
                        r.close()
                        g["members"].remove(e)
                        used -= need
                        n_active -= 1
                        done_jobs.append(j)
                        if self.on_finish:
                            self.on_finish(j)
                if g["members"] or g["joining"]:
                    launch(g)
                else:
                    groups.remove(g)
            if not progressed:
                if not groups and pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                else:
                    time.sleep(self.poll_interval)
        return self._stats(done_jobs, self.clock() - t0)

    def _stats(self, done_jobs, wall) -> dict:
        blocks = [x for j in done_jobs for x in j.block_latency]
        requests = [j.finished - j.arrival for j in done_jobs]
        waits = [j.admitted - j.arrival for j in done_jobs]
        moves = sum(j.walker_moves for j in done_jobs)
        return dict(jobs=len(done_jobs), wall_s=wall, walker_moves=moves,
                    walker_moves_per_s=moves / wall if wall > 0 else float("nan"),
                    p50_block_ms=1e3 * _pct(blocks, 50), p99_block_ms=1e3 * _pct(blocks, 99),
                    p50_request_s=_pct(requests, 50), p99_request_s=_pct(requests, 99),
                    p50_admission_wait_s=_pct(waits, 50), peak_active=self.peak_active,
                    peak_fused=self.peak_fused, peak_bytes=self.peak_bytes, budget_bytes=self.budget)
