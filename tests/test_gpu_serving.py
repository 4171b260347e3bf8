"""Serving on the device: a job's block energies are the same when it runs
alone and when it shares the GPU with other jobs admitted around it (each job
is its own walker batch on its own stream, with random numbers from its own
seed), and a stream of mixed-size requests completes with every block
measured."""
import numpy as np
import pytest

from alea_amd import serving
from alea_amd.walkers import FP64, MIXED

pytestmark = pytest.mark.gpu


def _job(i, n, nw, arrival=0.0, precision=MIXED, twist=(0.1, 0.2)):
    return serving.Job(i, arrival, n, 30.0, twist, nw, n_blocks=3, sweeps_per_block=4,
                       seed=1000 + i, precision=precision)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_job_results_same_alone_and_in_mix(precision):
    alone = _job(0, 26, 512, precision=precision)
    serving.Server(max_active=4).run([alone])
    mix = [_job(0, 26, 512, precision=precision), _job(1, 122, 2048, 0.0, precision, (0.13, 0.27)),
           _job(2, 58, 1024, 0.001, precision, (-0.21, 0.05)), _job(3, 26, 4096, 0.002, precision, (0.3, -0.3))]
    st = serving.Server(max_active=4).run(mix)
    assert st["jobs"] == 4 and st["peak_active"] >= 2          # they did share the device
    assert alone.block_energy == mix[0].block_energy            # bit for bit
    assert all(len(j.block_energy) == 3 and np.all(np.isfinite(j.block_energy)) for j in mix)


def test_poisson_stream_completes():
    jobs = serving.poisson_jobs(6, rate=50.0, seed=4, walkers=(512, 2048), n_blocks=2, sweeps_per_block=3,
                                dipole_fraction=0.5, dipoles=(64, 1.0))
    assert {j.kind for j in jobs} == {"heg", "dipoles"}
    st = serving.Server(max_active=4).run(jobs)
    assert st["jobs"] == 6 and st["p99_block_ms"] >= st["p50_block_ms"] > 0
    assert all(abs(e) < 1.0 for j in jobs if j.kind == "heg" for e in j.block_energy)
    # a few sweeps from a uniform start: the McMillan core still dominates
    # E_L of close pairs, so only finiteness is a property here
    assert all(np.isfinite(e) for j in jobs if j.kind == "dipoles" for e in j.block_energy)
