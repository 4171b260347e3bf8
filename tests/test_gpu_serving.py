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


@pytest.mark.parametrize("fuse", [True, False], ids=["fused", "per-job graphs"])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_job_results_same_alone_and_in_mix(precision, fuse):
    alone = _job(0, 26, 512, precision=precision)
    serving.Server(max_active=4, fuse=fuse).run([alone])
    mix = [_job(0, 26, 512, precision=precision), _job(1, 122, 2048, 0.0, precision, (0.13, 0.27)),
           _job(2, 58, 1024, 0.001, precision, (-0.21, 0.05)), _job(3, 26, 4096, 0.002, precision, (0.3, -0.3))]
    st = serving.Server(max_active=4, fuse=fuse).run(mix)
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


def _graph_pair(precision, nw=256):
    from alea_amd.system import heg_spec, uniform_positions
    from alea_amd.walkers import gpu_walkers, seed, set_sweep_graphs
    spec = heg_spec(13, 13, 30.0, twist=(0.13, 0.27))
    pos = uniform_positions(spec, nw, np.random.default_rng(5))
    pair = []
    for graphs in (False, True):
        w = gpu_walkers(spec, nw, precision)
        w.set_positions(pos)
        w.evaluate()
        seed(w, 11)
        set_sweep_graphs(w, graphs)
        pair.append(w)
    return spec, pair


def _same_state(a, b):
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.log_psi(), b.log_psi())
    ea, eb = a.local_energy()[0], b.local_energy()[0]
    assert np.array_equal(ea, eb)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_sweep_graph_replays_match_eager_sweeps(precision):
    """Blocks of device-stream sweeps replayed from a captured hipGraph give
    the eager launches' walkers bit for bit: repeated replays with eager work
    (energies, reads) between them, blocks of 3 sweeps that shift the refresh
    phase (refresh every 2 sweeps), and a permutation that halves the walker
    count, all of which must re-capture or replay the right graph."""
    spec, (eager, graph) = _graph_pair(precision)
    tau = 15.0 ** 2
    for block in range(4):
        na = eager.vmc_sweeps(3, tau)
        nb = graph.vmc_sweeps(3, tau)
        assert na == nb > 0, (block, na, nb)
        _same_state(eager, graph)
    src = np.arange(0, eager.nw, 2)
    for w in (eager, graph):
        w.permute(src)
    for block in range(2):
        assert eager.vmc_sweeps(2, tau) == graph.vmc_sweeps(2, tau)
        _same_state(eager, graph)
    for w in (eager, graph):
        w.close()


def test_sweep_graph_dmc_sweeps_match_eager():
    from alea_amd import dmc
    from alea_amd.walkers import seed
    spec, (eager, graph) = _graph_pair(FP64)
    for w in (eager, graph):
        seed(w, 11, dmc.K_DMC_THREAD_STREAM)
        w.dmc_begin()
    for block in range(3):
        assert eager.dmc_sweeps(2, 2.0) == graph.dmc_sweeps(2, 2.0)
        _same_state(eager, graph)
    for w in (eager, graph):
        w.close()


def _fused_systems(kind, precision, sizes, n):
    """Systems of one size bucket that differ in density, twist and walker
    count: HEG N=n at rs 30/40/60, or dipoles at three densities."""
    from alea_amd.system import dipole_spec, heg_spec, uniform_positions
    from alea_amd.walkers import gpu_walkers, seed
    out = []
    for i, (rs, nw) in enumerate(sizes):
        if kind == "dipoles":
            spec = dipole_spec(n, density=1.0 / (np.pi * rs * rs))
        else:
            spec = heg_spec(n // 2, n // 2, rs, twist=(0.1 * (i + 1), -0.07 * (i + 1)))
        w = gpu_walkers(spec, nw, precision)
        w.set_positions(uniform_positions(spec, nw, np.random.default_rng(20 + i)))
        w.evaluate()
        seed(w, 300 + i)
        out.append((w, (0.5 * rs) ** 2))
    return out


@pytest.mark.parametrize("kind,precision,n", [
    ("heg", MIXED, 26),      # 13-orbital blocks: 16-row register tile
    ("heg", MIXED, 74),      # 37-orbital blocks: 64-row register tile
    ("heg", FP64, 58),       # complex FP64, 29 orbitals: 32-row tile
    ("heg", FP64, 242),      # 121 orbitals: the inverse streams from HBM
    ("dipoles", FP64, 64),   # Jastrow only
])
def test_fused_sweeps_match_each_system_alone(kind, precision, n):
    """alea_vmc_sweeps_fused over three systems of one bucket gives every
    walker exactly what alea_vmc_sweeps gives on its own system: positions,
    log psi and local energies bit for bit, over two blocks (refresh phases
    included) with eager work between them."""
    from alea_amd.walkers import vmc_sweeps_fused
    sizes = [(30.0, 192), (40.0, 64), (60.0, 320)]
    alone = _fused_systems(kind, precision, sizes, n)
    fused = _fused_systems(kind, precision, sizes, n)
    for block in range(2):
        for w, tau in alone:
            w.vmc_sweeps(3, tau)
        vmc_sweeps_fused([w for w, _ in fused], 3, [tau for _, tau in fused])
        for (a, _), (b, _) in zip(alone, fused):
            _same_state(a, b)
    for w, _ in alone + fused:
        w.close()


def test_fused_launch_refuses_mixed_buckets():
    from alea_amd.walkers import vmc_sweeps_fused
    (a, ta), = _fused_systems("heg", MIXED, [(30.0, 64)], 26)
    (b, tb), = _fused_systems("heg", MIXED, [(30.0, 64)], 58)     # another register tile
    (c, tc), = _fused_systems("heg", FP64, [(30.0, 64)], 26)      # another precision
    for other, t in ((b, tb), (c, tc)):
        with pytest.raises(RuntimeError, match="share the precision"):
            vmc_sweeps_fused([a, other], 1, [ta, t])
    with pytest.raises(RuntimeError, match="appears twice"):
        vmc_sweeps_fused([a, a], 1, [ta, ta])
    for w in (a, b, c):
        w.close()


def test_serving_fuses_a_size_bucket():
    """Four N=26 jobs arriving together run as one fused group; each job's
    block energies equal its run alone."""
    twists = [(0.1, 0.2), (0.13, 0.27), (-0.21, 0.05), (0.3, -0.3)]
    jobs = [_job(i, 26, 256 * (i + 1), arrival=0.0, twist=twists[i]) for i in range(4)]
    st = serving.Server(max_active=8).run(jobs)
    assert st["jobs"] == 4 and st["peak_fused"] >= 2
    for i in range(4):
        alone = _job(i, 26, 256 * (i + 1), arrival=0.0, twist=twists[i])
        serving.Server(max_active=1).run([alone])
        assert alone.block_energy == jobs[i].block_energy, i
