"""alea_vmc_sweeps_static (the sweep entry point that allocates nothing) gives
bit for bit what alea_vmc_sweeps gives on the device streams: two calls in a
row, the first with more walkers than one chunk of the static tape holds
(N=122 twisted, mixed: several chunks), the second with a smaller batch
(N=26, FP64)."""
import numpy as np
import pytest

from alea_amd.system import heg_spec, uniform_positions
from alea_amd.walkers import FP64, MIXED, gpu_walkers, seed

pytestmark = pytest.mark.gpu


def _pair(spec, nw, prec, seed_):
    pos = uniform_positions(spec, nw, np.random.default_rng(seed_))
    out = []
    for _ in range(2):
        w = gpu_walkers(spec, nw, prec)
        w.set_positions(pos)
        w.evaluate()
        seed(w, 1000 + seed_)
        out.append(w)
    return out


@pytest.mark.parametrize("n_up, nw, prec, n_sweeps", [(61, 12000, MIXED, 2), (13, 300, FP64, 3)])
def test_static_sweeps_equal_the_allocating_entry_point(n_up, nw, prec, n_sweeps):
    spec = heg_spec(n_up, n_up, 30.0, twist=(0.21, -0.13))
    assert spec.n_particles * 3 * nw > (4 << 20) or n_up < 61   # the first case spans chunks
    a, b = _pair(spec, nw, prec, n_up)
    tau = (0.5 * 30.0) ** 2
    try:
        a.vmc_sweeps(n_sweeps, tau)
        b.api("vmc_sweeps_static", b.h, n_sweeps, tau)
        b.check_status()
        assert np.array_equal(a.positions(), b.positions())
        assert np.array_equal(a.log_psi(), b.log_psi())
        assert np.array_equal(a.local_energy()[0], b.local_energy()[0])
    finally:
        a.close()
        b.close()


def test_static_sweeps_refuse_an_unseeded_system():
    spec = heg_spec(5, 5, 2.0)
    w = gpu_walkers(spec, 8)
    try:
        w.set_positions(uniform_positions(spec, 8, np.random.default_rng(0)))
        w.evaluate()
        with pytest.raises(RuntimeError, match="seed"):
            w.api("vmc_sweeps_static", w.h, 1, 1.0)
    finally:
        w.close()
