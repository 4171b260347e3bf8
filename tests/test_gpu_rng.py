"""The per-walker device streams against the oracle's restatement: xoshiro256++
seeded by stream_seed(seed, stream, w) (random.h:52-56, xoshiro_random.h:
29-32), the ziggurat Gaussian with its wedge and tail paths
(xoshiro_random.h:90-108,170-190) and the 53-bit uniform, in the order the
sweep consumes them, over 1.5 million numbers per seed.

Every uniform and every Gaussian of the ziggurat's layers and wedges is
identical bit for bit.  A tail Gaussian (|x| beyond the last layer's edge r)
is r - log1p(-u)/r, and the device's log1p rounds about one tail argument in
a hundred differently from the host C library's (about one number in a
million overall): such a value may differ in its last bit, never by more,
and the stream goes on identically after it (every accept decision agrees).
The Gaussians and uniforms also pass moment and chi-square checks."""
import math

import numpy as np
import pytest
from scipy import special, stats

from alea_amd.system import heg_spec, uniform_positions
from alea_amd.walkers import MIXED, gpu_walkers, seed
from backends import oracle_device_tape

pytestmark = pytest.mark.gpu

ZIGGURAT_R = 3.6541528853610092   # the last layer's edge: beyond it, the tail path


@pytest.mark.parametrize("seed_value, stream", [(4242, 0x766d632d74687264), (7, 3)])
def test_device_streams_equal_the_oracle(seed_value, stream):
    spec = heg_spec(61, 61, 30.0, slater=False)
    nw, n_sweeps = 1024, 3
    w = gpu_walkers(spec, nw, MIXED)
    try:
        w.set_positions(uniform_positions(spec, nw, np.random.default_rng(0)))
        w.evaluate()
        seed(w, seed_value, stream)
        dev = w.draw_tape(n_sweeps)
        dev = np.concatenate([dev, w.draw_tape(1)])   # the streams continue across calls
    finally:
        w.close()
    n = n_sweeps + 1
    ref = oracle_device_tape(seed_value, stream, spec.n_particles, nw, n)
    assert dev.size == ref.size >= 1_000_000
    d = dev.reshape(n, spec.n_particles, 3, nw)
    r = ref.reshape(n, spec.n_particles, 3, nw)
    # uniforms: bit for bit
    assert np.array_equal(d[:, :, 2].view(np.uint64), r[:, :, 2].view(np.uint64))
    # Gaussians: bit for bit inside the last layer's edge; tail values within one ulp
    gd, gr = d[:, :, :2].ravel(), r[:, :, :2].ravel()
    differ = gd.view(np.int64) != gr.view(np.int64)
    tail = np.abs(gr) > ZIGGURAT_R
    assert not np.any(differ & ~tail)
    assert np.all(np.abs(gd[differ].view(np.int64) - gr[differ].view(np.int64)) <= 1)
    assert differ.sum() <= 1e-5 * gd.size and tail.sum() > 50
    # statistics of the device's own numbers
    g, u = gd, d[:, :, 2].ravel()
    assert abs(g.mean()) < 5.0 / math.sqrt(g.size) and abs(g.var() - 1.0) < 5.0 * math.sqrt(2.0 / g.size)
    edges = np.linspace(-4.0, 4.0, 33)
    p = np.diff(special.ndtr(np.concatenate([[-np.inf], edges[1:-1], [np.inf]])))
    assert stats.chisquare(np.histogram(np.clip(g, -3.999, 3.999), edges)[0], p * g.size).pvalue > 1e-4
    assert u.min() >= 0.0 and u.max() < 1.0
    assert stats.chisquare(np.histogram(u, 64, (0.0, 1.0))[0]).pvalue > 1e-4
