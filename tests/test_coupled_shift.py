"""The correlated-sampling estimate of the mixed form's DMC shift
(bench.reweighted_shift) on synthetic inputs.

Each walker carries the log-likelihood ratio of its window's path under the
mixed process against the FP64 one; the estimate is sum(W E)/sum(W) - mean(E)
with a jackknife error over walkers.  The checks fix what a slip in the
arithmetic would break: equal processes give exactly zero, a two-valued
weight gives the closed-form shift, a common factor in W changes nothing,
the jackknife error matches the delta-method error on a large sample, and a
process that favours low energies shifts E down.
"""
import math

import numpy as np

import bench


class _Batch:
    def __init__(self, e_total):
        self.e = np.asarray(e_total, dtype=np.float64)

    def local_energy(self):
        return self.e.copy(), np.zeros_like(self.e)


def test_equal_processes_give_no_shift():
    e = np.random.default_rng(1).normal(-1.0, 0.1, 4096) * 10
    est, err, spread = bench.reweighted_shift([_Batch(e)], 10, np.zeros(4096))
    assert est == 0.0 and spread == 0.0
    assert err == 0.0 or err < 1e-15


def test_two_valued_weights_give_the_closed_form():
    # half the walkers at E/N = a with weight w1, half at b with weight w2:
    # the shift is (w1 a + w2 b)/(w1 + w2) - (a + b)/2
    a, b, w1, w2, n = -0.5, 0.25, 1.5, 0.5, 74
    e = np.r_[np.full(100, a * n), np.full(100, b * n)]
    logw = np.r_[np.full(100, math.log(w1)), np.full(100, math.log(w2))]
    est, _, _ = bench.reweighted_shift([_Batch(e)], n, logw)
    assert abs(est - ((w1 * a + w2 * b) / (w1 + w2) - (a + b) / 2)) < 1e-14


def test_a_common_factor_in_the_weights_changes_nothing():
    rng = np.random.default_rng(2)
    e = rng.normal(size=1000)
    logw = rng.normal(scale=0.1, size=1000)
    x = bench.reweighted_shift([_Batch(e)], 1, logw)
    y = bench.reweighted_shift([_Batch(e)], 1, logw + 7.0)
    assert abs(x[0] - y[0]) < 1e-14 and abs(x[1] - y[1]) < 1e-14


def test_jackknife_error_matches_the_delta_method():
    # for W = 1 + eps with small independent eps, sum(W E)/sum(W) - mean(E)
    # is cov(eps, E) to first order; its standard error is
    # std((eps - mean eps)(E - mean E)) / sqrt(m)
    rng = np.random.default_rng(3)
    m = 20000
    e = rng.normal(size=m)
    eps = 1e-3 * rng.normal(size=m)
    est, err, _ = bench.reweighted_shift([_Batch(e)], 1, np.log1p(eps))
    prod = (eps - eps.mean()) * (e - e.mean())
    assert abs(est - prod.mean() / (1 + eps.mean())) < 1e-9
    assert abs(err / (prod.std() / math.sqrt(m)) - 1.0) < 0.05


def test_favouring_low_energies_shifts_down():
    rng = np.random.default_rng(4)
    e = rng.normal(size=5000)
    logw = -0.01 * e          # the perturbed process weights low E up
    est, err, _ = bench.reweighted_shift([_Batch(e)], 1, logw)
    # first order: -0.01 var(E)
    assert est < -5 * err
    assert abs(est + 0.01 * e.var()) < 5 * err
