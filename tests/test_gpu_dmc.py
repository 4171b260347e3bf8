"""DMC on the device against the oracle on the same tapes and branch stream.

The plain FP64 form must give identical Metropolis decisions, positions to
1e-12, rr sums and energies to 1e-10 relative, and — through weights,
integer branching and the E_T feedback — identical population counts and
source maps at every step.  The exact exciton keeps its population and
E_L = -1 with the device's own streams.
"""
import numpy as np
import pytest

from alea_amd import dmc
from alea_amd.system import exciton_spec, heg_drummond_spec, heg_spec
from alea_amd.walkers import gpu_walkers, seed
from backends import oracle_tape, oracle_walkers, pos_err, rel_err
from test_gpu_parity import _pos, specs

pytestmark = pytest.mark.gpu


def _tau(spec):
    return 0.05 if spec.lattice is None else (spec.lattice[0] / 40) ** 2


@pytest.mark.parametrize("spec", specs(), ids=lambda s: s.name)
def test_dmc_sweeps_parity(spec):
    nw, stride, n_sw = 24, 32, 3
    pos = _pos(spec, nw, 12)
    g, o = gpu_walkers(spec, nw), oracle_walkers(spec, nw)
    for x in (g, o):
        x.set_positions(pos)
        x.evaluate()
    tape = oracle_tape(31, spec.n_particles, stride, n_sw)
    tau = _tau(spec)
    rg = [np.zeros(nw), np.zeros(nw)]
    ro = [np.zeros(nw), np.zeros(nw)]
    ag = g.dmc_sweeps(n_sw, tau, tape, stride, *rg)
    ao = o.dmc_sweeps(n_sw, tau, tape, stride, *ro)
    assert ag == ao and ag > 0
    # Drifted positions carry the drift's 1e-10 bar.  The McMillan core
    # (b/r)^5 amplifies a 1e-12 position difference through three drifted
    # sweeps to about 1e-10 in log psi, so that system is held to 1e-9.
    tol = 1e-9 if spec.mcmillan else 1e-10
    assert pos_err(g.positions(), o.positions()) < tol
    assert rel_err(rg[0], ro[0]) < tol and rel_err(rg[1], ro[1]) < tol
    assert rel_err(g.log_psi(), o.log_psi()) < tol
    assert rel_err(g.local_energy()[0], o.local_energy()[0]) < tol


def _dmc_specs():
    a = heg_spec(5, 5, 2.0)
    a.name = "plane-wave N=10 (fixed node)"
    b = heg_spec(5, 5, 2.0, twist=(0.13, 0.27))
    b.name = "twisted N=10 (fixed phase)"
    return [a, b]


@pytest.mark.parametrize("spec", _dmc_specs(), ids=lambda s: s.name)
def test_dmc_steps_identical_populations(spec):
    cap, target, n_steps = 64, 24, 5
    pos = _pos(spec, cap, 21)
    tapes = lambda step, sub, n: oracle_tape(1000 + 10 * step + sub, spec.n_particles, cap, 1)
    runs = []
    for make in (gpu_walkers, oracle_walkers):
        x = make(spec, cap)
        x.set_positions(pos)
        x.evaluate()
        x.permute(np.arange(target))
        d = dmc.DMC(x, _tau(spec), target, n_substeps=2, seed=9, tapes=tapes, stride=cap)
        for i in range(n_steps):
            d.step(measure=i >= 2)
        runs.append((d, x))
    (dg, g), (do, o) = runs
    assert [h["nw"] for h in dg.history] == [h["nw"] for h in do.history]
    for hg, ho in zip(dg.history, do.history):
        assert list(hg["src"]) == list(ho["src"])
        assert hg["n_accepted"] == ho["n_accepted"] and hg["n_cut"] == ho["n_cut"]
        assert rel_err(hg["e_sum"], ho["e_sum"]) < 1e-10 and rel_err(hg["w_sum"], ho["w_sum"]) < 1e-10
        assert abs(hg["etrial"] - ho["etrial"]) < 1e-10 * abs(ho["etrial"])
    assert len(set(h["nw"] for h in dg.history)) > 1          # the population did move
    assert np.array_equal(dg.branch_state, do.branch_state)
    assert pos_err(g.positions(), o.positions()) < 1e-12


def test_c1_dmc_identical_populations():
    """The published DMC case's system (N = 74 paramagnetic HEG at rs = 40,
    plane-wave nodes, image-summed Padé plus an order-8 Drummond polynomial,
    Ewald) at tau = 2: target 32 in a capacity of 65, five steps of two
    drift-diffusion sub-steps from one tape per sub-step and one branch
    stream.  Device and oracle keep identical populations, source maps,
    accept and cutoff counts at every step."""
    spec = heg_drummond_spec(37, 37, 40.0)
    cap, target, n_steps = 65, 32, 5
    pos = _pos(spec, cap, 23)
    tapes = lambda step, sub, n: oracle_tape(2000 + 10 * step + sub, spec.n_particles, cap, 1)
    runs = []
    for make in (gpu_walkers, oracle_walkers):
        x = make(spec, cap)
        x.set_positions(pos)
        x.evaluate()
        x.permute(np.arange(target))
        d = dmc.DMC(x, 2.0, target, n_substeps=2, seed=19, tapes=tapes, stride=cap)
        for i in range(n_steps):
            d.step(measure=i >= 2)
        runs.append((d, x))
    (dg, g), (do, o) = runs
    assert [h["nw"] for h in dg.history] == [h["nw"] for h in do.history]
    for hg, ho in zip(dg.history, do.history):
        assert list(hg["src"]) == list(ho["src"])
        assert hg["n_accepted"] == ho["n_accepted"] and hg["n_cut"] == ho["n_cut"]
        assert rel_err(hg["e_sum"], ho["e_sum"]) < 1e-10 and rel_err(hg["w_sum"], ho["w_sum"]) < 1e-10
    assert len(set(h["nw"] for h in dg.history)) > 1
    assert np.array_equal(dg.branch_state, do.branch_state)
    assert pos_err(g.positions(), o.positions()) < 1e-10


def test_c1_mixed_drift_sweeps_track_fp64():
    """Config 1's system in the mixed form, whose 37-orbital real blocks take
    the 48-row register tile in drift sweeps (the FP64 form streams them from
    HBM): from the same walkers and tape, three drift-diffusion sweeps leave at
    least 95% of the walkers on the FP64 trajectory (positions within 1e-6 of
    the cell scale: the drift reads the single-precision inverse), with the
    accept counts within 2% and the local energy per particle on those walkers
    within 1e-5 Ha; and the mixed walkers' inverse, read fresh after
    refresh_inverse, gives the FP64 energies of the same positions to 1e-7 Ha
    per particle."""
    from alea_amd.walkers import FP64, MIXED
    spec = heg_drummond_spec(37, 37, 40.0)
    nw, n_sw = 128, 3
    pos = _pos(spec, nw, 29)
    tape = oracle_tape(37, spec.n_particles, nw, n_sw)
    out = []
    for prec in (FP64, MIXED):
        x = gpu_walkers(spec, nw, prec)
        try:
            x.set_positions(pos)
            x.evaluate()
            r = [np.zeros(nw), np.zeros(nw)]
            acc = x.dmc_sweeps(n_sw, 2.0, tape, nw, *r)
            swept = x.local_energy()[0].copy()
            p_ = x.positions().copy()
            x.refresh_inverse()
            out.append((acc, p_, swept, x.local_energy()[0].copy()))
        finally:
            x.close()
    (a64, p64, e64, f64_), (a32, p32, e32, f32_) = out
    assert abs(a32 - a64) <= 0.02 * a64 and a64 > 0
    scale = np.abs(p64).max()
    same = np.all(np.abs(p32 - p64).reshape(-1, nw) < 1e-6 * scale, axis=0)
    assert same.mean() >= 0.95
    n = spec.n_particles
    assert np.abs((e32 - e64)[same]).max() / n < 1e-5
    fresh = heg_drummond_spec(37, 37, 40.0)
    g = gpu_walkers(fresh, nw, FP64)
    try:
        g.set_positions(p32)
        g.evaluate()
        assert np.abs(f32_ - g.local_energy()[0]).max() / n < 1e-7
    finally:
        g.close()


def test_dmc_exciton_population_exact_on_device():
    s = exciton_spec()
    nw = 256
    g = gpu_walkers(s, 2 * nw + 1)          # capacity for the population bound 2 target + 1
    g.set_positions(_pos(s, 2 * nw + 1, 13))
    g.evaluate()
    g.permute(np.arange(nw))
    seed(g, 17)
    d = dmc.DMC(g, 0.05, nw, n_substeps=3, seed=3)
    for _ in range(4):
        e, w = d.step()
        assert w == nw and abs(e + nw) < 1e-10 * nw
    assert all(h["nw"] == nw and list(h["src"]) == list(range(nw)) for h in d.history)
    assert np.allclose(g.local_energy()[0], -1.0, atol=1e-12)
