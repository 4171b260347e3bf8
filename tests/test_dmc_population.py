"""The DMC population step on the CPU: the oracle's branching and the host
population controller against the reference's own WalkerSet::branch,
PopulationControl and robust_spread (oracle/_ref), the reference's
integer-weight branch cases (tests/test_walker_set.cpp:299-366), and the
exciton's exact population (E_L = -1 everywhere: every weight is one)."""
import ctypes as C
import math

import numpy as np
import pytest

from alea_amd import dmc
from alea_amd.system import exciton_spec, heg_spec, uniform_positions
from backends import oracle_lib, oracle_walkers, ref_lib

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def _next_uniform(state):
    lib = oracle_lib()[0]
    out = np.empty(1, dtype=np.uint64)
    st = state.copy()
    lib.oracle_xoshiro_state_u64(st.ctypes.data_as(C.POINTER(C.c_uint64)), 1,
                                 out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return float(int(out[0]) >> 11) * 2.0 ** -53


def _oracle_batch(weights):
    """An exciton (two-particle, open cell) oracle batch carrying the given weights."""
    s = exciton_spec()
    nw = len(weights)
    o = oracle_walkers(s, nw)
    pos = uniform_positions(s, nw, np.random.default_rng(0), box=3.0)
    o.set_positions(pos)
    o.evaluate()
    o.set_weights(np.asarray(weights, dtype=np.float64))
    return o, pos


def _oracle_branch(weights, seed, target):
    o, pos = _oracle_batch(weights)
    state = dmc.xoshiro_state(seed)
    src = o.branch(state, target)
    # the permutation carried the positions: walker j now holds walker src[j]'s
    P = pos.reshape(2, 2, len(weights))
    assert np.array_equal(o.positions().reshape(2, 2, -1), P[:, :, src])
    assert np.all(o.weights() == 1.0)
    return src, _next_uniform(state)


CASES = [
    ([0.0, 1.0, 2.0, 3.0], 123, 4),          # integer weights: m = k exactly
    ([5.0] * 4, 1, 2),                        # explosion: clipped to 2*target+1
    ([1.0, 0.0, 0.0, 0.0], 2, 20),            # collapse: lifted to target/5+1
    ([0.0, 0.0, 0.0], 3, 4),                  # extinction: the best walker survives
]


def test_branch_known_answers():
    """The reference's own branch cases (tests/test_walker_set.cpp:299-366)."""
    src, _ = _oracle_branch(*CASES[0])
    assert list(src) == [1, 2, 2, 3, 3, 3]
    assert len(_oracle_branch(*CASES[1])[0]) == 5
    assert len(_oracle_branch(*CASES[2])[0]) == 5
    assert len(_oracle_branch(*CASES[3])[0]) >= 1


def _random_cases():
    rng = np.random.default_rng(7)
    out = list(CASES)
    for nw, target, spread, seed in [(200, 200, 0.3, 11), (500, 300, 1.0, 12), (64, 8, 2.0, 13),
                                     (300, 1000, 0.1, 14), (40, 100, 0.05, 15)]:
        out.append((list(np.exp(rng.normal(0.0, spread, nw))), seed, target))
    return out


@pytest.mark.parametrize("case", range(len(_random_cases())))
def test_branch_matches_reference(case):
    weights, seed, target = _random_cases()[case]
    lib = ref_lib()
    w = np.asarray(weights, dtype=np.float64)
    src = np.empty(2 * target + 1, dtype=np.int32)
    n = C.c_int()
    nxt = C.c_double()
    assert lib.ref_branch(w.ctypes.data_as(D), len(w), seed, target,
                          src.ctypes.data_as(I), C.byref(n), C.byref(nxt)) == 0
    osrc, onext = _oracle_branch(weights, seed, target)
    assert list(osrc) == list(src[:n.value])
    assert onext == nxt.value


@pytest.mark.parametrize("cutoff", [0, 1, 2])
def test_population_control_matches_reference(cutoff):
    """The host controller against PopulationControl: weights, cut flags, E_T
    after every update, e_cut; compared exactly (the same libm calls)."""
    lib = ref_lib()
    names = ["none", "drv", "universal"]
    rng = np.random.default_rng(cutoff)
    tau, n_sub, g, target, alpha, npart = 0.7, 2, 10.0, 500, 2.0, 26
    e0, sigma = -3.1, 0.45
    n_steps, n_pairs = 12, 64
    ens = -3.1 + 0.05 * rng.standard_normal(n_steps)
    wsum = target * np.exp(0.1 * rng.standard_normal(n_steps))
    acc = (np.arange(n_steps) >= 6).astype(np.int32)
    e_new = -3.1 + 2.0 * rng.standard_normal(n_pairs)
    e_old = -3.1 + 2.0 * rng.standard_normal(n_pairs)
    wout = np.empty(n_steps * n_pairs)
    cout = np.empty(n_steps * n_pairs, dtype=np.int32)
    et = np.empty(n_steps)
    ecut = C.c_double()
    assert lib.ref_popctl_run(tau, n_sub, g, target, cutoff, alpha, npart, e0, sigma, n_steps,
                              ens.ctypes.data_as(D), wsum.ctypes.data_as(D), acc.ctypes.data_as(I),
                              n_pairs, e_new.ctypes.data_as(D), e_old.ctypes.data_as(D),
                              wout.ctypes.data_as(D), cout.ctypes.data_as(I), et.ctypes.data_as(D),
                              C.byref(ecut)) == 0
    pc = dmc.PopulationControl(tau, n_sub, g, target, names[cutoff], alpha, npart)
    pc.start(e0)
    pc.set_energy_spread(sigma)
    assert pc.e_cut == ecut.value
    for i in range(n_steps):
        for k in range(n_pairs):
            assert pc.branch_weight(e_new[k], e_old[k]) == wout[i * n_pairs + k]
            assert pc.is_cut(e_new[k], e_old[k]) == bool(cout[i * n_pairs + k])
        pc.update(ens[i], wsum[i], bool(acc[i]))
        assert pc.etrial == et[i]


@pytest.mark.parametrize("n", [1, 2, 7, 64, 1001])
def test_robust_spread_matches_reference(n):
    lib = ref_lib()
    x = np.random.default_rng(n).standard_normal(n) * 3.0 + 1.0
    assert dmc.robust_spread(x) == lib.ref_robust_spread(x.ctypes.data_as(D), n)


def test_oracle_weights_follow_population_control():
    """oracle_dmc_weights applies PopulationControl::branch_weight per walker."""
    s = heg_spec(5, 5, 2.0, slater=False)
    nw = 50
    o = oracle_walkers(s, nw)
    o.set_positions(uniform_positions(s, nw, np.random.default_rng(1)))
    o.evaluate()
    e0, _ = o.local_energy()
    o.dmc_begin()
    pc = dmc.PopulationControl(0.05, 1, 10.0, nw, "universal", 2.0, s.n_particles)
    pc.start(float(np.median(e0)) + 2.25)         # E_ref - E_L straddles e_cut
    pc.set_energy_spread(0.02)                     # small: some walkers are cut
    tape = np.random.default_rng(2).random(s.n_particles * 3 * nw)
    tape[: s.n_particles * 3 * nw] = np.random.default_rng(3).standard_normal(s.n_particles * 3 * nw)
    rr_p, rr_a = np.zeros(nw), np.zeros(nw)
    o.dmc_sweeps(1, 0.05, tape, nw, rr_p, rr_a)
    o.dmc_weights(pc.etrial, pc.eref, pc.e_cut, pc.e_max, pc.tau)
    e1, _ = o.local_energy()
    want = np.array([pc.branch_weight(e1[w], e0[w]) for w in range(nw)])
    assert np.array_equal(o.weights(), want)
    es, ws, nc = o.dmc_sums()
    assert nc == sum(pc.is_cut(e1[w], e0[w]) for w in range(nw)) and 0 < nc < nw
    assert ws == math.fsum(want) or abs(ws - math.fsum(want)) < 1e-12 * ws


def test_exciton_population_is_exact():
    """E_L = -1 for every configuration: E_T = E_ref = -1, every weight is
    exp(0) = 1, every multiplicity int(1 + xi) = 1, the population and the
    walker order never change."""
    s = exciton_spec()
    nw = 40
    o = oracle_walkers(s, nw)
    o.set_positions(uniform_positions(s, nw, np.random.default_rng(4), box=3.0))
    o.evaluate()
    tapes = lambda step, sub, n: np.random.default_rng(100 + step).standard_normal(s.n_particles * 3 * nw)
    d = dmc.DMC(o, 0.05, nw, n_substeps=2, seed=5, tapes=tapes, stride=nw)
    for _ in range(5):
        e, w = d.step()
        assert w == nw and abs(e + nw) < 1e-12 * nw       # E_L = -1 to rounding
    assert all(h["nw"] == nw and list(h["src"]) == list(range(nw)) for h in d.history)
    assert abs(d.pc.etrial + 1.0) < 1e-12
