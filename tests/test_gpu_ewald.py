"""The Ewald structure factor kept per walker on the device, and
Hamiltonian::delta_potential_particle through it.

ewald_rho keeps rho(k) = Sum_i q_i e^{ik.r_i} of each walker with the
positions it describes and folds in the particles moved since (the reference's
sync_delta_cache, ewald_term.cpp:675-747): delta_potential takes up to four
moved particles incrementally, the local energy up to half the particles.  The
oracle recomputes rho from scratch every time, so these tests hold the kept
sums to the fresh ones: 1e-10 relative, with identical Metropolis decisions.
"""
import numpy as np
import pytest

from alea_amd.system import heg_spec, uniform_positions
from alea_amd.walkers import FP64, gpu_walkers
from backends import oracle_tape, oracle_walkers, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(spec, nw, seed_):
    pos = uniform_positions(spec, nw, np.random.default_rng(seed_))
    g, o = gpu_walkers(spec, nw, FP64), oracle_walkers(spec, nw)
    for x in (g, o):
        x.set_positions(pos)
        x.evaluate()
    return g, o


@pytest.mark.parametrize("twist", [None, (0.13, 0.27)], ids=["plane-wave", "twisted"])
def test_delta_potential_matches_the_oracle(twist):
    """N=26 at rs=30 (Slater-Jastrow-Ewald): the delta of fresh walkers (the
    kept rho computed in full), after one committed particle on some walkers
    (folded in: one moved particle), after a whole sweep (about a third of
    the particles moved: recomputed), and on a walker range."""
    spec = heg_spec(13, 13, 30.0, twist=twist)
    nw = 32
    g, o = _pair(spec, nw, 3)
    rng = np.random.default_rng(4)

    def check(p):
        P = o.positions().reshape(spec.n_particles, 2, nw)
        newp = (P[p] + rng.normal(scale=15.0, size=(2, nw))).reshape(-1)
        dg, do = g.delta_potential(p, newp), o.delta_potential(p, newp)
        assert rel_err(dg, do) < TOL
        assert np.abs(do).max() > 0.0

    check(3)
    check(3)   # nothing moved since: the kept rho as it is
    P = o.positions().reshape(spec.n_particles, 2, nw)
    newp = (P[5] + rng.normal(scale=15.0, size=(2, nw))).reshape(-1)
    for x in (g, o):
        x.ratio(5, newp)
    mask = (rng.random(nw) < 0.5).astype(np.uint8)
    for x in (g, o):
        x.accept(5, mask)
    check(7)
    tau = (0.5 * 30.0) ** 2
    tape = oracle_tape(9, spec.n_particles, nw, 1)
    assert g.vmc_sweeps(1, tau, tape) == o.vmc_sweeps(1, tau, tape)
    check(11)
    # a walker range: only its entries are written
    P = o.positions().reshape(spec.n_particles, 2, nw)
    newp = (P[2] + rng.normal(scale=15.0, size=(2, nw))).reshape(-1)
    dg = g.delta_potential(2, newp, 8, 20)
    do = o.delta_potential(2, newp, 8, 20)
    assert rel_err(dg[8:20], do[8:20]) < TOL
    assert np.all(dg[:8] == 0.0) and np.all(dg[20:] == 0.0)
    g.close()


def test_local_energy_through_the_kept_structure_factor():
    """The benchmark system (twisted N=122, rs=30): a measurement after every
    sweep folds the accepted moves into the kept rho (a quarter to a half of
    the particles move per sweep); over eight sweeps E_L stays within 1e-10
    of the oracle's full sums, with identical decisions."""
    spec = heg_spec(61, 61, 30.0, twist=tuple(np.random.default_rng(3).uniform(-0.5, 0.5, 2)))
    nw = 8
    g, o = _pair(spec, nw, 17)
    tau = (0.5 * 30.0) ** 2
    assert rel_err(g.local_energy()[0], o.local_energy()[0]) < TOL
    for k in range(8):
        tape = oracle_tape(50 + k, spec.n_particles, nw, 1)
        assert g.vmc_sweeps(1, tau, tape) == o.vmc_sweeps(1, tau, tape)
        assert rel_err(g.local_energy()[0], o.local_energy()[0]) < TOL
    g.close()


def test_the_kept_structure_factor_follows_a_permutation():
    """After a DMC-style permutation every walker holds another's positions:
    the comparison with the kept snapshot sees them as moved, and E_L and the
    delta stay those of the oracle."""
    spec = heg_spec(13, 13, 30.0)
    nw = 16
    g, o = _pair(spec, nw, 21)
    g.local_energy()
    src = np.random.default_rng(22).integers(0, nw, size=nw).astype(np.int32)
    for x in (g, o):
        x.permute(src)
    assert rel_err(g.local_energy()[0], o.local_energy()[0]) < TOL
    P = o.positions().reshape(spec.n_particles, 2, nw)
    newp = (P[4] + 10.0).reshape(-1)
    assert rel_err(g.delta_potential(4, newp), o.delta_potential(4, newp)) < TOL
    g.close()
