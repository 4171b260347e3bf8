"""The batched inverse recompute on the matrix cores (k_refresh_mfma) against
the vector-unit Gauss-Jordan (k_refresh_reg) and the oracle.

Both kernels run the reference's pivoted Gauss-Jordan
(slater_determinant.cpp:905-993) in FP64; they differ in the order of the
additions, so the FP64 form agrees with the oracle to 1e-10 relative either
way and the two kernels agree with each other to 1e-12 on log|psi|.  The
mixed form stores the inverse in single precision, so the kernels agree to
single-precision rounding there.
"""
import numpy as np
import pytest

from alea_amd.system import heg_spec, uniform_positions
from alea_amd.walkers import FP64, MIXED, gpu_walkers
from backends import oracle_tape, oracle_walkers, pos_err, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _twisted122():
    s = heg_spec(61, 61, 30.0, twist=tuple(np.random.default_rng(3).uniform(-0.5, 0.5, 2)))
    s.name = "twisted N=122"
    return s


def _plane74():
    s = heg_spec(37, 37, 40.0)
    s.name = "plane-wave N=74"
    return s


def _pos(spec, nw, seed_):
    return uniform_positions(spec, nw, np.random.default_rng(seed_))


def _run(spec, nw, prec, kind, pos, tape, n_sw, tau):
    x = gpu_walkers(spec, nw, prec)
    try:
        x.set_refresh_kernel(kind)
        x.set_positions(pos)
        x.evaluate()
        first = (x.log_psi().copy(), x.local_energy()[0].copy())
        acc = x.vmc_sweeps(n_sw, tau, tape)
        x.refresh_inverse()
        return first, acc, x.positions(), x.log_psi().copy(), x.local_energy()[0].copy()
    finally:
        x.close()


@pytest.mark.parametrize("spec", [_twisted122(), _plane74()], ids=lambda s: s.name)
def test_matrix_cores_refresh_like_the_vector_units_fp64(spec):
    """Evaluate (every block rebuilt), four sweeps from one tape (the blocks
    that reach 64 updates rebuilt after each), then refresh_inverse: the two
    kernels give the same accept counts and positions, and log|psi| and E_L
    to 1e-12 relative."""
    nw, n_sw = 64, 4
    rs = 30.0 if spec.n_particles == 122 else 40.0
    tau = (0.5 * rs) ** 2
    pos = _pos(spec, nw, 5)
    tape = oracle_tape(11, spec.n_particles, nw, n_sw)
    a = _run(spec, nw, FP64, 0, pos, tape, n_sw, tau)
    b = _run(spec, nw, FP64, 1, pos, tape, n_sw, tau)
    assert rel_err(b[0][0], a[0][0]) < 1e-12
    assert rel_err(b[0][1], a[0][1]) < 1e-12
    assert a[1] == b[1] and a[1] > 0
    assert pos_err(b[2], a[2]) < 1e-12
    assert rel_err(b[3], a[3]) < 1e-12
    assert rel_err(b[4], a[4]) < 1e-12


def test_matrix_cores_refresh_like_the_vector_units_mixed():
    """The mixed form (inverse stored in single precision) on the benchmark
    system: from the same positions, log|psi| after the rebuild agrees to
    1e-9 relative (it is FP64 in both), and the local energies read through
    the two single-precision inverses to 1e-6 relative."""
    spec = _twisted122()
    nw = 64
    pos = _pos(spec, nw, 6)
    out = []
    for kind in (0, 1):
        x = gpu_walkers(spec, nw, MIXED)
        try:
            x.set_refresh_kernel(kind)
            x.set_positions(pos)
            x.evaluate()
            out.append((x.log_psi().copy(), x.local_energy()[0].copy()))
        finally:
            x.close()
    assert rel_err(out[1][0], out[0][0]) < 1e-9
    assert rel_err(out[1][1], out[0][1]) < 1e-6


def test_matrix_cores_refresh_matches_the_oracle():
    """Real 37-orbital blocks (N=74 plane waves, the k_refresh_mfma<double,
    false> instantiation) and complex 61-orbital blocks against the oracle:
    evaluate, two sweeps from one tape, refresh_inverse over a walker range;
    identical accepts, log|psi| and E_L to 1e-10."""
    for spec, rs in ((_plane74(), 40.0), (_twisted122(), 30.0)):
        nw = 12
        pos = _pos(spec, nw, 9)
        tape = oracle_tape(13, spec.n_particles, nw, 2)
        g, o = gpu_walkers(spec, nw), oracle_walkers(spec, nw)
        g.set_refresh_kernel(1)
        for x in (g, o):
            x.set_positions(pos)
            x.evaluate()
        assert rel_err(g.log_psi(), o.log_psi()) < TOL
        assert g.vmc_sweeps(2, (0.5 * rs) ** 2, tape) == o.vmc_sweeps(2, (0.5 * rs) ** 2, tape)
        for x in (g, o):
            x.refresh_inverse(3, 10)
        assert rel_err(g.log_psi(), o.log_psi()) < TOL
        assert rel_err(g.local_energy()[0], o.local_energy()[0]) < TOL
        g.close()


def test_matrix_cores_report_a_singular_block():
    """Two same-spin particles on one spot make a 37-orbital block singular:
    the matrix-core rebuild flags the walker, and the status check raises the
    reference's runtime_error (slater_determinant.cpp:927-928)."""
    spec = _plane74()
    nw = 4
    pos = _pos(spec, nw, 2).reshape(spec.n_particles, 2, nw)
    pos[1, :, 2] = pos[0, :, 2]   # walker 2: particle 1 onto particle 0
    g = gpu_walkers(spec, nw)
    try:
        g.set_refresh_kernel(1)
        g.set_positions(pos.reshape(-1))
        with pytest.raises(RuntimeError, match="singular"):
            g.evaluate()
            g.check_status()
    finally:
        g.close()
