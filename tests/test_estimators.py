"""Pair correlation g(r) and structure factor S(k) (alea_amd.estimators over
alea_add_gofr / alea_add_sofk / alea_estimators_*).

The oracle's restatement is pinned against the reference's own
PairCorrelation and StructureFactor (src/estimators/pair_correlation.cpp,
structure_factor.cpp, compiled into oracle/_ref): the same positions and
weights, several blocks, every finalized series equal to the reference's;
and against closed forms (one pair at a known distance lands in its bin;
one particle has |rho_k|^2 = 1, so S(k) = 1 and rho_k = e^{i k.r}).  On the
GPU the HIP kernels must give the oracle's series: bit for bit with unit
weights, to 1e-12 with DMC weights (the device sums walkers in chunks).
"""
import ctypes as C
import math

import numpy as np
import pytest

from alea_amd.estimators import Estimators
from alea_amd.system import Spec, hexagonal_lattice, square_lattice
from backends import RefWalkers, oracle_walkers, ref_lib


def _spec(lattice, sizes, name):
    s = Spec(n_particles=sum(sizes), lattice=lattice, name=name)
    first = 0
    for i, n in enumerate(sizes):
        s.groups.append((first, first + n, -1.0 if i % 2 == 0 else 1.0, 1.0))
        first += n
    return s


SYSTEMS = [
    pytest.param(_spec(hexagonal_lattice(10, 1.5), (6, 4), "hex"), dict(gofr=(3.0, 24), sofk=9.0), id="hex 6+4"),
    pytest.param(_spec(square_lattice(9, 2.0), (9,), "square"), dict(gofr=(2.9, 17), sofk=6.0), id="square 9"),
    pytest.param(_spec(None, (3, 2), "open"), dict(gofr=(2.0, 20)), id="open 3+2"),
]


def _positions(spec, nw, rng):
    n = spec.n_particles
    if spec.lattice is None:
        return rng.uniform(-1.5, 1.5, size=n * 2 * nw)
    a = np.array(spec.lattice).reshape(2, 2)
    f = rng.uniform(-0.5, 0.5, size=(n, 2, nw))
    return np.einsum("pin,id->pdn", f, a).reshape(-1)


def _run(w, spec, kw, pos_sets, weight_sets, blocks, steps):
    est = Estimators(w, **kw)
    m = 0
    for _ in range(blocks):
        est.start_block()
        for _ in range(steps):
            w.set_positions(pos_sets[m])
            if weight_sets is not None:
                w.set_weights(weight_sets[m])
            est.accumulate(use_weights=weight_sets is not None)
            m += 1
        est.stop_block()
    return est.finalize()


def _ref_series(spec, kw, pos_sets, weight_sets, blocks, steps, nw):
    lib = ref_lib()
    r = RefWalkers(spec, nw)
    try:
        D = C.POINTER(C.c_double)
        f = lib.ref_estimators
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, D, D, C.c_char_p, C.c_int,
                      D, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
        pos = np.ascontiguousarray(np.concatenate(pos_sets))
        wts = np.ascontiguousarray(np.concatenate(weight_sets) if weight_sets is not None
                                   else np.ones(blocks * steps * nw))
        names = C.create_string_buffer(1 << 16)
        vals = np.empty(1 << 18)
        lens = np.zeros(4096, dtype=np.int32)
        ns = C.c_int()
        rmax, nb = kw.get("gofr", (0.0, 0))
        rc = f(r.h, rmax, nb, kw.get("sofk", 0.0), blocks, steps, pos.ctypes.data_as(D), wts.ctypes.data_as(D),
               names, len(names), vals.ctypes.data_as(D), vals.size, lens.ctypes.data_as(C.POINTER(C.c_int)),
               lens.size, C.byref(ns))
        assert rc == 0, lib.ref_last_error()
        out, off = {}, 0
        for name, n in zip(names.value.decode().split("\n"), lens[:ns.value]):
            out[name] = list(vals[off:off + n])
            off += n
        return out
    finally:
        r.close()


def _compare(got, want, rtol=0.0, atol=0.0):
    assert set(got) == set(want), (sorted(set(got) ^ set(want)))
    for name in want:
        a, b = np.array(got[name]), np.array(want[name])
        assert a.shape == b.shape, name
        if rtol == 0.0 and atol == 0.0:
            assert np.array_equal(a, b), (name, np.abs(a - b).max())
        else:
            scale = max(1.0, float(np.abs(b).max()))
            assert np.abs(a - b).max() <= atol + rtol * scale, (name, np.abs(a - b).max())


def _inputs(spec, nw, blocks, steps, weighted, seed):
    rng = np.random.default_rng(seed)
    pos = [_positions(spec, nw, rng) for _ in range(blocks * steps)]
    wts = [rng.uniform(0.2, 2.0, size=nw) for _ in range(blocks * steps)] if weighted else None
    return pos, wts


@pytest.mark.parametrize("weighted", [False, True], ids=["vmc", "dmc"])
@pytest.mark.parametrize("spec,kw", SYSTEMS)
def test_oracle_matches_reference_estimators(spec, kw, weighted):
    nw, blocks, steps = 7, 3, 2
    pos, wts = _inputs(spec, nw, blocks, steps, weighted, 5)
    want = _ref_series(spec, kw, pos, wts, blocks, steps, nw)
    w = oracle_walkers(spec, nw)
    try:
        got = _run(w, spec, kw, pos, wts, blocks, steps)
    finally:
        w.close()
    # gofr is bit for bit; S(k) too (the same phase ladder), kept to 1e-14
    # should a libm differ in the last bit
    _compare({k: v for k, v in got.items() if k.startswith("gofr")},
             {k: v for k, v in want.items() if k.startswith("gofr")})
    _compare({k: v for k, v in got.items() if not k.startswith("gofr")},
             {k: v for k, v in want.items() if not k.startswith("gofr")}, rtol=1e-14, atol=1e-14)


def test_oracle_closed_forms():
    # two particles 1.23 apart in a large square cell: one count in bin
    # int(1.23 / 0.1) of the one function, g = 1 / (1 pair * shell / area)
    spec = _spec(square_lattice(2, 10.0), (2,), "pair")
    w = oracle_walkers(spec, 1)
    try:
        est = Estimators(w, gofr=(5.0, 50), sofk=1.5)
        est.start_block()
        x = np.array([0.3, -0.2, 0.3 + 1.23, -0.2])   # SoA [p][d][w], nw = 1
        w.set_positions(x)
        est.accumulate()
        est.stop_block()
        out = est.finalize()
    finally:
        w.close()
    g = np.array(out["gofr_g0_g0"])
    assert np.count_nonzero(g) == 1 and np.argmax(g) == 12
    area = square_lattice(2, 10.0)[0] ** 2
    shell = math.pi * (1.3 ** 2 - 1.2 ** 2)
    assert math.isclose(g[12], area / shell, rel_tol=1e-12)
    # S(k) of two particles: |e^{ik.r1} + e^{ik.r2}|^2 / 2 = 1 + cos(k.(r1 - r2))
    kx, ky = np.array(out["sofk_kx"]), np.array(out["sofk_ky"])
    s_vec = 1.0 + np.cos(kx * (-1.23) + ky * 0.0)
    for q, sq in enumerate(out["sofk_g0_g0"]):
        members = np.array(out["sofk_shell"]) == q
        assert math.isclose(sq, s_vec[members].mean(), rel_tol=1e-12, abs_tol=1e-12)
    rho = np.cos(kx * 0.3 - ky * 0.2) + np.cos(kx * 1.53 - ky * 0.2)
    assert np.allclose(out["rhok_g0_re"], rho, atol=1e-12)


def test_refusals():
    spec = _spec(square_lattice(4, 1.0), (4,), "small")
    w = oracle_walkers(spec, 2)
    try:
        with pytest.raises(RuntimeError, match="Wigner-Seitz"):
            Estimators(w, gofr=(5.0, 10))
        with pytest.raises(RuntimeError, match="shortest reciprocal"):
            Estimators(w, sofk=0.1)
    finally:
        w.close()
    w = oracle_walkers(_spec(None, (2,), "open"), 2)
    try:
        with pytest.raises(RuntimeError, match="periodic"):
            Estimators(w, sofk=1.0)
    finally:
        w.close()


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [False, True], ids=["vmc", "dmc"])
@pytest.mark.parametrize("spec,kw", SYSTEMS)
def test_gpu_estimators_match_oracle(spec, kw, weighted):
    from alea_amd.walkers import gpu_walkers
    nw, blocks, steps = 37, 3, 2
    pos, wts = _inputs(spec, nw, blocks, steps, weighted, 9)
    print(f"estimators {spec.name} weighted={weighted}: oracle", flush=True)
    o = oracle_walkers(spec, nw)
    try:
        want = _run(o, spec, kw, pos, wts, blocks, steps)
    finally:
        o.close()
    print(f"estimators {spec.name} weighted={weighted}: device", flush=True)
    g = gpu_walkers(spec, nw)
    try:
        got = _run(g, spec, kw, pos, wts, blocks, steps)
    finally:
        g.close()
    gofr = {k: v for k, v in want.items() if k.startswith("gofr")}
    if weighted:
        _compare({k: got[k] for k in gofr}, gofr, rtol=1e-12)
    else:   # integer counts: exact
        _compare({k: got[k] for k in gofr}, gofr)
    rest = {k: v for k, v in want.items() if not k.startswith("gofr")}
    _compare({k: got[k] for k in rest}, rest, rtol=1e-11, atol=1e-11)


@pytest.mark.gpu
def test_gpu_estimators_deterministic_at_full_size():
    """The N=122 workload's 65,536 walkers: two measurements of the same
    walkers give the same bits (the chunked reduction does not depend on
    scheduling), and g(r) counts every pair below rmax exactly once."""
    from alea_amd.system import heg_spec, uniform_positions
    from alea_amd.walkers import gpu_walkers, MIXED
    spec = heg_spec(61, 61, 30.0, twist=(0.21, -0.37))
    nw = 65536
    pos = uniform_positions(spec, nw, np.random.default_rng(3))
    print("estimators: N=122, 65536 walkers", flush=True)
    w = gpu_walkers(spec, nw, MIXED)
    try:
        w.set_positions(pos)
        rws = 0.5 * spec.lattice[0]
        est = Estimators(w, gofr=(rws, 100), sofk=0.05)
        sums = []
        for _ in range(2):
            est.start_block()
            est.accumulate()
            g = np.zeros(3 * 100)
            s = np.zeros(3 * est.nk)
            ws = np.zeros(2)
            w.api("estimators_read", w.h, g.ctypes.data_as(C.POINTER(C.c_double)),
                  s.ctypes.data_as(C.POINTER(C.c_double)), None, None, ws.ctypes.data_as(C.POINTER(C.c_double)))
            sums.append((g.copy(), s.copy(), ws.copy()))
        assert all(np.array_equal(a, b) for a, b in zip(sums[0], sums[1]))
        g, s, ws = sums[0]
        assert ws[0] == nw
        # every count is an integer, and the total is at most the pair count
        assert np.array_equal(g, np.round(g)) and g.sum() <= nw * 122 * 121 / 2
        # in the disk r < L/2 of area pi L^2/4 a uniform ensemble holds a
        # fraction pi/4 of all pairs (to a few parts in 1e4 at this size)
        assert abs(g.sum() / (nw * 122 * 121 / 2) - math.pi / 4) < 2e-3
    finally:
        w.close()
