"""One-body density (alea_amd.estimators over alea_add_density /
alea_estimators_read_density).

The oracle's restatement is pinned against the reference's own Density
(src/estimators/density.cpp, compiled into oracle/_ref): the same positions
and weights over several blocks, every finalized series equal bit for bit;
and against closed forms (a particle at a known fractional position lands in
its cell; in a periodic cell the density integrates to the group's count).
On the GPU the HIP kernel must give the oracle's series: bit for bit with
unit weights, to 1e-12 with DMC weights (the device sums walkers in chunks).
"""
import ctypes as C

import numpy as np
import pytest

from alea_amd.estimators import Estimators
from alea_amd.system import hexagonal_lattice, square_lattice
from backends import RefWalkers, oracle_walkers, ref_lib
from test_estimators import _compare, _inputs, _run, _spec

SYSTEMS = [
    pytest.param(_spec(hexagonal_lattice(10, 1.5), (6, 4), "hex"), dict(density=(0.0, 12)), id="hex 6+4"),
    pytest.param(_spec(square_lattice(9, 2.0), (9,), "square"), dict(density=(0.0, 16)), id="square 9"),
    # the open window [-1.2, 1.2)^2 inside positions drawn on [-1.5, 1.5)^2:
    # some particles fall outside and drop
    pytest.param(_spec(None, (3, 2), "open"), dict(density=(1.2, 10)), id="open 3+2"),
    # one grid per particle, named <group>_p<i> (density.cpp:42-51)
    pytest.param(_spec(hexagonal_lattice(10, 1.5), (6, 4), "hex"), dict(density=(0.0, 8, True)),
                 id="hex 6+4 per particle"),
]


def _ref_density(spec, kw, pos_sets, weight_sets, blocks, steps, nw):
    lib = ref_lib()
    r = RefWalkers(spec, nw)
    try:
        D = C.POINTER(C.c_double)
        f = lib.ref_density
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, D, D, C.c_char_p, C.c_int, D,
                      C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
        pos = np.ascontiguousarray(np.concatenate(pos_sets))
        wts = np.ascontiguousarray(np.concatenate(weight_sets) if weight_sets is not None
                                   else np.ones(blocks * steps * nw))
        names = C.create_string_buffer(1 << 16)
        vals = np.empty(1 << 18)
        lens = np.zeros(4096, dtype=np.int32)
        ns = C.c_int()
        rmax, nb = kw["density"][:2]
        per_particle = 1 if len(kw["density"]) > 2 and kw["density"][2] else 0
        rc = f(r.h, rmax, nb, per_particle, blocks, steps, pos.ctypes.data_as(D), wts.ctypes.data_as(D), names, len(names),
               vals.ctypes.data_as(D), vals.size, lens.ctypes.data_as(C.POINTER(C.c_int)), lens.size, C.byref(ns))
        assert rc == 0, lib.ref_last_error()
        out, off = {}, 0
        for name, n in zip(names.value.decode().split("\n"), lens[:ns.value]):
            out[name] = list(vals[off:off + n])
            off += n
        return out
    finally:
        r.close()


def _density_series(d):
    # the grid and the per-group densities; density_lattice is the cell's
    # vectors, which the spec carries
    return {k: v for k, v in d.items() if k == "density_edges" or (k.startswith("density_") and k != "density_lattice")}


@pytest.mark.parametrize("weighted", [False, True], ids=["vmc", "dmc"])
@pytest.mark.parametrize("spec,kw", SYSTEMS)
def test_oracle_matches_reference_density(spec, kw, weighted):
    nw, blocks, steps = 7, 3, 2
    pos, wts = _inputs(spec, nw, blocks, steps, weighted, 5)
    want = _density_series(_ref_density(spec, kw, pos, wts, blocks, steps, nw))
    w = oracle_walkers(spec, nw)
    try:
        got = _run(w, spec, kw, pos, wts, blocks, steps)
    finally:
        w.close()
    assert want and any(k.endswith("_error") for k in want)
    _compare(got, want)   # bit for bit


def test_density_closed_forms():
    """A particle at fractional position f lands in cell int((f + 1/2) n)
    per axis; a periodic density integrates to the group's count."""
    lat = square_lattice(4, 1.0)
    spec = _spec(lat, (4,), "sq")
    L = lat[0]
    fr = np.array([[-0.43, -0.41], [-0.07, 0.33], [0.23, -0.17], [0.47, 0.04]])   # inside their cells
    pos = np.ascontiguousarray((fr * L)[:, :, None].reshape(-1))       # nw = 1
    w = oracle_walkers(spec, 1)
    try:
        est = Estimators(w, density=(0.0, 10))
        est.start_block()
        w.set_positions(pos)
        est.accumulate()
        h = np.zeros(100)
        ws = np.zeros(1)
        w.api("estimators_read_density", w.h, h.ctypes.data_as(C.POINTER(C.c_double)),
              ws.ctypes.data_as(C.POINTER(C.c_double)))
        want = np.zeros(100)
        for f0, f1 in fr:
            want[int((f0 + 0.5) * 10) * 10 + int((f1 + 0.5) * 10)] += 1.0
        assert np.array_equal(h, want) and ws[0] == 1.0
        est.stop_block()
        rho = np.array(est.finalize()["density_g0"])
        assert abs(rho.sum() * (L * L / 100) - 4.0) < 1e-12
    finally:
        w.close()


def _out_of_order(make):
    """Per-particle grids are numbered in group declaration order
    (density.cpp:46-56): with the groups declared out of particle order the
    library refuses them, and keeps the per-group grids."""
    spec = _spec(square_lattice(9, 2.0), (5, 4), "square")
    spec.groups = spec.groups[::-1]
    w = make(spec, 2)
    try:
        with pytest.raises(RuntimeError, match="ascending"):
            w.api("add_density", w.h, 0.0, 8, 1)
        w.api("add_density", w.h, 0.0, 8, 0)
    finally:
        w.close()


def test_density_refusals():
    _out_of_order(oracle_walkers)
    spec = _spec(square_lattice(9, 2.0), (9,), "square")
    w = oracle_walkers(spec, 2)
    try:
        with pytest.raises(RuntimeError):
            w.api("add_density", w.h, 0.0, 0, 0)
    finally:
        w.close()
    spec = _spec(None, (3, 2), "open")
    w = oracle_walkers(spec, 2)
    try:
        with pytest.raises(RuntimeError):
            w.api("add_density", w.h, 0.0, 8, 0)     # an open cell needs a window
    finally:
        w.close()


@pytest.mark.gpu
def test_gpu_density_refuses_out_of_order_groups():
    from alea_amd.walkers import gpu_walkers
    _out_of_order(gpu_walkers)


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [False, True], ids=["vmc", "dmc"])
@pytest.mark.parametrize("spec,kw", SYSTEMS)
def test_gpu_density_matches_oracle(spec, kw, weighted):
    from alea_amd.walkers import gpu_walkers
    nw, blocks, steps = 37, 3, 2
    pos, wts = _inputs(spec, nw, blocks, steps, weighted, 9)
    print(f"density {spec.name} weighted={weighted}: oracle", flush=True)
    o = oracle_walkers(spec, nw)
    try:
        want = _run(o, spec, kw, pos, wts, blocks, steps)
    finally:
        o.close()
    print(f"density {spec.name} weighted={weighted}: device", flush=True)
    g = gpu_walkers(spec, nw)
    try:
        got = _run(g, spec, kw, pos, wts, blocks, steps)
    finally:
        g.close()
    if weighted:
        _compare(got, want, rtol=1e-12, atol=1e-12)
    else:
        _compare(got, want)


@pytest.mark.gpu
def test_gpu_density_with_gofr_at_full_size():
    """The N=122 workload's 65,536 walkers with g(r) and the density on one
    accumulate: the density's counts are integers that add up to every
    particle once, and g(r) reads the same as without the density."""
    from alea_amd.system import heg_spec, uniform_positions
    from alea_amd.walkers import MIXED, gpu_walkers
    spec = heg_spec(61, 61, 30.0, twist=(0.21, -0.37))
    nw = 65536
    pos = uniform_positions(spec, nw, np.random.default_rng(3))
    print("density: N=122, 65536 walkers", flush=True)
    P = C.POINTER(C.c_double)
    g_alone = None
    for with_density in (False, True):
        w = gpu_walkers(spec, nw, MIXED)
        try:
            w.set_positions(pos)
            kw = dict(gofr=(0.5 * spec.lattice[0], 100))
            if with_density:
                kw["density"] = (0.0, 32)
            est = Estimators(w, **kw)
            est.start_block()
            est.accumulate()
            g = np.zeros(300)
            ws = np.zeros(2)
            w.api("estimators_read", w.h, g.ctypes.data_as(P), None, None, None, ws.ctypes.data_as(P))
            if with_density:
                assert np.array_equal(g, g_alone)
                h = np.zeros(2 * 32 * 32)
                wd = np.zeros(1)
                w.api("estimators_read_density", w.h, h.ctypes.data_as(P), wd.ctypes.data_as(P))
                assert np.array_equal(h, np.round(h)) and h.sum() == nw * 122 and wd[0] == nw
                assert h[:1024].sum() == nw * 61
            else:
                g_alone = g.copy()
        finally:
            w.close()
