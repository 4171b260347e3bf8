"""Block statistics (alea_amd.statistics) against the reference's own
numerics::reblock / reblock_chains, compiled into oracle/_ref, and closed forms."""
import ctypes as C
import math

import numpy as np
import pytest

from alea_amd import statistics as st
from backends import ref_lib


def _ref(x, n_chains=0):
    lib = ref_lib()
    f = lib.ref_reblock
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(7)
    iout = np.zeros(3, dtype=np.int32)
    rc = f(a.ctypes.data_as(C.POINTER(C.c_double)), a.size, n_chains, out.ctypes.data_as(C.POINTER(C.c_double)),
           iout.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0
    return out, iout


def _ar1(n, phi, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(n)
    v = 0.0
    for i in range(n):
        v = phi * v + rng.normal()
        x[i] = v - 1.5
    return x


def _same(r, out, iout):
    got = [r.mean, r.naive_error, r.error, r.tau_int, r.error_on_error, float(r.block_length), float(r.level)]
    assert got == list(out), (got, list(out))
    assert [int(r.plateau_found), r.selected_level, r.selected_n_blocks] == list(iout)


@pytest.mark.parametrize("n,phi,seed", [(1000, 0.9, 1), (257, 0.0, 2), (4096, 0.5, 3), (37, 0.3, 4),
                                        (2, 0.0, 5), (1, 0.0, 6), (5, 0.0, 7)])
def test_reblock_matches_reference(n, phi, seed):
    x = _ar1(n, phi, seed)
    _same(st.reblock(x), *_ref(x))


@pytest.mark.parametrize("k,length", [(4, 256), (3, 101), (8, 64), (2, 1)])
def test_reblock_chains_matches_reference(k, length):
    x = np.concatenate([_ar1(length, 0.7, 10 + i) for i in range(k)])
    _same(st.reblock_chains(x, k), *_ref(x, k))


def test_reblock_closed_forms():
    # a constant series has no error at any level
    r = st.reblock([2.5] * 64)
    assert r.mean == 2.5 and r.error == 0.0 and r.naive_error == 0.0
    # level 0 is the textbook standard error of the mean
    x = np.random.default_rng(11).normal(size=512)
    assert math.isclose(st.reblock(x).naive_error, np.std(x, ddof=1) / math.sqrt(x.size), rel_tol=1e-12)
    # a strongly correlated series needs longer blocks: tau_int of AR(1) is
    # phi/(1-phi) = 9 for phi = 0.9 (the estimate is noisy; the order is fixed)
    r = st.reblock(_ar1(1 << 14, 0.9, 12))
    assert 4.0 < r.tau_int < 20.0 and r.block_length > 8
    # chains with K = 1 are reblock itself
    assert st.reblock_chains(x, 1) == st.reblock(x)
    with pytest.raises(ValueError):
        st.reblock_chains(np.zeros(10), 3)


def test_block_mean_error_closed_form():
    m, e = st.block_mean_error([1.0, 2.0, 4.0])
    assert m == 7.0 / 3.0
    assert math.isclose(e, math.sqrt(((1 - m) ** 2 + (2 - m) ** 2 + (4 - m) ** 2) / 2 / 3), rel_tol=1e-15)
    assert st.block_mean_error([3.0]) == (3.0, 0.0)
