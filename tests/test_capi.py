"""The C-ABI library loads and exports every entry point include/alea_amd.h
declares (no compute call: this runs without a GPU), and the oracle exports
the same set under its own prefix."""
import ctypes as C
import os
import re

import pytest

from backends import ORACLE_SO, ROOT

HEADER = os.path.join(ROOT, "include", "alea_amd.h")
LIB = os.path.join(ROOT, "alea_amd", "libalea_amd.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(alea_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "alea_create" in names and "alea_vmc_sweeps" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("alea_amd/libalea_amd.so is not built: run __graft_entry__.build()")
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_the_same_entry_points():
    lib = C.CDLL(ORACLE_SO)
    # the oracle has no device RNG to seed, no stream to time or poll, and
    # rebuilds on the reference's own schedule (every 64 updates), which it
    # does not count
    missing = [n for n in declared() if n not in ("alea_seed", "alea_synchronize", "alea_last_sweep_ms",
                                                 "alea_set_refresh_interval", "alea_set_sweep_graphs", "alea_vmc_sweeps_fused",
                                                 "alea_query", "alea_rebuild_counts", "alea_set_cancellation_factor",
                                                 "alea_set_refresh_updates", "alea_set_refresh_kernel", "alea_sweep_kernel_ms", "alea_vmc_sweeps_static",
                                                 "alea_draw_tape")
               and not hasattr(lib, n.replace("alea_", "oracle_", 1))]
    assert not missing, missing


def test_refusals_without_a_device():
    """Invalid input is refused before any device work, with the reference's
    error category (invalid_argument -> ALEA_ERR_INVALID)."""
    lib = C.CDLL(LIB)
    lib.alea_last_error.restype = C.c_char_p
    h = C.c_void_p()
    assert lib.alea_create(C.byref(h), 0, 4, 2, None, 0) == 1
    assert b"empty" in lib.alea_last_error()
    assert lib.alea_create(C.byref(h), 4, 4, 3, None, 0) == 1
    assert lib.alea_create(C.byref(h), 4, 1000, 2, None, 0) == 1
