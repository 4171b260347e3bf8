"""alea_amd: MI355X-native batched VMC/DMC walker sweep.

The compute path is libalea_amd.so (HIP kernels for gfx950 behind the C-ABI
in include/alea_amd.h).  Importing this package never builds anything;
load() raises if the library is missing, so nothing can silently fall back to
a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

from ._capi import CApi

HERE = os.path.dirname(os.path.abspath(__file__))
# ALEA_AMD_LIB overrides the library path (tuning builds); default in-tree.
LIB_PATH = os.environ.get("ALEA_AMD_LIB", os.path.join(HERE, "libalea_amd.so"))

_api = None


def load() -> CApi:
    """The bound C-ABI of libalea_amd.so (raises if it is not built)."""
    global _api
    if _api is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m alea_amd.build "
                               "(or __graft_entry__.build()) before using the GPU path")
        lib = C.CDLL(LIB_PATH)
        api = CApi(lib, "alea_")
        fn = lib.alea_seed
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        api.fns["seed"] = fn
        for name, args in (("synchronize", [C.c_void_p]),
                           ("last_sweep_ms", [C.c_void_p, C.POINTER(C.c_double)]),
                           ("set_refresh_interval", [C.c_void_p, C.c_int]),
                           ("set_sweep_graphs", [C.c_void_p, C.c_int]),
                           ("query", [C.c_void_p, C.POINTER(C.c_int)])):
            f = getattr(lib, "alea_" + name)
            f.restype = C.c_int
            f.argtypes = args
            api.fns[name] = f
        f = lib.alea_vmc_sweeps_fused
        f.restype = C.c_int
        f.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_double)]
        api.fns["vmc_sweeps_fused"] = f
        _api = api
    return _api
