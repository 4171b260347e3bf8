"""ctypes binding of the C-ABI declared in include/alea_amd.h.

The same table binds any library that exports the C-ABI under another prefix
(the oracle exports it as ``oracle_*``); the product loads ``libalea_amd.so``
with prefix ``alea_``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_P_D = C.POINTER(C.c_double)
_P_I = C.POINTER(C.c_int)
_P_U8 = C.POINTER(C.c_ubyte)
_P_U64 = C.POINTER(C.c_uint64)
_P_LL = C.POINTER(C.c_longlong)
_V = C.c_void_p

# name -> (restype, argtypes); the handle is an opaque pointer.
class AleaParam(C.Structure):
    """alea_param: one free parameter of alea_parameter_derivatives."""
    _fields_ = [("factor", C.c_int), ("index", C.c_int), ("weights", _P_D), ("spline", _P_D),
                ("n_knots", C.c_int), ("r_cut", C.c_double), ("inv_delta", C.c_double), ("limit", C.c_double)]


SIGNATURES = {
    "last_error": (C.c_char_p, []),
    "create": (C.c_int, [C.POINTER(_V), C.c_int, C.c_int, C.c_int, _P_D, C.c_int]),
    "destroy": (C.c_int, [_V]),
    "add_group": (C.c_int, [_V, C.c_int, C.c_int, C.c_double, C.c_double]),
    "add_pade_jastrow": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, C.c_int, _P_D, C.c_int, _P_I, _P_I,
                                   _P_D, _P_D, _P_D, C.c_int, _P_D, C.c_double, C.c_double, C.c_double]),
    "add_drummond_jastrow": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _P_D]),
    "add_nosanow_jastrow": (C.c_int, [_V, C.c_int, C.c_double, _P_D]),
    "add_slater_plane_wave": (C.c_int, [_V, C.c_int, C.c_int, _P_D, _P_U8]),
    "add_slater_gaussian": (C.c_int, [_V, C.c_int, C.c_int, _P_D, _P_D]),
    "add_slater_twisted": (C.c_int, [_V, C.c_int, C.c_int, _P_D]),
    "add_ewald": (C.c_int, [_V, C.c_int, _P_I, _P_I, _P_D, _P_D, _P_D, C.c_int, _P_D, C.c_double,
                            C.c_double, C.c_double, C.c_double]),
    "add_coulomb": (C.c_int, [_V, C.c_int, C.c_int, C.c_double]),
    "add_inverse_power": (C.c_int, [_V, C.c_int, C.c_int, C.c_double, C.c_double]),
    "add_mcmillan_jastrow": (C.c_int, [_V, C.c_int, C.c_int, C.c_double, C.c_int]),
    "set_positions": (C.c_int, [_V, _P_D]),
    "get_positions": (C.c_int, [_V, _P_D]),
    "evaluate": (C.c_int, [_V]),
    "get_log_psi": (C.c_int, [_V, _P_D]),
    "get_gradients": (C.c_int, [_V, _P_D, _P_D]),
    "get_phase_gradients": (C.c_int, [_V, _P_D, _P_D]),
    "local_energy": (C.c_int, [_V, _P_D, _P_D]),
    "ratio": (C.c_int, [_V, C.c_int, _P_D, _P_D, _P_D, _P_D, _P_D]),
    "accept": (C.c_int, [_V, C.c_int, _P_U8]),
    "ratio_range": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, _P_D, _P_D, _P_D, _P_D, _P_D]),
    "accept_range": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, _P_U8]),
    "ratio_range_phase": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, _P_D, _P_D, _P_D, _P_D, _P_D, _P_D, _P_D]),
    "get_derivatives_range": (C.c_int, [_V, C.c_int, C.c_int, _P_D, _P_D, _P_D, _P_D, _P_D]),
    "get_particle_derivatives_range": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, _P_D, _P_D, _P_D, _P_D]),
    "delta_potential": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, _P_D, _P_D]),
    "parameter_derivatives": (C.c_int, [_V, C.POINTER(AleaParam), C.c_int, _P_D]),
    "local_energy_parameter_derivatives": (C.c_int, [_V, C.POINTER(AleaParam), C.c_int, _P_D, _P_D]),
    "vmc_sweeps": (C.c_int, [_V, C.c_int, C.c_double, _P_D, _P_LL]),
    "snapshots_create": (C.c_int, [_V, C.c_int, C.POINTER(C.c_void_p)]),
    "snapshots_destroy": (C.c_int, [_V]),
    "snapshots_record": (C.c_int, [_V, C.c_int]),
    "snapshots_load": (C.c_int, [_V, C.c_int, _V, _P_I, _P_D, C.c_int]),
    "vmc_drift_sweeps": (C.c_int, [_V, C.c_int, C.c_double, _P_D, _P_LL]),
    "dmc_sweeps": (C.c_int, [_V, C.c_int, C.c_double, _P_D, C.c_int, _P_D, _P_D, _P_LL]),
    "permute": (C.c_int, [_V, C.c_int, _P_I]),
    "get_num_walkers": (C.c_int, [_V, _P_I]),
    "dmc_begin": (C.c_int, [_V]),
    "dmc_weights": (C.c_int, [_V, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "dmc_sums": (C.c_int, [_V, _P_D, _P_D, _P_LL]),
    "get_weights": (C.c_int, [_V, _P_D]),
    "set_weights": (C.c_int, [_V, _P_D]),
    "branch": (C.c_int, [_V, _P_U64, C.c_int, _P_I, _P_I]),
    "rebuild_counts": (C.c_int, [_V, _P_LL, _P_I, C.c_int]),
    "set_cancellation_factor": (C.c_int, [_V, C.c_double]),
    "refresh_inverse": (C.c_int, [_V, C.c_int, C.c_int]),
    "check_status": (C.c_int, [_V]),
    "vmc_sweeps_static": (C.c_int, [_V, C.c_int, C.c_double]),
    "draw_tape": (C.c_int, [_V, C.c_int, _P_D]),
    "set_refresh_updates": (C.c_int, [_V, C.c_int]),
    "set_refresh_kernel": (C.c_int, [_V, C.c_int]),
    "sweep_kernel_ms": (C.c_int, [_V, _P_D, C.c_int, _P_I]),
    "add_gofr": (C.c_int, [_V, C.c_double, C.c_int]),
    "add_density": (C.c_int, [_V, C.c_double, C.c_int, C.c_int]),
    "estimators_read_density": (C.c_int, [_V, _P_D, _P_D]),
    "add_sofk": (C.c_int, [_V, C.c_double, _P_I]),
    "get_sofk_vectors": (C.c_int, [_V, _P_D, _P_D, _P_I]),
    "estimators_start_block": (C.c_int, [_V]),
    "estimators_accumulate": (C.c_int, [_V, C.c_int]),
    "estimators_read": (C.c_int, [_V, _P_D, _P_D, _P_D, _P_D, _P_D]),
}


def ptr(a, kind=_P_D):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class CApi:
    """Bound entry points of one library; raises RuntimeError on a nonzero status."""

    def __init__(self, lib: C.CDLL, prefix: str, optional=()):
        self.lib = lib
        self.prefix = prefix
        self.fns = {}
        for name, (res, args) in SIGNATURES.items():
            sym = prefix + name
            if not hasattr(lib, sym):
                if name in optional:
                    continue
                raise AttributeError(f"{lib._name}: missing symbol {sym}")
            fn = getattr(lib, sym)
            fn.restype = res
            fn.argtypes = args
            self.fns[name] = fn

    def has(self, name):
        return name in self.fns

    def error(self) -> str:
        msg = self.fns["last_error"]()
        return msg.decode() if msg else ""

    def __call__(self, name, *args):
        rc = self.fns[name](*args)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}{name} failed ({rc}): {self.error()}")
        return rc


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)
