"""Loaders for the three libraries the tests compare.

  alea    libalea_amd.so        the HIP product (GPU tests only)
  oracle  oracle/liboracle.so   the FP64 CPU restatement (the checker)
  ref     oracle/_ref/libalea_ref.so   the reference's own sources, built
          only where /root/reference exists; tests that need it skip elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from alea_amd._capi import CApi, f64, i32, ptr, u8, _P_I, _P_U8, _P_U64
from alea_amd.system import Spec
from alea_amd.walkers import Walkers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libalea_ref.so")

_oracle = None
_fixture_cells = None


def fixture_tables(lattice):
    """One cell's breakup tables as the reference's own set-up code wrote
    them (tests/golden/cells.json, by tests/golden/make_tables.py):
    {"ewald": ..., "pade": {pade_key: ...}}, or None for a cell with no tables
    (no Coulomb term)."""
    global _fixture_cells
    if _fixture_cells is None:
        import json
        with open(os.path.join(ROOT, "tests", "golden", "cells.json")) as f:
            _fixture_cells = json.load(f)["cells"]
    return _fixture_cells.get(",".join(f"{x:.12g}" for x in lattice))


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            pytest.fail("oracle/liboracle.so is not built: run __graft_entry__.build()")
        lib = C.CDLL(ORACLE_SO)
        _oracle = (lib, CApi(lib, "oracle_", optional=("rebuild_counts", "set_cancellation_factor", "set_refresh_updates", "set_refresh_kernel",
                                                      "sweep_kernel_ms",
                                                      "vmc_sweeps_static", "draw_tape")))
        for name, res, args in [
            ("oracle_stream_seed", C.c_uint64, [C.c_uint64] * 3),
            ("oracle_xoshiro_state_u64", C.c_int, [_P_U64, C.c_int, _P_U64]),
            ("oracle_xoshiro_draws", C.c_int, [C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            ("oracle_sweep_tape", C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_double)]),
            ("oracle_device_tape", C.c_int, [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                             C.POINTER(C.c_double)]),
        ]:
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
    return _oracle


def oracle_walkers(spec: Spec, nw: int) -> Walkers:
    return Walkers(oracle_lib()[1], spec, nw)


def pos_err(a, b):
    """Largest difference of two position arrays over their largest
    coordinate: a coordinate's own size depends on where the origin sits, so
    positions are compared on the cell's scale."""
    return rel_err(a, b, rel_floor=1.0)


def oracle_device_tape(seed: int, stream: int, np_: int, nw: int, n_sweeps: int):
    """The per-walker device streams' numbers, restated on the CPU."""
    lib = oracle_lib()[0]
    out = np.empty(n_sweeps * np_ * 3 * nw)
    lib.oracle_device_tape(seed, stream, np_, nw, n_sweeps, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def oracle_tape(seed: int, np_: int, nw: int, n_sweeps: int):
    lib = oracle_lib()[0]
    out = np.empty(n_sweeps * np_ * 3 * nw)
    lib.oracle_sweep_tape(seed, np_, 2, nw, n_sweeps, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


_ref = None


def ref_lib():
    """The reference build; skips the calling test where it is absent."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            pytest.skip("oracle/_ref is not built here (needs /root/reference)")
        lib = C.CDLL(REF_SO)
        lib.ref_last_error.restype = C.c_char_p
        D = C.POINTER(C.c_double)
        sigs = {
            "ref_stream_seed": (C.c_uint64, [C.c_uint64] * 3),
            "ref_xoshiro_state_u64": (C.c_int, [_P_U64, C.c_int, _P_U64]),
            "ref_xoshiro_seed_u64": (C.c_int, [C.c_uint64, C.c_int, _P_U64]),
            "ref_xoshiro_draws": (C.c_int, [C.c_uint64, C.c_int, C.c_int, D]),
            "ref_sweep_tape": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, D]),
            "ref_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, D]),
            "ref_destroy": (C.c_int, [C.c_void_p]),
            "ref_add_group": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]),
            "ref_add_pade_jastrow": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, D, C.c_int]),
            "ref_add_drummond_jastrow": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                                   C.c_double, C.c_int, D]),
            "ref_add_nosanow_jastrow": (C.c_int, [C.c_void_p, C.c_int, C.c_double, D]),
            "ref_add_ewald": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
            "ref_add_coulomb": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
            "ref_set_positions": (C.c_int, [C.c_void_p, D]),
            "ref_evaluate": (C.c_int, [C.c_void_p]),
            "ref_get_positions": (C.c_int, [C.c_void_p, D]),
            "ref_get_log_psi": (C.c_int, [C.c_void_p, D]),
            "ref_get_gradients": (C.c_int, [C.c_void_p, D, D]),
            "ref_local_energy": (C.c_int, [C.c_void_p, D]),
            "ref_ratio": (C.c_int, [C.c_void_p, C.c_int, D, D, D, D]),
            "ref_accept": (C.c_int, [C.c_void_p, C.c_int, _P_U8]),
            "ref_vmc_sweeps": (C.c_int, [C.c_void_p, C.c_int, C.c_double, D, C.POINTER(C.c_longlong)]),
            "ref_vmc_drift_sweeps": (C.c_int, [C.c_void_p, C.c_int, C.c_double, D, C.POINTER(C.c_longlong)]),
            "ref_branch": (C.c_int, [D, C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), D]),
            "ref_popctl_run": (C.c_int, [C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                         C.c_int, C.c_double, C.c_double, C.c_int, D, D,
                                         C.POINTER(C.c_int), C.c_int, D, D, D, C.POINTER(C.c_int), D, D]),
            "ref_robust_spread": (C.c_double, [D, C.c_int]),
            "ref_add_inverse_power": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]),
            "ref_delta_potential": (C.c_int, [C.c_void_p, C.c_int, D, D]),
            "ref_set_free": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_ubyte)]),
            "ref_parameter_derivatives": (C.c_int, [C.c_void_p, D, C.c_int, C.POINTER(C.c_int)]),
            "ref_local_energy_parameter_derivatives": (C.c_int, [C.c_void_p, D, C.c_int, C.POINTER(C.c_int)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _ref = lib
    return _ref


class RefWalkers:
    """The reference's own objects behind the Walkers interface (Jastrow,
    Nosanow, Ewald and Coulomb factors; no Slater: Eigen is absent)."""

    def __init__(self, spec: Spec, nw: int, ewald_tol: float | None = None, free_flags=None):
        lib = ref_lib()
        self.lib, self.spec, self.nw, self.np = lib, spec, nw, spec.n_particles
        h = C.c_void_p()
        lat = None if spec.lattice is None else f64(spec.lattice)
        self._chk(lib.ref_create(C.byref(h), nw, self.np, 2, ptr(lat)))
        self.h = h
        if free_flags is not None:
            # free flags of the factors' parameters in the order they are
            # added (Padé coefficients, Drummond coefficients, Nosanow alpha)
            fl = u8(free_flags)
            self._chk(lib.ref_set_free(h, len(fl), ptr(fl, _P_U8)))
        for first, last, q, im in spec.groups:
            lib.ref_add_group(h, first, last, q, im)
        for j in spec.pade:
            c = f64(j["coeffs"])
            self._chk(lib.ref_add_pade_jastrow(h, j["ga"], j["gb"], j["num"], j["den"], ptr(c),
                                               1 if j.get("image") else 0))
        for d in spec.drummond:
            c = f64(d["coeffs"])
            self._chk(lib.ref_add_drummond_jastrow(h, d["ga"], d["gb"], d["C"], d["gamma"], d["L"],
                                                   len(c), ptr(c)))
        for n in spec.nosanow:
            st = f64(n["sites"])
            self._chk(lib.ref_add_nosanow_jastrow(h, n["group"], n["alpha"], ptr(st)))
        if spec.slater_pw or spec.slater_twisted or spec.slater_gauss:
            raise ValueError("the reference build has no Slater determinant (Eigen is absent)")
        if spec.ewald:
            self._chk(lib.ref_add_ewald(h, ewald_tol or spec.ewald.get("tol", 1e-6), 1.0))
        for ga, gb, eps in spec.coulomb:
            self._chk(lib.ref_add_coulomb(h, ga, gb, eps))
        for ga, gb, c, p in spec.inverse_power:
            self._chk(lib.ref_add_inverse_power(h, ga, gb, c, p))
        if spec.mcmillan:
            raise ValueError("the reference has no McMillan Jastrow")

    def _chk(self, rc):
        if rc:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def set_positions(self, pos):
        p = f64(pos)
        self._chk(self.lib.ref_set_positions(self.h, ptr(p)))

    def positions(self):
        out = np.empty(self.np * 2 * self.nw)
        self._chk(self.lib.ref_get_positions(self.h, ptr(out)))
        return out

    def evaluate(self):
        self._chk(self.lib.ref_evaluate(self.h))

    def log_psi(self):
        out = np.empty(self.nw)
        self._chk(self.lib.ref_get_log_psi(self.h, ptr(out)))
        return out

    def gradients(self):
        g = np.empty(self.np * 2 * self.nw)
        l = np.empty_like(g)
        self._chk(self.lib.ref_get_gradients(self.h, ptr(g), ptr(l)))
        return g, l

    def local_energy(self):
        re = np.empty(self.nw)
        self._chk(self.lib.ref_local_energy(self.h, ptr(re)))
        return re, np.zeros(self.nw)

    def ratio(self, p, new_pos):
        npos = f64(new_pos)
        lr, ng, nl = np.empty(self.nw), np.empty(2 * self.nw), np.empty(2 * self.nw)
        self._chk(self.lib.ref_ratio(self.h, p, ptr(npos), ptr(lr), ptr(ng), ptr(nl)))
        return lr, np.zeros(self.nw), ng, nl

    def accept(self, p, mask):
        m = u8(mask)
        self._chk(self.lib.ref_accept(self.h, p, ptr(m, _P_U8)))

    def parameter_derivatives(self):
        n = C.c_int()
        cap = 64 * self.nw
        out = np.zeros(cap)
        self._chk(self.lib.ref_parameter_derivatives(self.h, ptr(out), cap, C.byref(n)))
        return out[:n.value * self.nw].reshape(n.value, self.nw)

    def local_energy_parameter_derivatives(self):
        n = C.c_int()
        cap = 64 * self.nw
        out = np.zeros(cap)
        self._chk(self.lib.ref_local_energy_parameter_derivatives(self.h, ptr(out), cap, C.byref(n)))
        return out[:n.value * self.nw].reshape(n.value, self.nw)

    def delta_potential(self, p, new_pos):
        npos = f64(new_pos)
        out = np.empty(self.nw)
        self._chk(self.lib.ref_delta_potential(self.h, p, ptr(npos), ptr(out)))
        return out

    def vmc_sweeps(self, n_sweeps, tau, tape, drift=False):
        acc = C.c_longlong()
        t = f64(tape)
        f = self.lib.ref_vmc_drift_sweeps if drift else self.lib.ref_vmc_sweeps
        self._chk(f(self.h, n_sweeps, tau, ptr(t), C.byref(acc)))
        return acc.value

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_err(a, b, floor=1e-300, rel_floor=1e-3):
    """Largest elementwise relative difference of a from b: each component
    is held relative to itself, with a floor of rel_floor times the array's
    largest magnitude for components that are themselves the small result
    of a cancellation (a gradient component near zero is the sum of terms of
    the array's size, so its rounding is on that scale)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    if b.size == 0:
        return 0.0
    big = max(float(np.max(np.abs(b))), floor)
    den = np.maximum(np.abs(b), rel_floor * big)
    return float(np.max(np.abs(a - b) / den))
