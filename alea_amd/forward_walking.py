"""Pure (forward-walking) estimators of a DMC run: ForwardWalking
(src/qmc/forward_walking.h:60-147, forward_walking.cpp) over a population on
the device.

The configuration sampled at step t is weighted by the weight its
descendants carry at t + T (Casulleras & Boronat, PRB 52, 3654 (1995)).  The
snapshot positions stay on the population's device in a ring
(alea_snapshots_*); the ancestry is integers on the host and follows each
branch's source map; a measurement gathers the ancestors that still have
descendants into a measurement batch (a system with the Hamiltonian's terms
and no wavefunction factor, so its E_L is the potential energy), which the
estimators of the lag then accumulate with the descendant weights.
"""
from __future__ import annotations

import ctypes as C
import copy
import math

import numpy as np

from alea_amd import statistics
from alea_amd._capi import f64, i32, ptr, _P_D, _P_I
from alea_amd.walkers import Walkers


# The result namespaces of the estimators (Estimator::output_namespaces:
# pair_correlation.h:59, structure_factor.h:80, density.h:79).
NAMESPACES = ("gofr", "sofk", "rhok", "density")


def relabel(key: str, label: str) -> str:
    """Estimator::write_results under an output label (estimator.h:139-162):
    the label follows the longest namespace the key equals or extends with
    '_' (gofr_r -> gofr_fw0_r)."""
    match = None
    for ns in NAMESPACES:
        if (key == ns or (key.startswith(ns) and len(key) > len(ns) and key[len(ns)] == "_")) and \
                (match is None or len(ns) > len(match)):
            match = ns
    if match is None:
        raise ValueError(f"result key '{key}' lies outside the estimators' namespaces")
    return match + "_" + label + key[len(match):]


def potential_spec(spec):
    """The spec's particles and Hamiltonian terms without its wavefunction
    factors: a batch of it measures V (Hamiltonian::evaluate_potential)."""
    s = copy.copy(spec)
    s.pade, s.drummond, s.nosanow, s.mcmillan = [], [], [], []
    s.slater_pw, s.slater_twisted, s.slater_gauss = [], [], []
    s.name = (spec.name or "system") + " (potential)"
    return s


class ForwardWalking:
    """forward_walking.h:60-110.  lag_steps: strictly increasing lags in DMC
    steps; step_time: tau * n_substeps (output only); interval: steps between
    snapshots; estimators: per lag a dict of Estimators keyword arguments
    (gofr=, sofk=, density=) or None.  The potential energy is always
    measured."""

    def __init__(self, lag_steps, step_time, interval, estimators=None):
        lag_steps = [int(x) for x in lag_steps]
        if not lag_steps:
            raise ValueError("forward walking needs at least one lag")
        for a, b in zip(lag_steps, lag_steps[1:]):
            if b <= a:
                raise ValueError("forward walking lags must be strictly increasing")
        if interval < 1:
            raise ValueError("forward walking interval must be at least one step")
        if not (math.isfinite(step_time) and step_time > 0.0):
            raise ValueError("forward walking needs a positive step time (tau * sub_steps)")
        estimators = estimators if estimators is not None else [None] * len(lag_steps)
        if len(estimators) != len(lag_steps):
            raise ValueError("forward walking: one estimator set per lag")
        self.lag_steps = lag_steps
        self.step_time = float(step_time)
        self.interval = int(interval)
        self.est_args = list(estimators)
        n = len(lag_steps)
        self.block_weighted = [0.0] * n
        self.block_weight = [0.0] * n
        self.block_potential = [[] for _ in range(n)]
        self.measurements = [0] * n
        self.surviving_sum = [0.0] * n
        self.ring = None
        self.batches, self.est = [], []

    # ── set-up ───────────────────────────────────────────────────────────────
    def prepare(self, population: Walkers, first_measured_step: int, group_names=None):
        """forward_walking.cpp:50-81: the ring of floor(max lag / interval) + 1
        snapshots of the population's capacity on its device, and one
        measurement batch per lag."""
        from alea_amd.estimators import Estimators
        max_lag = self.lag_steps[-1]
        if first_measured_step < max_lag:
            raise ValueError(f"forward walking: the largest lag ({max_lag} steps) exceeds the equilibration "
                             f"({first_measured_step} steps); the first accumulation block would have no "
                             "snapshot that old")
        self.pop = population
        self.first_measured = int(first_measured_step)
        self.first_snapshot = self.first_measured - max_lag
        n_slots = max_lag // self.interval + 1
        self.slots = [dict(step=None, nw=0, ancestor=None) for _ in range(n_slots)]
        h = C.c_void_p()
        population.api("snapshots_create", population.h, n_slots, C.byref(h))
        self.ring = h
        pspec = potential_spec(population.spec)
        for args in self.est_args:
            b = Walkers(population.api, pspec, population.capacity)
            self.batches.append(b)
            self.est.append(Estimators(b, group_names=group_names, **args) if args else None)

    def close(self):
        if self.ring is not None:
            self.pop.api("snapshots_destroy", self.ring)
            self.ring = None
        for b in self.batches:
            b.close()
        self.batches = []

    # ── schedule ────────────────────────────────────────────────────────────
    def _snapshot_step(self, step):
        return step >= self.first_snapshot and (step - self.first_snapshot) % self.interval == 0

    def _slot_of(self, step):
        return ((step - self.first_snapshot) // self.interval) % len(self.slots)

    def observe(self, step: int):
        """forward_walking.cpp:110-127: record a snapshot if one is due, then,
        at an accumulation step, measure every lag whose snapshot is that old.
        Call after the step's moves and reweighting, before branching."""
        if self.ring is None:
            raise RuntimeError("ForwardWalking.observe before prepare")
        if self._snapshot_step(step):
            slot = self._slot_of(step)
            self.pop.api("snapshots_record", self.ring, slot)
            self.slots[slot] = dict(step=step, nw=self.pop.nw, ancestor=np.arange(self.pop.nw, dtype=np.int64))
        if step < self.first_measured:
            return
        weights = None
        for i, lag in enumerate(self.lag_steps):
            taken = step - lag
            if not self._snapshot_step(taken):
                continue
            slot = self._slot_of(taken)
            s = self.slots[slot]
            if s["step"] != taken:
                raise RuntimeError("ForwardWalking: snapshot ring overwritten too early")
            if weights is None:
                weights = self.pop.weights()
            self._measure(i, slot, s, weights)

    def _measure(self, i, slot, snap, weights):
        """forward_walking.cpp:129-173."""
        anc = snap["ancestor"]
        if anc.size != weights.size:
            raise RuntimeError("ForwardWalking: ancestry out of step with the population")
        # D[a] = sum of the current weights over a's descendants, in walker
        # order (np.add.at adds unbuffered, index by index)
        d = np.zeros(snap["nw"])
        np.add.at(d, anc, weights)
        kept = np.flatnonzero(d > 0.0)
        if kept.size == 0:
            return
        kw = d[kept]
        b = self.batches[i]
        k32 = i32(kept)
        kwf = f64(kw)
        self.pop.api("snapshots_load", self.ring, slot, b.h, ptr(k32, _P_I), ptr(kwf, _P_D), int(kept.size))
        b.nw = int(kept.size)
        if self.est[i] is not None:
            self.est[i].accumulate(use_weights=True)
        v = b.local_energy()[0]
        weighted = total = 0.0
        for wc, vc in zip(kw.tolist(), v.tolist()):
            weighted += wc * vc
            total += wc
        self.block_weighted[i] += weighted
        self.block_weight[i] += total
        self.measurements[i] += 1
        self.surviving_sum[i] += kept.size / snap["nw"]

    def branch(self, src):
        """forward_walking.cpp:175-182: new walker c descends from old walker
        src[c]."""
        src = np.asarray(src, dtype=np.int64)
        for s in self.slots:
            if s["step"] is not None:
                s["ancestor"] = s["ancestor"][src]

    def start_block(self):
        n = len(self.lag_steps)
        self.block_weighted = [0.0] * n
        self.block_weight = [0.0] * n
        for e in self.est:
            if e is not None:
                e.start_block()

    def stop_block(self):
        for i in range(len(self.lag_steps)):
            if self.block_weight[i] > 0.0:
                self.block_potential[i].append(self.block_weighted[i] / self.block_weight[i])
        for e in self.est:
            if e is not None:
                e.stop_block()

    # ── results ─────────────────────────────────────────────────────────────
    def results(self) -> dict:
        """forward_walking.cpp:193-223's keys: per lag i, fw<i>_lag_time /
        _lag_steps / _measurements / _surviving_fraction, potential_fw<i>
        with its error and reblocking diagnostics, acc_block_potential_fw<i>
        and the lag's estimator
        series labelled fw<i>; fw_interval and fw_lag_times."""
        out = {}
        for i, lag in enumerate(self.lag_steps):
            tag = f"fw{i}"
            out[tag + "_lag_time"] = lag * self.step_time
            out[tag + "_lag_steps"] = float(lag)
            out[tag + "_measurements"] = float(self.measurements[i])
            m = self.measurements[i]
            out[tag + "_surviving_fraction"] = self.surviving_sum[i] / m if m else 0.0
            chain = self.block_potential[i]
            name = "potential_" + tag
            if chain:
                rb = statistics.reblock(chain)
                out[name] = rb.mean
                out[name + "_error"] = rb.error
                out.update(statistics.reblock_diagnostics(name, rb, len(chain)))
            out["acc_block_" + name] = list(chain)
            if self.est[i] is not None:
                for k, v in self.est[i].finalize().items():
                    out[relabel(k, tag)] = v
        out["fw_interval"] = float(self.interval)
        out["fw_lag_times"] = [lag * self.step_time for lag in self.lag_steps]
        return out
