"""End-to-end runs of a reference config on the GPU path: the `qmc:` stages in
order, as the reference's main loop runs them (main.cpp:62-90), each stage
starting where the previous one ended (carry_over_walkers,
src/qmc/app.cpp:1066-1098).

  VMC stage   VMC::run, src/qmc/vmc.cpp:157-420: equilibration blocks, then
              accumulation blocks of steps_per_block steps; a step is
              sub_steps Metropolis sweeps and one measurement of E_L over the
              walkers; a block's energy is its steps' walker mean
  DMC stage   DMC::run, src/qmc/dmc.cpp:736-830 through alea_amd.dmc.DMC
              (run_one_step): equilibration blocks refresh the energy spread,
              accumulation blocks give the weighted-mean block energies
  errors      numerics::reblock (alea_amd.statistics) over the block energies

A VMC stage walks by Metropolis moves or, with random_walk: langevin, by
the drift proposal (app.cpp:162).  A DMC stage's forward_walking {lags,
interval} adds the pure estimators (alea_amd.forward_walking): the pure
potential and the stage's estimators again for every lag.

The walkers draw from the device streams (one xoshiro256++ stream per
walker, seeded from the stage seed), so a run reproduces the reference's
statistics, not its trajectories (those are replayed through tapes in the
tests).  The `gofr`, `density` and `sofk` estimators of a stage are measured on the
device (alea_amd.estimators) after each accumulation step, with the DMC
weights before the branch in DMC.  Run settings the GPU path does not hold
(non-bare reweighting, other estimators, walker files, optimizer and PIMC
stages) are refused with a ValueError.
"""
from __future__ import annotations

import math

import numpy as np

from . import config, dmc, statistics
from .estimators import Estimators
from .forward_walking import ForwardWalking
from .walkers import FP64, gpu_walkers, seed as seed_walkers


def _refuse(what: str):
    raise ValueError(f"run: {what} is not on the GPU path")


def _rows(pos_soa: np.ndarray, np_: int, nw: int) -> np.ndarray:
    """SoA [p*2*nw + d*nw + w] -> rows [w][p*2 + d]."""
    return pos_soa.reshape(np_, 2, nw).transpose(2, 0, 1).reshape(nw, np_ * 2)


def _soa(rows: np.ndarray, np_: int) -> np.ndarray:
    nw = rows.shape[0]
    return np.ascontiguousarray(rows.reshape(nw, np_, 2).transpose(1, 2, 0).reshape(-1))


def _seed_rows(rows: np.ndarray, nw: int) -> np.ndarray:
    """seed_walker_rows: walker w takes row w mod n_rows (app.cpp:1066-1075)."""
    return rows[np.arange(nw) % rows.shape[0]]


def _check_stage(st: dict):
    """Every refusal of a stage, before any stage runs."""
    method = st.get("method")
    if method not in ("vmc", "dmc"):
        _refuse(f"method '{method}'")
    for key in ("walkers_in", "walkers_out", "paths_in", "paths_out"):
        if st.get(key):
            _refuse(f"`{key}` of a {method} stage")
    if st.get("forward_walking") and method != "dmc":
        raise ValueError(f"run: {method}: forward_walking belongs to a dmc stage")
    _estimator_args(st)
    for key in ("samples", "accumulation_blocks", "steps_per_block", "sub_steps", "time_step"):
        if not st.get(key):
            raise ValueError(f"run: a {method} stage needs `{key}` > 0")
    # parser.cpp:1184-1197: VMC walks by 'metropolis' or 'langevin' (the
    # drift proposal, app.cpp:162); DMC by 'drift_diffusion' or its synonym
    # 'langevin'
    walk = st.get("random_walk") or ("metropolis" if method == "vmc" else "drift_diffusion")
    if walk not in (("metropolis", "langevin") if method == "vmc" else ("drift_diffusion", "langevin")):
        raise ValueError(f"run: {method}: random_walk '{walk}' is not one the reference accepts")
    if method == "dmc" and st.get("reweighting", "bare") != "bare":
        _refuse(f"reweighting '{st.get('reweighting')}'")
    if method == "dmc":
        _forward_walking(st)


def _forward_walking(st: dict):
    """The stage's forward_walking {lags, interval} as (lag steps, interval),
    or None (parser.cpp:1310-1336, dmc.cpp:102-139): every lag a whole number
    of steps of time_step * sub_steps, strictly increasing; the interval
    dividing steps_per_block; the largest lag within the equilibration."""
    fw = st.get("forward_walking")
    if not fw:
        return None
    if not isinstance(fw, dict) or set(fw) != {"lags", "interval"}:
        raise ValueError("dmc: forward_walking must be a map {lags: [...], interval: n}")
    lags = fw["lags"]
    if not isinstance(lags, (list, tuple)) or not lags:
        raise ValueError("dmc: forward_walking needs `lags`, a nonempty list of imaginary times")
    interval = int(fw["interval"])
    step_time = float(st["time_step"]) * int(st["sub_steps"])
    n_steps = int(st["steps_per_block"])
    if interval < 1 or n_steps % interval != 0:
        raise ValueError(f"dmc: forward_walking interval ({interval}) must be >= 1 and divide steps_per_block "
                         f"({n_steps})")
    out = []
    for i, lag in enumerate(float(x) for x in lags):
        if not (math.isfinite(lag) and lag >= 0.0):
            raise ValueError("dmc: forward_walking lags must be finite and >= 0")
        if i > 0 and not lag > float(lags[i - 1]):
            raise ValueError("dmc: forward_walking lags must be strictly increasing")
        steps = lag / step_time
        whole = round(steps)
        if abs(steps - whole) > 1e-9 * max(1.0, steps):
            raise ValueError(f"dmc: forward_walking lag {lag} is not a whole number of steps of "
                             f"time_step * sub_steps = {step_time}")
        out.append(int(whole))
    eq_steps = max(0, int(st.get("equilibration_blocks", 0))) * n_steps
    if out[-1] > eq_steps:
        raise ValueError(f"dmc: the largest forward_walking lag ({out[-1]} steps) exceeds the equilibration "
                         f"({eq_steps} steps): snapshots start that many steps before accumulation")
    return out, interval


def _estimator_args(st: dict) -> dict:
    """The stage's `estimators:` as Estimators arguments: one `gofr` (rmax,
    n_bins), one `density` (rmax, n_bins) and one `sofk` (kmax) at most
    (parser.cpp:148-171)."""
    args = {}
    for e in st.get("estimators", []) or []:
        kind = e.get("type")
        if kind not in ("gofr", "sofk", "density"):
            _refuse(f"estimator '{kind}'")
        if e.get("name"):
            _refuse(f"a named {kind} estimator")
        if kind in args:
            _refuse(f"a second {kind} estimator in one stage")
        if kind == "gofr":
            args["gofr"] = (float(e.get("rmax", 0.0)), int(e.get("n_bins", 0)))
        elif kind == "density":
            args["density"] = (float(e.get("rmax", 0.0)), int(e.get("n_bins", 0)), bool(e.get("per_particle", False)))
        else:
            args["sofk"] = float(e.get("kmax", 0.0))
    return args


def _estimators(w, st, group_names):
    args = _estimator_args(st)
    return Estimators(w, group_names=group_names, **args) if args else None


def run_vmc(spec, st: dict, rows: np.ndarray, stage_seed: int, precision=FP64, nw=None,
            group_names=None) -> dict:
    _check_stage(st)
    nw = nw or int(st["samples"])
    n = spec.n_particles
    w = gpu_walkers(spec, nw, precision)
    try:
        w.set_positions(_soa(_seed_rows(rows, nw), n))
        w.evaluate()
        seed_walkers(w, stage_seed)
        est = _estimators(w, st, group_names)
        tau = float(st["time_step"])
        drift = st.get("random_walk") == "langevin"
        n_sub, n_steps = int(st["sub_steps"]), int(st["steps_per_block"])
        for _ in range(int(st.get("equilibration_blocks", 0))):
            for _ in range(n_steps):
                w.vmc_sweeps(n_sub, tau, drift=drift)
        blocks, blocks_im = [], []
        e_sum = e2_sum = 0.0
        n_tot = acc = prop = 0
        for _ in range(int(st["accumulation_blocks"])):
            b_sum = b_im = 0.0
            b_n = 0
            if est:
                est.start_block()
            for _ in range(n_steps):
                acc += w.vmc_sweeps(n_sub, tau, drift=drift)
                prop += n_sub * n * nw
                if est:
                    est.accumulate()
                e, e_im = w.local_energy()
                b_sum += float(np.sum(e))
                b_im += float(np.sum(e_im))
                b_n += nw
                e_sum += float(np.sum(e))
                e2_sum += float(np.sum(e * e))
                n_tot += nw
            blocks.append(b_sum / b_n)
            blocks_im.append(b_im / b_n)
            if est:
                est.stop_block()
        rb = statistics.reblock(blocks)
        mean = e_sum / n_tot
        out = dict(method="vmc", walkers=nw, energy=rb.mean, error=rb.error, tau_int=rb.tau_int,
                   variance=e2_sum / n_tot - mean * mean, acceptance=acc / prop,
                   block_energies=blocks)
        if spec.slater_twisted:
            ri = statistics.reblock(blocks_im)
            out.update(energy_imag=ri.mean, energy_imag_error=ri.error)
        if est:
            out["estimators"] = est.finalize()
        out["final_rows"] = _rows(w.positions(), n, nw)
        return out
    finally:
        w.close()


def run_dmc(spec, st: dict, rows: np.ndarray, stage_seed: int, precision=FP64, target=None,
            group_names=None) -> dict:
    _check_stage(st)
    target = target or int(st["samples"])
    n = spec.n_particles
    cap = 2 * target + 1
    w = gpu_walkers(spec, cap, precision)
    fw = None
    try:
        w.set_positions(_soa(_seed_rows(rows, cap), n))
        w.evaluate()
        w.permute(np.arange(target))
        seed_walkers(w, stage_seed, dmc.K_DMC_THREAD_STREAM)
        est = _estimators(w, st, group_names)
        measure = (lambda: est.accumulate(use_weights=True)) if est else None
        cutoff = st.get("branch_cutoff", "universal")
        alpha = st.get("branch_cutoff_alpha")
        n_steps = int(st["steps_per_block"])
        sched = _forward_walking(st)
        if sched is not None:
            # the stage's estimators again for every lag, labelled fw<i>
            # (builders/estimators.cpp:158-195), and the pure potential
            args = _estimator_args(st)
            fw = ForwardWalking(sched[0], float(st["time_step"]) * int(st["sub_steps"]), sched[1],
                                [args or None] * len(sched[0]))
            fw.prepare(w, int(st.get("equilibration_blocks", 0)) * n_steps, group_names)
        d = dmc.DMC(w, float(st["time_step"]), target, n_substeps=int(st["sub_steps"]),
                    generations=float(st.get("population_generations", 10.0)), cutoff=cutoff,
                    alpha=dmc.K_DEFAULT_BRANCH_CUTOFF_ALPHA if alpha is None else float(alpha),
                    seed=stage_seed, forward_walking=fw)
        eq_blocks = []
        for _ in range(int(st.get("equilibration_blocks", 0))):
            be = bw = 0.0
            for _ in range(n_steps):
                e, ws = d.step(measure=False)
                be += e
                bw += ws
            eq_blocks.append(be / bw if bw > 0.0 else 0.0)
            d.refresh_energy_spread()
        d.pc.begin_accumulation()
        blocks, pops = [], []
        acc = att = 0
        for _ in range(int(st["accumulation_blocks"])):
            be = bw = 0.0
            if est:
                est.start_block()
            if fw:
                fw.start_block()
            for _ in range(n_steps):
                att += w.nw * n * int(st["sub_steps"])
                e, ws = d.step(measure=True, on_measure=measure)
                acc += d.history[-1]["n_accepted"]
                be += e
                bw += ws
            blocks.append(be / bw if bw > 0.0 else 0.0)
            pops.append(w.nw)
            if est:
                est.stop_block()
            if fw:
                fw.stop_block()
        rb = statistics.reblock(blocks)
        out = dict(method="dmc", target=target, energy=rb.mean, error=rb.error, tau_int=rb.tau_int,
                   acceptance=acc / att if att else 0.0, mean_population=float(np.mean(pops)),
                   e_trial=d.pc.etrial, block_energies=blocks, equilibration_block_energies=eq_blocks)
        if est:
            out["estimators"] = est.finalize()
        if fw:
            out["forward_walking"] = fw.results()
        out["final_rows"] = _rows(w.positions(), n, w.nw)
        return out
    finally:
        if fw:
            fw.close()
        w.close()


def run(cfg, precision=FP64, seed=None, walkers=None) -> list[dict]:
    """Every `qmc:` stage of a config, in order; `walkers` overrides each
    stage's `samples` (VMC walkers, DMC target)."""
    spec = config.build_spec(cfg)
    stages = list(cfg.get("qmc", []) or [])
    if any(st.get("method") in ("optimizer", "pimc") for st in stages):
        _refuse("an optimizer or PIMC stage")
    for st in stages:
        _check_stage(st)
    base = int(seed if seed is not None else (cfg.get("project", {}) or {}).get("seed", 0))
    names = [g["name"] for g in config.build_groups(cfg)]
    rows, results = None, []
    for i, st in enumerate(stages):
        stage_seed = base + i + 1           # app.cpp:815-817, no optimizer or PIMC stage
        nw = walkers or int(st["samples"])
        if rows is None:
            rows = _rows(config.initial_positions(cfg, nw, stage_seed), spec.n_particles, nw)
        if st.get("method") == "vmc":
            res = run_vmc(spec, st, rows, stage_seed, precision, nw, names)
        else:
            res = run_dmc(spec, st, rows, stage_seed, precision, nw, names)
        rows = res.pop("final_rows")
        res["stage"] = i
        results.append(res)
    return results
