"""Headline benchmark: walker-moves/s of the batched VMC sweep on the N=122
twisted 2-D electron gas, mixed precision vs plain FP64, on one MI355X.

Workload (the largest of the published configurations that fits one GPU):
  2-D paramagnetic HEG, N = 122 (61 up, 61 down), rs = 30, square cell
  L = sqrt(pi N) rs; complex plane-wave orbitals at a twist drawn uniformly in
  the first Brillouin zone (seed 3); image-summed Padé Jastrow
  u = a r/(1 + r/(2 rs)) with the 2-D Kato cusps (a = 1/3 parallel, 1
  antiparallel); 2-D Ewald; 65,536 walkers; single-particle Metropolis with
  step 0.5 rs and per-walker xoshiro256++ streams.
A step is one sweep: every particle of every walker proposed, ratioed,
accepted or rejected, committed.  One walker-move = one proposed move.

Prints ONE JSON line (the bench contract).  value = walker-moves/s of the mixed
form; the FP64 form is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
N_UP = N_DN = 61
RS = 30.0
TWIST_SEED = 3
POS_SEED = 1


def workload_spec():
    from alea_amd.system import heg_spec
    twist = tuple(np.random.default_rng(TWIST_SEED).uniform(-0.5, 0.5, 2))
    return heg_spec(N_UP, N_DN, RS, twist=twist), twist


def bytes_per_move(word, acceptance, spec):
    """Algorithmic HBM bytes of one walker-move (the reference's state traffic):
    s*[N*dim + N + a*2N + c*Nb + a*2c*Nb^2 + F_J*2*n_kJ*(1+a) + 2*N*dim*(1+a)]
    with c = 2 (complex), F_J = 2 k-space factors touching the moved particle."""
    N, dim, c = spec.n_particles, 2, 2
    nb = N_UP
    n_kj = len(spec.pade[0]["image"]["w"])
    a = acceptance
    words = (N * dim + N + a * 2 * N + c * nb + a * 2 * c * nb * nb
             + 2 * 2 * n_kj * (1 + a) + 2 * N * dim * (1 + a))
    return word * words


EQUIL_SWEEPS = 150   # FP64 VMC sweeps from the uniform start before either form is timed
EQUIL_SEED = 7


def equilibrated_positions(nw, tau, sweeps=EQUIL_SWEEPS):
    """Uniform walkers (seed 1) relaxed by `sweeps` FP64 VMC sweeps: the
    acceptance falls from 0.52 to a plateau of 0.27 and E/N settles by about
    sweep 120 on this workload.  Both forms start from these walkers."""
    from alea_amd.system import uniform_positions
    from alea_amd.walkers import gpu_walkers, seed
    spec, _ = workload_spec()
    w = gpu_walkers(spec, nw, 0)
    try:
        w.set_positions(uniform_positions(spec, nw, np.random.default_rng(POS_SEED)))
        w.evaluate()
        seed(w, EQUIL_SEED)
        for done in range(0, sweeps, 50):
            acc = w.vmc_sweeps(min(50, sweeps - done), tau)
            print(f"bench: equilibration sweep {min(done + 50, sweeps)}/{sweeps}, acceptance "
                  f"{acc / (nw * spec.n_particles * min(50, sweeps - done)):.4f}", file=sys.stderr, flush=True)
        return w.positions()
    finally:
        w.close()


def run_gpu(precision, nw, steps, warmup, tau, pos):
    from alea_amd.walkers import gpu_walkers, seed
    spec, _ = workload_spec()
    w = gpu_walkers(spec, nw, precision)
    try:
        w.set_positions(pos)
        w.evaluate()
        seed(w, 42)
        for _ in range(warmup):
            w.vmc_sweeps(1, tau)
        w.api("synchronize", w.h)
        w.rebuild_counts(reset=True)
        # the timed steps are queued back to back, as a VMC run queues its
        # sweeps (draw, sweep, then the batched refresh of the blocks that
        # reached 64 updates), and read back once: the acceptance count of
        # the same steps; each sweep kernel is bracketed by its own events
        t0 = time.perf_counter()
        acc = w.vmc_sweeps(steps, tau)
        w.api("synchronize", w.h)
        wall = time.perf_counter() - t0
        kernel_ms = w.sweep_kernel_ms()
        triggers, _ = w.rebuild_counts(reset=True)
        acceptance = acc / (nw * spec.n_particles * steps)
        # one rebuild of every block (the batched inverse recompute on the
        # matrix cores), timed on its own
        t0 = time.perf_counter()
        w.refresh_inverse()
        w.api("synchronize", w.h)
        refresh_ms = 1e3 * (time.perf_counter() - t0)
        e, e_im = w.local_energy()
        out = dict(moves_per_s=nw * spec.n_particles * steps / wall, ms_per_step=1e3 * wall / steps,
                   kernel_ms=float(np.mean(kernel_ms)), kernel_ms_min=float(np.min(kernel_ms)),
                   kernel_ms_max=float(np.max(kernel_ms)), acceptance=acceptance, refresh_ms=refresh_ms,
                   rebuild_triggers_per_sweep=[t / steps for t in triggers],
                   e_per_particle=float(np.mean(e)) / spec.n_particles,
                   e_per_particle_err=float(np.std(e) / math.sqrt(nw)) / spec.n_particles,
                   im_e_per_particle=float(np.mean(e_im)) / spec.n_particles)
    finally:
        w.close()
    return out, spec


def _cpu_worker(args):
    nw, tau, seed_ = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from alea_amd.system import uniform_positions
    from backends import oracle_tape, oracle_walkers
    spec, _ = workload_spec()
    o = oracle_walkers(spec, nw)
    o.set_positions(uniform_positions(spec, nw, np.random.default_rng(seed_)))
    o.evaluate()
    tape = oracle_tape(seed_, spec.n_particles, nw, 1)
    t0 = time.perf_counter()
    o.vmc_sweeps(1, tau, tape)
    return nw * spec.n_particles, time.perf_counter() - t0


CPU_SHARE = 16   # host cores one GPU of the box is allotted


def cpu_baseline(tau, walkers_per_core=384):
    """The oracle (the FP64 CPU restatement, which follows the reference's
    full complex re-inversion per accepted move) on the host cores: one
    single-threaded process per core of the GPU's host-core share, each
    sweeping `walkers_per_core` walkers once.  Per-walker CPU cost does not
    depend on the walker count."""
    nproc = os.cpu_count() or 1
    cores = max(1, min(CPU_SHARE, nproc))
    jobs = [(walkers_per_core, tau, 100 + i) for i in range(cores)]
    print(f"bench: CPU baseline, {cores} processes x {walkers_per_core} walkers on the oracle", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    moves = sum(r[0] for r in res)
    sweep_time = max(r[1] for r in res)
    return dict(value=moves / sweep_time, unit="walker-moves/s", cores=cores, nproc=nproc, kind="port",
                sample=f"{cores} processes (the {CPU_SHARE}-core host share of one GPU; nproc {nproc}) x "
                       f"{walkers_per_core} walkers x 1 sweep of the N=122 workload from uniform walkers on "
                       f"the oracle (FP64, complex re-inversion per accept); {wall:.1f} s wall incl. setup")


SERVING = dict(n_jobs=16, rate_per_s=4.0, seed=4, n_blocks=4, sweeps_per_block=10,
               walkers=(512, 8192), sizes=(26, 58, 74, 122, 242), rs=(30.0, 40.0, 60.0),
               dipole_fraction=0.2, dipoles=(256, 1.0))


# the stream's own job size: 20 blocks of 50 sweeps per job
SERVING_FULL = dict(SERVING, n_blocks=20, sweeps_per_block=50)


def serving_leg(precision=1, full=False):
    """The mixed-size request stream: Poisson arrivals of VMC jobs (N, rs,
    twist, walker count drawn from the seed), continuously batched on the
    GPU.  p50/p99 block and request latency, aggregate walker-moves/s.  The
    default is a reduced stream (4 blocks of 10 sweeps per job) that fits
    the bench's few minutes; full=True runs every job at 20 x 50."""
    from alea_amd import serving
    c = SERVING_FULL if full else SERVING
    jobs = serving.poisson_jobs(c["n_jobs"], c["rate_per_s"], seed=c["seed"], sizes=c["sizes"],
                                rs_choices=c["rs"], walkers=c["walkers"], n_blocks=c["n_blocks"],
                                sweeps_per_block=c["sweeps_per_block"], precision=precision,
                                dipole_fraction=c["dipole_fraction"], dipoles=c["dipoles"])
    def progress(j):
        print(f"bench: serving job {j.job_id} (N={j.n_particles}, {j.n_walkers} walkers) done at "
              f"{j.finished:.1f} s", file=sys.stderr, flush=True)
    last = [time.perf_counter()]

    def block_done(j):
        # a line at most every 20 s while blocks finish (the full stream runs minutes)
        if time.perf_counter() - last[0] > 20.0:
            last[0] = time.perf_counter()
            print(f"bench: serving job {j.job_id} block {len(j.block_energy)}/{j.n_blocks}", file=sys.stderr,
                  flush=True)
    st = serving.Server(hbm_budget_bytes=96 << 30, max_active=8, on_finish=progress,
                        on_block=block_done).run(jobs)
    st["config"] = dict(c, reduced=not full,
                        note="each job 20 blocks x 50 sweeps" if full else
                        "reduced stream: each job 4 blocks x 10 sweeps (the stream's jobs are 20 x 50: "
                        "bench.py --serving-full)")
    return st


DMC_CFG = dict(n_up=37, n_dn=37, rs=40.0, target=16384, tau=2.0, steps=500, equilibration=100,
               vmc_sweeps=200, drummond_order=8, drummond_sigma=1e-3, drummond_seed=2, pos_seed=1,
               seed=7, generations=10.0, alpha=2.0, replicas=6)


def dmc_spec(c=DMC_CFG):
    """The N=74, rs=40 paramagnetic HEG of the published DMC case
    (alea_amd.system.heg_drummond_spec)."""
    from alea_amd.system import heg_drummond_spec
    return heg_drummond_spec(c["n_up"], c["n_dn"], c["rs"], order=c["drummond_order"],
                             sigma=c["drummond_sigma"], seed=c["drummond_seed"])


def dmc_run(spec, c, prec, pos_seed, seed):
    """One fixed-node DMC run: uniform walkers from pos_seed, `vmc_sweeps`
    VMC sweeps, then `steps` DMC steps (one drift-diffusion sweep each, bare
    reweighting, integer branching, universal cutoff) on the streams of
    `seed`.  Returns the per-step ensemble E/N after `equilibration`, the
    populations and the walker-moves/s."""
    from alea_amd import dmc
    from alea_amd.system import uniform_positions
    from alea_amd.walkers import gpu_walkers, seed as seed_walkers
    n = spec.n_particles
    cap = 2 * c["target"] + 1
    w = gpu_walkers(spec, cap, prec)
    try:
        w.set_positions(uniform_positions(spec, cap, np.random.default_rng(pos_seed)))
        w.evaluate()
        w.permute(np.arange(c["target"]))
        seed_walkers(w, seed)
        w.vmc_sweeps(c["vmc_sweeps"], (0.5 * c["rs"]) ** 2)
        seed_walkers(w, seed, dmc.K_DMC_THREAD_STREAM)
        d = dmc.DMC(w, c["tau"], c["target"], n_substeps=1, generations=c["generations"],
                    cutoff="universal", alpha=c["alpha"], seed=seed)
        series, pops, moves = [], [], 0
        t0 = time.perf_counter()
        for step in range(c["steps"]):
            moves += w.nw * n
            e_sum, w_sum = d.step(measure=step >= c["equilibration"])
            if step >= c["equilibration"]:
                series.append(e_sum / w_sum / n)
                pops.append(w.nw)
        wall = time.perf_counter() - t0
        return series, pops, moves / wall, (w.nw, w.positions())
    finally:
        w.close()


def same_configuration_energies(spec, nw, pos):
    """E_L/N of the same walkers evaluated in FP64 and in mixed form (fresh
    inverses, no trajectory): the mixed form's local-energy error alone."""
    from alea_amd.walkers import gpu_walkers
    e = []
    for prec in (0, 1):
        w = gpu_walkers(spec, nw, prec)
        try:
            w.set_positions(pos)
            w.evaluate()
            e.append(w.local_energy()[0] / spec.n_particles)
        finally:
            w.close()
    d = e[1] - e[0]
    return dict(walkers=nw, mean=float(np.mean(d)), max_abs=float(np.max(np.abs(d))),
                fp64_e_per_particle=float(np.mean(e[0])))


def _mean_err(xs):
    """Mean of independent replica values and its standard error."""
    xs = np.asarray(xs, dtype=np.float64)
    return float(xs.mean()), float(xs.std(ddof=1) / math.sqrt(len(xs))) if len(xs) > 1 else float("nan")


def dmc_leg(c=DMC_CFG):
    """Fixed-node DMC on the N=74, rs=40 HEG in FP64 and in mixed form.
    `replicas` independent runs, replica r on walker seed pos_seed + r and
    stream seed seed + r; FP64 and mixed share each replica's walkers and
    streams.  E/N of one run is the mean of its steps after `equilibration`
    (with its reblocked error); a form's E/N is the mean over replicas, and
    its error is the spread of the replica means, which stays honest where
    population-control correlations outlast one run's reblocking.  The
    mixed - FP64 difference is taken per replica, then averaged the same way.
    Beside it, the local-energy difference of the two forms on identical
    configurations (replica 0's final FP64 walkers), which no sampling
    noise enters."""
    from alea_amd import statistics
    spec = dmc_spec(c)
    out = {"fp64": {}, "mixed": {}}
    per = {"fp64": [], "mixed": []}
    final = None
    for r in range(c["replicas"]):
        for name, prec in (("fp64", 0), ("mixed", 1)):
            series, pops, rate, last = dmc_run(spec, c, prec, c["pos_seed"] + r, c["seed"] + r)
            if r == 0 and prec == 0:
                final = last
            rb = statistics.reblock(series)
            per[name].append(dict(e_per_particle=rb.mean, reblocked_error=rb.error, tau_int=rb.tau_int,
                                  plateau_found=rb.plateau_found, mean_population=float(np.mean(pops)),
                                  final_population=pops[-1], walker_moves_per_s=rate))
            print(f"bench: DMC replica {r} {name}: E/N {rb.mean:.7f} +- {rb.error:.1e}, "
                  f"final population {pops[-1]}", file=sys.stderr, flush=True)
    for name in per:
        m, e = _mean_err([x["e_per_particle"] for x in per[name]])
        out[name] = dict(e_per_particle=m, error=e,
                         walker_moves_per_s=float(np.mean([x["walker_moves_per_s"] for x in per[name]])),
                         replicas=per[name])
    d, de = _mean_err([a["e_per_particle"] - b["e_per_particle"] for a, b in zip(per["mixed"], per["fp64"])])
    out["mixed_minus_fp64"] = d
    out["error"] = de
    out["within_1e-5"] = abs(d) <= 1e-5
    out["within_1_sigma"] = abs(d) <= de
    # the final DMC walkers of replica 0's FP64 run, evaluated in both forms
    out["same_configuration_difference"] = same_configuration_energies(spec, *final)
    out["config"] = dict(c, system="N=74 (37+37) paramagnetic HEG, rs=40, plane-wave nodes, Pade + Drummond, Ewald")
    return out


CRYSTAL_CFG = dict(rs=60.0, walkers=32768, steps=300, equilibration=60, tau=2.0, vmc_sweeps=20, seed=11,
                   pos_seed=5, spread=0.15, generations=10.0, alpha=2.0, refresh_updates=32)
COUPLED_CFG = dict(walkers=4096, dmc_steps=150, sweeps=200, window=20, seed=23)
COUPLED_CRYSTAL_CFG = dict(walkers=2048, dmc_steps=100, sweeps=100, window=20, seed=29)


def reweighted_shift(forms, n, logw):
    """E/N under the mixed process minus E/N under the FP64 one, from the
    FP64 walkers at the end of a window and each walker's likelihood ratio
    of its window's path (correlated sampling): sum(W E)/sum(W) - mean(E),
    with a jackknife error over the walkers, and the spread of W."""
    e = forms[0].local_energy()[0] / n
    w = np.exp(logw - np.max(logw))
    w /= w.mean()
    m = len(e)
    tot_we, tot_w, tot_e = float(np.sum(w * e)), float(np.sum(w)), float(np.sum(e))
    est = tot_we / tot_w - tot_e / m
    jack = ((tot_we - w * e) / (tot_w - w)) - (tot_e - e) / (m - 1)
    err = math.sqrt((m - 1) / m * float(np.sum((jack - jack.mean()) ** 2)))
    return est, err, float(np.std(w))


def coupled_leg(c=COUPLED_CFG, d=DMC_CFG):
    """The mixed form's effect on DMC sampling, measured without sampling
    noise, on the N=74, rs=40 system of the DMC leg: walkers from a short
    FP64 DMC run, inverses rebuilt every second sweep (as the DMC sweeps
    rebuild a 37-orbital block after 64 updates).  See coupled_run."""
    spec = dmc_spec(d)
    n = spec.n_particles
    run = dict(d, target=c["walkers"], steps=c["dmc_steps"], equilibration=c["dmc_steps"])
    _, _, _, (nw_f, pos_f) = dmc_run(spec, run, 0, d["pos_seed"], d["seed"])
    nw = min(nw_f, c["walkers"])
    pos = pos_f.reshape(n * 2, nw_f)[:, :nw].reshape(-1).copy()
    out = coupled_run(spec, pos, nw, d["tau"], c["sweeps"], c["window"], c["seed"], refresh_moves=2 * n)
    out["config"] = dict(c, tau=d["tau"], system="N=74 (37+37) paramagnetic HEG, rs=40 (the DMC leg's)")
    return out


def coupled_crystal_leg(c=COUPLED_CRYSTAL_CFG, d=CRYSTAL_CFG):
    """coupled_run on the Wigner crystal of the crystal leg: walkers from a
    short FP64 DMC run from the sites, the block's inverse rebuilt every 32
    moves (the config's refresh)."""
    from alea_amd import dmc
    from alea_amd.system import site_positions, wigner_crystal_spec
    from alea_amd.walkers import gpu_walkers, seed as seed_walkers
    spec = wigner_crystal_spec(d["rs"])
    n = spec.n_particles
    a0 = spec.lattice[0] / 16
    target = c["walkers"]
    cap = 2 * target + 1
    w = gpu_walkers(spec, cap, 0)
    try:
        w.set_refresh_updates(d["refresh_updates"])
        w.set_positions(site_positions(spec, cap, np.random.default_rng(d["pos_seed"]), d["spread"] * a0))
        w.evaluate()
        w.permute(np.arange(target))
        seed_walkers(w, d["seed"])
        w.vmc_sweeps(d["vmc_sweeps"], (0.25 * a0) ** 2)
        seed_walkers(w, d["seed"], dmc.K_DMC_THREAD_STREAM)
        dm = dmc.DMC(w, d["tau"], target, n_substeps=1, generations=d["generations"], cutoff="universal",
                     alpha=d["alpha"], seed=d["seed"])
        for _ in range(c["dmc_steps"]):
            dm.step()
        nw_f, pos_f = w.nw, w.positions()
    finally:
        w.close()
    nw = min(nw_f, c["walkers"])
    pos = pos_f.reshape(n * 2, nw_f)[:, :nw].reshape(-1).copy()
    out = coupled_run(spec, pos, nw, d["tau"], c["sweeps"], c["window"], c["seed"],
                      refresh_moves=d["refresh_updates"])
    out["config"] = dict(c, tau=d["tau"], system="polarized Wigner crystal N=128, rs=60 (the crystal leg's)")
    return out


def coupled_run(spec, pos, nw, tau, sweeps, window, seed, refresh_moves):
    """An FP64 and a mixed batch loaded with the same walkers; every move
    proposes the same drift-diffusion step (FP64 drift, shared Gaussians) to
    both, each form computes its own acceptance min(1, |psi'/psi|^2
    G(R<-R')/G(R'<-R)) from its own ratio and drifts at both ends (Langevin
    Green's function, green_function.h:20-26; node crossings rejected), one
    shared uniform decides both, and both batches commit the FP64 decision so
    they stay on one trajectory.  Both inverses are rebuilt every
    `refresh_moves` moves.  Reported: the fraction of decisions the two forms
    take differently and which form accepted, the log-ratio difference
    (mean, median, signed mean, max), the drift difference, the local-energy
    difference on the final walkers, and the shift of E/N that the mixed
    process makes, by reweighting the FP64 path with the mixed process's
    likelihood of each step (its proposal density and its probability of the
    decision taken) over windows of `window` sweeps (longer than the DMC
    energy's correlation time)."""
    from alea_amd.walkers import FP64, MIXED, gpu_walkers
    n = spec.n_particles
    rng = np.random.default_rng(seed)
    forms = [gpu_walkers(spec, nw, FP64), gpu_walkers(spec, nw, MIXED)]
    try:
        for w in forms:
            w.set_positions(pos)
            w.evaluate()
        flips = moves = mixed_only = 0
        signed_sum = 0.0
        dlr_all = []
        dlr_max = dlr_sum = 0.0
        ddrift_max = ddrift_sum = 0.0
        accepted = 0
        logw = np.zeros(nw)
        windows = []
        moved = 0
        for sweep in range(sweeps):
            if sweep % window == 0 and sweep > 0:
                windows.append(reweighted_shift(forms, n, logw))
                logw[:] = 0.0
            for p in range(n):
                if moved % refresh_moves == 0:
                    for w in forms:
                        w.refresh_inverse()
                moved += 1
                old = forms[0].positions().reshape(n, 2, nw)[p]
                grads = [w.gradients()[0].reshape(n, 2, nw)[p] for w in forms]
                step = np.sqrt(tau) * rng.standard_normal((2, nw))
                new = old + tau * grads[0] + step
                u = rng.random(nw)
                acc, prob = [], []
                outs = [w.ratio(p, new.reshape(-1)) for w in forms]
                for (lr, ph, ng, _), v_old in zip(outs, grads):
                    v_new = ng.reshape(2, nw)
                    fwd = np.sum((new - old - tau * v_old) ** 2, axis=0)
                    back = np.sum((old - new - tau * v_new) ** 2, axis=0)
                    log_a = 2.0 * lr + (fwd - back) / (2.0 * tau)
                    crossed = np.cos(ph) < 0.0
                    acc.append((np.log(u) < log_a) & ~crossed)
                    prob.append(np.where(crossed, 0.0, np.exp(np.minimum(log_a, 0.0))))
                # the mixed process's likelihood of this step relative to the
                # FP64 one: its proposal density (its own drift) and its
                # probability of the FP64 decision
                disp = new - old
                log_t = -(np.sum((disp - tau * grads[1]) ** 2, axis=0) -
                          np.sum((disp - tau * grads[0]) ** 2, axis=0)) / (2.0 * tau)
                with np.errstate(divide="ignore", invalid="ignore"):
                    step_w = np.where(acc[0], prob[1] / prob[0], (1.0 - prob[1]) / (1.0 - prob[0]))
                    logw += log_t + np.log(np.maximum(step_w, 0.0))
                flips += int(np.count_nonzero(acc[0] != acc[1]))
                mixed_only += int(np.count_nonzero(acc[1] & ~acc[0]))
                signed_sum += float(np.sum(outs[1][0] - outs[0][0]))
                dlr_all.append(np.abs(outs[1][0] - outs[0][0]).astype(np.float32))
                moves += nw
                accepted += int(np.count_nonzero(acc[0]))
                dlr = np.abs(outs[1][0] - outs[0][0])
                dlr_max = max(dlr_max, float(dlr.max()))
                dlr_sum += float(dlr.sum())
                g64, g32 = outs[0][2].reshape(2, nw), outs[1][2].reshape(2, nw)
                rel = np.linalg.norm(g32 - g64, axis=0) / np.maximum(np.linalg.norm(g64, axis=0), 1e-300)
                ddrift_max = max(ddrift_max, float(rel.max()))
                ddrift_sum += float(rel.sum())
                mask = acc[0].astype(np.uint8)
                for w in forms:
                    w.accept(p, mask)
        windows.append(reweighted_shift(forms, n, logw))
        shift = np.array([x[0] for x in windows])
        shift_err = math.sqrt(sum(x[1] ** 2 for x in windows)) / len(windows)
        e = [w.local_energy()[0] / n for w in forms]
        de = e[1] - e[0]
        return dict(walkers=nw, sweeps=sweeps, moves=moves, acceptance=accepted / moves,
                    decision_flips=flips, flip_fraction=flips / moves,
                    flips_mixed_accepts=mixed_only, flips_fp64_accepts=flips - mixed_only,
                    log_ratio_diff_mean=dlr_sum / moves, log_ratio_diff_max=dlr_max,
                    log_ratio_diff_median=float(np.median(np.concatenate(dlr_all))),
                    log_ratio_diff_signed_mean=signed_sum / moves,
                    drift_rel_diff_mean=ddrift_sum / moves, drift_rel_diff_max=ddrift_max,
                    e_per_particle_diff_mean=float(np.mean(de)), e_per_particle_diff_max=float(np.max(np.abs(de))),
                    sampling_shift=float(shift.mean()), sampling_shift_error=shift_err,
                    sampling_shift_windows=[dict(shift=a, error=b, weight_spread=sw) for a, b, sw in windows])
    finally:
        for w in forms:
            w.close()


def crystal_leg(c=CRYSTAL_CFG):
    """Fixed-node DMC of the polarized Wigner crystal (N = 128 on an 8 x 16
    triangular lattice at rs = 60; one 128-orbital real Gaussian determinant,
    Nosanow, Pade + Drummond, Ewald) in FP64 and in mixed form from the same
    walkers and streams: walkers start at the sites plus N(0, (0.15 a)^2),
    VMC sweeps, then `steps` DMC steps (one drift-diffusion sweep each) at a
    target population of `walkers`.  E/N is the mean of the steps after
    `equilibration` with a reblocked error; walker-moves/s over the DMC
    steps."""
    from alea_amd import dmc, statistics
    from alea_amd.system import site_positions, wigner_crystal_spec
    from alea_amd.walkers import gpu_walkers, seed as seed_walkers
    spec = wigner_crystal_spec(c["rs"])
    n = spec.n_particles
    a0 = spec.lattice[0] / 16
    target = c["walkers"]
    cap = 2 * target + 1
    pos = site_positions(spec, cap, np.random.default_rng(c["pos_seed"]), c["spread"] * a0)
    out = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        print(f"bench: crystal DMC N={n} {name}", file=sys.stderr, flush=True)
        w = gpu_walkers(spec, cap, prec)
        try:
            # the config's inverse refresh every 32 moves: a block is rebuilt
            # once it has taken 32 Sherman-Morrison updates
            w.set_refresh_updates(c["refresh_updates"])
            w.set_positions(pos)
            w.evaluate()
            w.permute(np.arange(target))
            seed_walkers(w, c["seed"])
            w.vmc_sweeps(c["vmc_sweeps"], (0.25 * a0) ** 2)
            seed_walkers(w, c["seed"], dmc.K_DMC_THREAD_STREAM)
            d = dmc.DMC(w, c["tau"], target, n_substeps=1, generations=c["generations"], cutoff="universal",
                        alpha=c["alpha"], seed=c["seed"])
            series, moves = [], 0
            w.api("synchronize", w.h)
            t0 = time.perf_counter()
            for step in range(c["steps"]):
                moves += w.nw * n
                e_sum, w_sum = d.step(measure=step >= c["equilibration"])
                if step >= c["equilibration"]:
                    series.append(e_sum / w_sum / n)
                if step % 100 == 99:
                    print(f"bench: crystal {name} step {step + 1}, population {w.nw}", file=sys.stderr, flush=True)
            wall = time.perf_counter() - t0
            rb = statistics.reblock(series)
            out[name] = dict(e_per_particle=rb.mean, error=rb.error, plateau_found=rb.plateau_found,
                             walker_moves_per_s=moves / wall, wall_s=wall, final_population=w.nw)
        finally:
            w.close()
    out["mixed_over_fp64"] = out["mixed"]["walker_moves_per_s"] / out["fp64"]["walker_moves_per_s"]
    out["mixed_minus_fp64"] = out["mixed"]["e_per_particle"] - out["fp64"]["e_per_particle"]
    out["config"] = dict(c, system="polarized Wigner crystal N=128 (8 x 16 triangular), rs=60, Gaussian orbitals "
                                   "c=0.1/rs^1.5, Nosanow, Pade + Drummond, Ewald")
    return out


def estimators_leg(nw=65536, reps=5):
    """One measurement of g(r) (rmax = L/2, 100 bins), S(k) (|k| <= 3 k_F)
    and the density (32 x 32 cells per group) of the N=122 workload's
    walkers on the device, `reps` times; milliseconds per measurement from
    the wall clock around synchronized batches."""
    from alea_amd.estimators import Estimators
    from alea_amd.system import uniform_positions
    from alea_amd.walkers import gpu_walkers
    spec, _ = workload_spec()
    w = gpu_walkers(spec, nw, 1)
    try:
        w.set_positions(uniform_positions(spec, nw, np.random.default_rng(POS_SEED)))
        L = spec.lattice[0]
        k_f = math.sqrt(4.0 * math.pi * N_UP / (L * L))
        est = Estimators(w, gofr=(0.5 * L, 100), sofk=3.0 * k_f, density=(0.0, 32))
        est.start_block()
        est.accumulate()
        est.stop_block()     # warm-up; reading the sums waits for the stream
        est.start_block()
        t0 = time.perf_counter()
        for _ in range(reps):
            est.accumulate()
        est.stop_block()     # reads the sums back: waits for the stream
        ms = (time.perf_counter() - t0) * 1e3 / reps
        return dict(walkers=nw, n_particles=spec.n_particles, gofr_bins=100, sofk_vectors=est.nk,
                    density_cells=32 * 32, ms_per_measurement=ms)
    finally:
        w.close()


def derivatives_leg(nw=65536, crystal_nw=8192, reps=3):
    """The optimizer's derivatives on the device (d log|psi|/dp and dE_L/dp
    of TrialWavefunction::compute_parameter_derivatives and
    Hamiltonian::local_energy_parameter_derivatives): every Padé coefficient
    of the N=122 workload (image-summed, complex trial) on `nw` walkers, and
    the Wigner crystal's (N=128) Gaussian width, Drummond L and Nosanow alpha
    on `crystal_nw` walkers.  Milliseconds per call from the wall clock
    around each call (which reads the results back); the kernels' own times
    come from a kernel trace of the same command."""
    from alea_amd.system import free_parameters, site_positions, uniform_positions, wigner_crystal_spec
    from alea_amd.walkers import gpu_walkers
    out = {}
    spec, _ = workload_spec()
    free = [(f, i) for f, j in enumerate(spec.pade) for i in range(j["num"] + j["den"])]
    params = free_parameters(spec, free)
    w = gpu_walkers(spec, nw, 0)
    try:
        w.set_positions(uniform_positions(spec, nw, np.random.default_rng(POS_SEED)))
        w.evaluate()
        w.parameter_derivatives(params)
        w.local_energy_parameter_derivatives(params)   # warm-up
        t = []
        for fn in (w.parameter_derivatives, w.local_energy_parameter_derivatives):
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(params)
            t.append((time.perf_counter() - t0) * 1e3 / reps)
        out["heg122"] = dict(walkers=nw, parameters=len(params), dlogpsi_ms=t[0], de_ms=t[1])
    finally:
        w.close()
    cs = wigner_crystal_spec()
    nd = len(cs.drummond[0]["coeffs"])
    params = free_parameters(cs, [(1, nd), (2, 0), (3, 0)])
    a = math.sqrt(cs.lattice[0] * cs.lattice[3] / cs.n_particles)
    w = gpu_walkers(cs, crystal_nw, 0)
    try:
        w.set_positions(site_positions(cs, crystal_nw, np.random.default_rng(POS_SEED), 0.15 * a))
        w.evaluate()
        w.local_energy_parameter_derivatives(params)   # warm-up
        t = []
        for fn in (w.parameter_derivatives, w.local_energy_parameter_derivatives):
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(params)
            t.append((time.perf_counter() - t0) * 1e3 / reps)
        n = cs.n_particles
        out["crystal128"] = dict(walkers=crystal_nw, parameters=len(params), dlogpsi_ms=t[0], de_ms=t[1],
                                 gauss_width_flops_per_walker=4 * n ** 3)
    finally:
        w.close()
    out["cpu_baseline"] = cpu_baseline_derivatives()
    return out


def cpu_baseline_derivatives(heg_walkers=64, crystal_walkers=16):
    """The oracle's d log|psi|/dp and dE_L/dp on one host core, on a sample
    of each system's walkers (the same parameters as derivatives_leg); cost
    per walker does not depend on the walker count, so the figures are ms
    per 1,000 walkers."""
    from alea_amd.system import free_parameters, site_positions, uniform_positions, wigner_crystal_spec
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    from backends import oracle_walkers
    out = dict(kind="port", cores=1)
    spec, _ = workload_spec()
    cs = wigner_crystal_spec()
    a = math.sqrt(cs.lattice[0] * cs.lattice[3] / cs.n_particles)
    nd = len(cs.drummond[0]["coeffs"])
    for name, sp, free, nw, pos in (
            ("heg122", spec, [(f, i) for f, j in enumerate(spec.pade) for i in range(j["num"] + j["den"])],
             heg_walkers, lambda n: uniform_positions(spec, n, np.random.default_rng(POS_SEED))),
            ("crystal128", cs, [(1, nd), (2, 0), (3, 0)], crystal_walkers,
             lambda n: site_positions(cs, n, np.random.default_rng(POS_SEED), 0.15 * a))):
        params = free_parameters(sp, free)
        o = oracle_walkers(sp, nw)
        try:
            o.set_positions(pos(nw))
            o.evaluate()
            t0 = time.perf_counter()
            o.parameter_derivatives(params)
            t1 = time.perf_counter()
            o.local_energy_parameter_derivatives(params)
            t2 = time.perf_counter()
        finally:
            o.close()
        out[name] = dict(walkers=nw, parameters=len(params), dlogpsi_ms_per_1k_walkers=(t1 - t0) * 1e6 / nw,
                         de_ms_per_1k_walkers=(t2 - t1) * 1e6 / nw)
    out["sample"] = (f"{heg_walkers} N=122 walkers (6 Padé parameters) and {crystal_walkers} Wigner-crystal "
                     f"walkers (3 parameters) on the oracle, one thread")
    return out


def copy_calibration(n_bytes, reps=3):
    """A plain device-to-device copy of n_bytes, reps times: the known-size
    reference that scales the FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    import torch
    src = torch.empty(n_bytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--equilibration", type=int, default=EQUIL_SWEEPS,
                    help="FP64 sweeps from the uniform start before either form is timed")
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="time one form only (profiling runs); prints its dict, not the bench line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run a plain copy of BYTES bytes (counter calibration) and exit")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serving-only", action="store_true", help="run the serving leg alone, print its dict")
    ap.add_argument("--serving-full", action="store_true",
                    help="run the serving stream at its full job size (20 blocks x 50 sweeps) alone, print its dict")
    ap.add_argument("--no-dmc", action="store_true", help="skip the DMC accuracy leg")
    ap.add_argument("--dmc-only", action="store_true", help="run the DMC leg alone, print its dict")
    ap.add_argument("--crystal-only", action="store_true",
                    help="run the Wigner-crystal DMC leg (FP64 and mixed) alone, print its dict")
    ap.add_argument("--coupled-only", action="store_true",
                    help="count the DMC decisions the mixed form takes differently from FP64 on one "
                         "shared trajectory, print its dict")
    ap.add_argument("--coupled-crystal-only", action="store_true",
                    help="the same on the Wigner crystal, print its dict")
    ap.add_argument("--estimators-only", action="store_true",
                    help="time one g(r) + S(k) measurement of the workload's walkers, print its dict")
    ap.add_argument("--derivatives-only", action="store_true",
                    help="time the optimizer's parameter derivatives (N=122 Padé, Wigner-crystal width), "
                         "print their dict")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(walkers shard by GPU with no exchange inside a sweep)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.serving_only or args.serving_full:
        print(json.dumps(serving_leg(full=args.serving_full)), flush=True)
        return
    if args.dmc_only:
        print(json.dumps(dmc_leg()), flush=True)
        return
    if args.crystal_only:
        print(json.dumps(crystal_leg()), flush=True)
        return
    if args.coupled_only:
        print(json.dumps(coupled_leg()), flush=True)
        return
    if args.coupled_crystal_only:
        print(json.dumps(coupled_crystal_leg()), flush=True)
        return
    if args.estimators_only:
        print(json.dumps(estimators_leg(args.walkers)), flush=True)
        return
    if args.derivatives_only:
        print(json.dumps(derivatives_leg(args.walkers)), flush=True)
        return
    tau = (0.5 * RS) ** 2   # sigma = sqrt(tau/m) = 0.5 rs
    pos = equilibrated_positions(args.walkers, tau, args.equilibration)
    if args.only:
        out, _ = run_gpu(0 if args.only == "fp64" else 1, args.walkers, args.steps, args.warmup, tau, pos)
        print(json.dumps(dict(form=args.only, **out)), flush=True)
        return
    print("bench: FP64 sweeps", file=sys.stderr, flush=True)
    fp64, spec = run_gpu(0, args.walkers, args.steps, args.warmup, tau, pos)
    print(f"bench: FP64 {fp64['moves_per_s']:.4g} walker-moves/s; mixed sweeps", file=sys.stderr, flush=True)
    mixed, _ = run_gpu(1, args.walkers, args.steps, args.warmup, tau, pos)
    print(f"bench: mixed {mixed['moves_per_s']:.4g} walker-moves/s", file=sys.stderr, flush=True)
    b_move = bytes_per_move(4, mixed["acceptance"], spec)
    moves_per_launch = args.walkers * spec.n_particles
    achieved = b_move * moves_per_launch / (mixed["kernel_ms"] * 1e-3) / 1e9
    traffic = traffic_fp64 = None
    tf = os.path.join(ROOT, "profiles", "sweep_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        if t.get("walkers") == args.walkers and t.get("precision") == "mixed":
            traffic = t.get("hbm_bytes_per_launch")
            traffic_fp64 = t.get("fp64", {}).get("hbm_bytes_per_launch")
    # the FP64 complex sweep streams the inverse through HBM on every accepted
    # move: the survey's byte model is its roofline as it stands
    b_move64 = bytes_per_move(8, fp64["acceptance"], spec)
    achieved64 = b_move64 * moves_per_launch / (fp64["kernel_ms"] * 1e-3) / 1e9
    # the sweep holds the swept block in registers, so HBM does not bound it:
    # its VALU issue time (SQ_INSTS_VALU from rocprofv3, profiles/) against
    # the kernel time measured here
    valu = refresh_roof = None
    vf = os.path.join(ROOT, "profiles", "sweep_counters.json")
    if os.path.exists(vf):
        with open(vf) as f:
            v = json.load(f)
        if v.get("walkers") == args.walkers and v.get("precision") == "mixed":
            valu = {"insts_per_launch": v["valu_insts_per_launch"], "issue_ms": v["valu_issue_ms"],
                    "issue_frac": v["valu_issue_ms"] / mixed["kernel_ms"], "wait_frac": v["wait_fraction"]}
            rm = v.get("refresh_mfma")
            if rm:
                # the full rebuild's matrix-core flops (SQ_INSTS_VALU_MFMA_MOPS_F64,
                # profiles/) over the rebuild timed here
                tf_ = rm["mfma_flops"] / (mixed["refresh_ms"] * 1e-3) / 1e12
                refresh_roof = {"kernel": rm["kernel"], "ms": mixed["refresh_ms"], "bound": "mfma",
                                "achieved": tf_, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": tf_ / FP64_MFMA_PEAK_TFLOPS, "flops": rm["mfma_flops"],
                                "fp64_form_ms": fp64["refresh_ms"]}
    line = {
        "metric": "walker-moves/s, 2D HEG N=122, mixed vs FP64 GPU",
        "value": mixed["moves_per_s"],
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64 (mixed)",
        "data": f"synthetic: uniform initial walkers (seed 1) equilibrated by {args.equilibration} FP64 VMC "
                "sweeps, shared by both forms; twist from seed 3",
        "config": {"workload": "twisted 2D HEG VMC sweep, N=122 (61+61), rs=30, Slater-Jastrow-Ewald",
                   "n_walkers": args.walkers, "n_particles": spec.n_particles,
                   "step": "one sweep (every particle of every walker)", "tau": tau},
        "fp64": {"value": fp64["moves_per_s"], "ms_per_step": fp64["ms_per_step"],
                 "kernel_ms": fp64["kernel_ms"], "kernel_ms_range": [fp64["kernel_ms_min"], fp64["kernel_ms_max"]],
                 "acceptance": fp64["acceptance"], "e_per_particle": fp64["e_per_particle"],
                 "e_per_particle_err": fp64["e_per_particle_err"],
                 "rebuild_triggers_per_sweep": fp64["rebuild_triggers_per_sweep"],
                 "roofline": {"bound": "hbm", "achieved": achieved64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved64 / HBM_PEAK_GBS, "traffic": traffic_fp64,
                              "kernel": "k_sweep<double,true,false,0>", "bytes_per_move": b_move64}},
        "mixed_over_fp64": mixed["moves_per_s"] / fp64["moves_per_s"],
        "acceptance": mixed["acceptance"],
        "e_per_particle": mixed["e_per_particle"],
        "e_per_particle_err": mixed["e_per_particle_err"],
        "rebuild_triggers_per_sweep": mixed["rebuild_triggers_per_sweep"],
        "kernel_ms_range": [mixed["kernel_ms_min"], mixed["kernel_ms_max"]],
        "im_e_per_particle": mixed["im_e_per_particle"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sweep<float,true,false,64>", "kernel_ms": mixed["kernel_ms"],
                     "bytes_per_move": b_move,
                     "model": "achieved = the survey's algorithmic bytes per move (A read and written "
                              "through HBM on every accepted move) x moves / kernel time.  This kernel "
                              "keeps the swept block in registers, so those bytes never reach HBM "
                              "(measured: `traffic` per launch) and frac can pass 1; its binding "
                              "roofline is VALU issue (`valu.issue_frac`) and latency (`valu.wait_frac`)",
                     "valu": valu},
        "refresh_roofline": refresh_roof,
        "serving": None if args.no_serving else serving_leg(),
        "dmc": None if args.no_dmc else dmc_leg(),
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(tau),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
