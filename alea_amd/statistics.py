"""Block statistics of the energy series (host side).

  block_mean_error   numerics/statistics.cpp:18-34
  reblock            numerics/statistics.cpp:52-193 (Flyvbjerg-Petersen, with
                     the reference's plateau rule: the first level whose step
                     to the next is within the combined error-on-error,
                     among levels with >= 16 blocks, falling back to >= 4)
  reblock_chains     numerics/statistics.cpp:195-230 (K equal-length chains,
                     pooled level by level)
  robust_spread      numerics/statistics.cpp:297-315 (in alea_amd.dmc)

Sums run left to right in series order, as the reference's loops do.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

K_MIN_BLOCKS = 16
K_MIN_EFFECTIVE_BLOCKS = 50     # qmc/reblock_diagnostics.h:28 kMinEffectiveBlocks


@dataclass
class ReblockResult:
    mean: float = 0.0
    naive_error: float = 0.0
    error: float = 0.0
    tau_int: float = 0.0
    block_length: int = 1
    level: int = 0
    plateau_found: bool = False
    selected_level: int = 0
    selected_n_blocks: int = 0
    error_on_error: float = 0.0


def _sum(xs) -> float:
    s = 0.0
    for v in xs:
        s += v
    return s


def block_mean_error(block_means):
    """(mean, standard error of the mean) of decorrelated block means."""
    x = [float(v) for v in block_means]
    m = len(x)
    if m == 0:
        return 0.0, 0.0
    mean = _sum(x) / m
    if m < 2:
        return mean, 0.0
    sq = 0.0
    for v in x:
        sq += (v - mean) * (v - mean)
    return mean, math.sqrt(sq / (m - 1) / m)


def _table(series):
    """Rows (level, n_blocks, sem, sem_err) of the pairwise coarse-graining."""
    cur = [float(v) for v in series]
    rows, level = [], 0
    while len(cur) >= 2:
        n = len(cur)
        mean = _sum(cur) / n
        sq = 0.0
        for v in cur:
            d = v - mean
            sq += d * d
        sem = math.sqrt(sq / (n - 1) / n)
        rows.append((level, n, sem, sem / math.sqrt(2.0 * (n - 1))))
        if len(cur) % 2:
            cur.pop()
        cur = [0.5 * (cur[i] + cur[i + 1]) for i in range(0, len(cur), 2)]
        level += 1
    return rows


def _select_plateau(table):
    cand = [r for r in table if r[1] >= K_MIN_BLOCKS]
    if len(cand) < 4:
        cand = [r for r in table if r[1] >= 4]
    if not cand:
        return table[-1][0], False
    for a, b in zip(cand, cand[1:]):
        if b[2] - a[2] <= a[3] + b[3]:
            return a[0], True
    return cand[-1][0], False


def _plateau_error(table, onset):
    for min_blocks in (K_MIN_BLOCKS, 4):
        best, sel = None, None
        for r in table:
            if r[0] < onset or r[1] < min_blocks:
                continue
            if best is None or r[2] > best:
                sel = r
            best = r[2] if best is None else max(best, r[2])
        if best is not None:
            return best, sel
    for r in table:
        if r[0] == onset:
            return r[2], r
    return (table[0][2], table[0]) if table else (0.0, None)


def _assemble(res: ReblockResult, table):
    res.naive_error = table[0][2]
    onset, res.plateau_found = _select_plateau(table)
    res.level = onset
    res.error, sel = _plateau_error(table, onset)
    if sel is not None:
        res.selected_level, res.selected_n_blocks, res.error_on_error = sel[0], sel[1], sel[3]
    if res.naive_error > 0.0 and res.error > 0.0:
        s = (res.error / res.naive_error) ** 2
        res.tau_int = max(0.0, 0.5 * (s - 1.0))
        res.block_length = max(1, int(math.ceil(s)))
    return res


def reblock(series) -> ReblockResult:
    x = [float(v) for v in series]
    res = ReblockResult()
    if not x:
        return res
    res.mean = _sum(x) / len(x)
    res.selected_n_blocks = len(x)
    if len(x) < 2:
        return res
    return _assemble(res, _table(x))


def reblock_diagnostics(name: str, res: ReblockResult, n_samples: int,
                        min_effective_blocks: int = K_MIN_EFFECTIVE_BLOCKS) -> dict:
    """add_reblock_diagnostics (qmc/reblock_diagnostics.h:30-66): the
    reliability figures written beside a reblocked mean."""
    eff = n_samples / (1.0 + 2.0 * res.tau_int)
    reliable = res.plateau_found and res.selected_n_blocks >= K_MIN_BLOCKS and eff >= min_effective_blocks
    return {name + "_effective_blocks": eff, name + "_tau_int": res.tau_int,
            name + "_error_naive": res.naive_error, name + "_reblock_plateau": 1.0 if res.plateau_found else 0.0,
            name + "_error_reliable": 1.0 if reliable else 0.0, name + "_reblock_level": float(res.selected_level),
            name + "_reblock_n_blocks": float(res.selected_n_blocks), name + "_reblock_onset_level": float(res.level),
            name + "_error_on_error": res.error_on_error}


def reblock_chains(chains, n_chains: int) -> ReblockResult:
    """chains: [chain * length + sample], n_chains equal-length chains."""
    x = [float(v) for v in chains]
    if n_chains <= 0 or len(x) % n_chains:
        raise ValueError("reblock_chains: the series does not split into equal-length chains")
    if n_chains == 1:
        return reblock(x)
    res = ReblockResult()
    if not x:
        return res
    res.mean = _sum(x) / len(x)
    res.selected_n_blocks = len(x)
    length = len(x) // n_chains
    if length < 2:
        return res
    per = [_table(x[i * length:(i + 1) * length]) for i in range(n_chains)]
    pooled = []
    for t in range(len(per[0])):
        sem2, nb, dof = 0.0, 0, 0
        for tab in per:
            sem2 += tab[t][2] * tab[t][2]
            nb += tab[t][1]
            dof += tab[t][1] - 1
        sem = math.sqrt(sem2) / n_chains
        pooled.append((t, nb, sem, sem / math.sqrt(2.0 * dof)))
    return _assemble(res, pooled)
