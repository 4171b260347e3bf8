"""Observables measured on the device: the pair correlation g(r), the
structure factor S(k) with the densities rho_k, and the one-body density.

The device holds each block's running sums (alea_add_gofr / alea_add_sofk,
alea_estimators_start_block / _accumulate / _read); this module does what
the reference does on the host with them:

  stop_block   pair_correlation.cpp:112-128 (the histogram over the block's
               weight sum), structure_factor.cpp:113-141 (S per shell: the sum
               over the shell's vectors of acc / W / sqrt(N_a N_b) / count;
               rho over W)
               density.cpp:120-170 (the grid over W V_cell, streamed into
               per-cell Welford means and second moments)
  finalize     pair_correlation.cpp:140-190 (bin centres; g = block mean /
               (pairs * shell area [/ cell area])), structure_factor.cpp:
               169-221 (series per group pair and group), with
               numerics::block_mean_error over the blocks

Series names follow the reference's RunResult keys: gofr_r,
gofr_<A>_<B>[_error], sofk_k, sofk_shell_count, sofk_<A>_<B>[_error],
sofk_kx, sofk_ky, sofk_shell, rhok_<A>_re/_im[_error], density_edges,
density_<A>[_error] (density.cpp:220-266).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._capi import ptr
from .config import cell_vectors
from .statistics import block_mean_error


class Estimators:
    """g(r), S(k) and/or the density of one Walkers batch (product or oracle
    library)."""

    def __init__(self, walkers, gofr=None, sofk=None, group_names=None, density=None):
        """gofr: (rmax, n_bins) or None; sofk: kmax or None; density:
        (rmax, n_bins) or (rmax, n_bins, per_particle) or None (rmax unused in
        a periodic cell); group_names: one name per group (default
        g<first>)."""
        self.w = walkers
        api, h = walkers.api, walkers.h
        spec = walkers.spec
        self.groups = [(g[0], g[1]) for g in spec.groups]
        self.names = group_names or [f"g{first}" for first, _ in self.groups]
        G = len(self.groups)
        self.G, self.npg = G, G * (G + 1) // 2
        self.lattice = spec.lattice
        self.gofr = gofr is not None
        self.sofk = sofk is not None
        if self.gofr:
            self.rmax, self.n_bins = float(gofr[0]), int(gofr[1])
            api("add_gofr", h, self.rmax, self.n_bins)
            self.delta = self.rmax / self.n_bins
        if self.sofk:
            n = C.c_int()
            api("add_sofk", h, float(sofk), C.byref(n))
            self.nk = n.value
            self.kx, self.ky = np.empty(self.nk), np.empty(self.nk)
            self.shell = np.empty(self.nk, dtype=np.int32)
            api("get_sofk_vectors", h, ptr(self.kx), ptr(self.ky), ptr(self.shell, C.POINTER(C.c_int)))
            self.n_shells = int(self.shell.max()) + 1
            self.shell_count = np.bincount(self.shell, minlength=self.n_shells)
            first = [int(np.argmax(self.shell == q)) for q in range(self.n_shells)]
            self.shell_k = [math.hypot(self.kx[i], self.ky[i]) for i in first]
        self.density = density is not None
        if self.density:
            self.d_rmax, self.d_bins = float(density[0]), int(density[1])
            per_particle = len(density) > 2 and bool(density[2])
            api("add_density", h, self.d_rmax, self.d_bins, 1 if per_particle else 0)
            self.d_cells = self.d_bins * self.d_bins
            # one grid per group, or per particle named <group>_p<i> with i
            # local to the group (density.cpp:42-51)
            if per_particle:
                self.d_names = [f"{self.names[g]}_p{p - first}" for g, (first, last) in enumerate(self.groups)
                                for p in range(first, last)]
            else:
                self.d_names = list(self.names)
            if self.lattice is not None:
                a = cell_vectors(self.lattice)
                self.d_volume = abs(a[0] * a[3] - a[1] * a[2]) / float(self.d_cells)
            else:
                delta = 2.0 * self.d_rmax / float(self.d_bins)
                self.d_volume = 1.0 * delta * delta
            n = len(self.d_names) * self.d_cells
            self.d_mean, self.d_m2, self.d_blocks = [0.0] * n, [0.0] * n, 0
        self.hist_g, self.hist_s, self.hist_re, self.hist_im = [], [], [], []

    # ── per block ────────────────────────────────────────────────────────────
    def start_block(self):
        self.w.api("estimators_start_block", self.w.h)

    def accumulate(self, use_weights=False):
        """One measurement: weight 1 per walker (VMC) or the DMC weights."""
        self.w.api("estimators_accumulate", self.w.h, 1 if use_weights else 0)

    def stop_block(self):
        if self.density:
            self._stop_density()
        if not (self.gofr or self.sofk):
            return
        G, npg = self.G, self.npg
        g = np.zeros(npg * self.n_bins) if self.gofr else None
        s = np.zeros(npg * self.nk) if self.sofk else None
        re = np.zeros(G * self.nk) if self.sofk else None
        im = np.zeros(G * self.nk) if self.sofk else None
        wsum = np.zeros(2)
        self.w.api("estimators_read", self.w.h, ptr(g), ptr(s), ptr(re), ptr(im), ptr(wsum))
        if self.gofr:
            blk = [0.0] * (npg * self.n_bins)
            if wsum[0] > 0.0:
                inv_w = 1.0 / wsum[0]
                blk = [float(x) * inv_w for x in g]
            self.hist_g.append(blk)
        if self.sofk:
            ns, nk = self.n_shells, self.nk
            sb = [0.0] * (npg * ns)
            rb = [0.0] * (G * nk)
            ib = [0.0] * (G * nk)
            if wsum[1] > 0.0:
                inv_w = 1.0 / wsum[1]
                pid = 0
                for a in range(G):
                    for b in range(a, G):
                        na = self.groups[a][1] - self.groups[a][0]
                        nb = self.groups[b][1] - self.groups[b][0]
                        norm = 1.0 / math.sqrt(float(na) * float(nb)) if na > 0 and nb > 0 else 0.0
                        for k in range(nk):
                            q = int(self.shell[k])
                            sb[pid * ns + q] += float(s[pid * nk + k]) * inv_w * norm / float(self.shell_count[q])
                        pid += 1
                rb = [float(x) * inv_w for x in re]
                ib = [float(x) * inv_w for x in im]
            self.hist_s.append(sb)
            self.hist_re.append(rb)
            self.hist_im.append(ib)

    def _stop_density(self):
        """density.cpp:120-170: x = counts / (W_block V_cell), then the
        singleton Welford update of each cell's mean and second moment."""
        n = len(self.d_names) * self.d_cells
        h = np.zeros(n)
        wsum = np.zeros(1)
        self.w.api("estimators_read_density", self.w.h, ptr(h), ptr(wsum))
        w = float(wsum[0])
        norm = 1.0 / (w * self.d_volume) if w > 0.0 else 0.0
        self.d_blocks += 1
        if self.d_blocks == 1:
            self.d_mean = [float(x) * norm for x in h]
            self.d_m2 = [0.0] * n
            return
        inv_n = 1.0 / float(self.d_blocks)
        vw = float(self.d_blocks - 1) * inv_n
        for k in range(n):
            x = float(h[k]) * norm
            d = x - self.d_mean[k]
            self.d_mean[k] += d * inv_n
            self.d_m2[k] += d * (d * vw)

    # ── results ──────────────────────────────────────────────────────────────
    def finalize(self) -> dict:
        out = {}
        if self.density and self.d_blocks:
            nb = self.d_bins
            if self.lattice is not None:
                out["density_edges"] = [-0.5 + float(b) / float(nb) for b in range(nb + 1)]
            else:
                delta = 2.0 * self.d_rmax / float(nb)
                out["density_edges"] = [-self.d_rmax + float(b) * delta for b in range(nb + 1)]
            n = float(self.d_blocks)
            err_norm = 1.0 / ((n - 1.0) * n) if self.d_blocks >= 2 else 0.0
            for g, name in enumerate(self.d_names):
                sl = slice(g * self.d_cells, (g + 1) * self.d_cells)
                out[f"density_{name}"] = list(self.d_mean[sl])
                out[f"density_{name}_error"] = [math.sqrt(m * err_norm) for m in self.d_m2[sl]]
        if self.gofr and self.hist_g:
            nb = self.n_bins
            out["gofr_r"] = [(b + 0.5) * self.delta for b in range(nb)]
            area = 0.0
            if self.lattice is not None:
                a = cell_vectors(self.lattice)
                area = abs(a[0] * a[3] - a[1] * a[2])
            pid = 0
            for ga in range(self.G):
                for gb in range(ga, self.G):
                    n_a = self.groups[ga][1] - self.groups[ga][0]
                    n_b = self.groups[gb][1] - self.groups[gb][0]
                    pc = n_a * (n_a - 1) // 2 if ga == gb else n_a * n_b
                    if pc:
                        g, e = [], []
                        for b in range(nb):
                            lo, hi = b * self.delta, (b + 1) * self.delta
                            shell = math.pi * (hi * hi - lo * lo)
                            norm = pc * shell / area if area > 0.0 else pc * shell
                            m, err = block_mean_error([blk[pid * nb + b] for blk in self.hist_g])
                            g.append(m / norm)
                            e.append(err / norm)
                        name = f"gofr_{self.names[ga]}_{self.names[gb]}"
                        out[name] = g
                        out[name + "_error"] = e
                    pid += 1
        if self.sofk and self.hist_s:
            ns, nk = self.n_shells, self.nk

            def emit(name, hist, offset, count):
                mean, err = [], []
                for q in range(count):
                    m, e = block_mean_error([blk[offset + q] for blk in hist])
                    mean.append(m)
                    err.append(e)
                out[name] = mean
                out[name + "_error"] = err

            out["sofk_k"] = list(self.shell_k)
            out["sofk_shell_count"] = [float(c) for c in self.shell_count]
            pid = 0
            for a in range(self.G):
                for b in range(a, self.G):
                    emit(f"sofk_{self.names[a]}_{self.names[b]}", self.hist_s, pid * ns, ns)
                    pid += 1
            out["sofk_kx"] = [float(x) for x in self.kx]
            out["sofk_ky"] = [float(x) for x in self.ky]
            out["sofk_shell"] = [float(x) for x in self.shell]
            for g in range(self.G):
                emit(f"rhok_{self.names[g]}_re", self.hist_re, g * nk, nk)
                emit(f"rhok_{self.names[g]}_im", self.hist_im, g * nk, nk)
        return out
