"""Walker batch handle over a C-ABI library (include/alea_amd.h).

Mirrors the reference's host-side interfaces at the granularity the C-ABI
exposes them: WalkerSet (positions, log psi, gradients), TrialWavefunction
(evaluate_batch, ratio_batch, apply_move_batch), Hamiltonian
(evaluate_batch) and the drivers' sweeps.  Arrays use the reference's
walker-fastest SoA layouts:

  positions, gradients, laplacians   [p*dim*nw + d*nw + w]
  proposed position, new_grad        [d*nw + w]
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import CApi, f64, i32, ptr, u8, _P_I, _P_U8, _P_U64
from .system import Spec

FP64, MIXED = 0, 1


class Walkers:
    def __init__(self, api: CApi, spec: Spec, nw: int, precision: int = FP64):
        self.api = api
        self.spec = spec
        self.nw = nw
        self.capacity = nw          # alea_create's walker count is the capacity
        self.np = spec.n_particles
        self.dim = 2
        h = C.c_void_p()
        lat = None if spec.lattice is None else f64(spec.lattice)
        api("create", C.byref(h), nw, self.np, 2, ptr(lat), precision)
        self.h = h
        self._keep = []
        try:
            self._register(spec)
        except Exception:
            self.close()
            raise

    def _k(self, a):
        self._keep.append(a)
        return a

    def _register(self, s: Spec):
        api, h = self.api, self.h
        for first, last, q, im in s.groups:
            api("add_group", h, first, last, q, im)
        for j in s.pade:
            c = self._k(f64(j["coeffs"]))
            t = j.get("image")
            if t:
                n1, n2 = self._k(i32(t["n1"])), self._k(i32(t["n2"]))
                kx, ky, w = self._k(f64(t["kx"])), self._k(f64(t["ky"])), self._k(f64(t["w"]))
                sp = self._k(f64(t["spline"]))
                api("add_pade_jastrow", h, j["ga"], j["gb"], j["num"], j["den"], ptr(c), len(w),
                    ptr(n1, _P_I), ptr(n2, _P_I), ptr(kx), ptr(ky), ptr(w), t["n_knots"], ptr(sp),
                    t["r_cut"], t["inv_delta"], t["u_inf"])
            else:
                api("add_pade_jastrow", h, j["ga"], j["gb"], j["num"], j["den"], ptr(c), 0, None,
                    None, None, None, None, 0, None, 0.0, 0.0, 0.0)
        for d in s.drummond:
            c = self._k(f64(d["coeffs"]))
            api("add_drummond_jastrow", h, d["ga"], d["gb"], d["C"], d["gamma"], d["L"], len(c), ptr(c))
        for n in s.nosanow:
            st = self._k(f64(n["sites"]))
            api("add_nosanow_jastrow", h, n["group"], n["alpha"], ptr(st))
        for first, last, k, sn in s.slater_pw:
            kk, ss = self._k(f64(k)), self._k(u8(sn))
            api("add_slater_plane_wave", h, first, last, ptr(kk), ptr(ss, _P_U8))
        for first, last, q in s.slater_twisted:
            qq = self._k(f64(q))
            api("add_slater_twisted", h, first, last, ptr(qq))
        for first, last, cen, al in s.slater_gauss:
            cc, aa = self._k(f64(cen)), self._k(f64(al))
            api("add_slater_gaussian", h, first, last, ptr(cc), ptr(aa))
        if s.ewald:
            e = s.ewald
            n1, n2 = self._k(i32(e["n1"])), self._k(i32(e["n2"]))
            kx, ky, w = self._k(f64(e["kx"])), self._k(f64(e["ky"])), self._k(f64(e["w"]))
            sp = self._k(f64(e["spline"]))
            api("add_ewald", h, len(w), ptr(n1, _P_I), ptr(n2, _P_I), ptr(kx), ptr(ky), ptr(w),
                e["n_knots"], ptr(sp), e["r_cut"], e["inv_delta"], e["v_long_zero"], e["sr_k0"])
        for ga, gb, eps in s.coulomb:
            api("add_coulomb", h, ga, gb, eps)
        for ga, gb, c, p in s.inverse_power:
            api("add_inverse_power", h, ga, gb, c, p)
        for ga, gb, b, n in s.mcmillan:
            api("add_mcmillan_jastrow", h, ga, gb, b, n)

    # ── state ──────────────────────────────────────────────────────────────
    def set_positions(self, pos):
        p = f64(pos)
        assert p.size == self.np * 2 * self.nw
        self.api("set_positions", self.h, ptr(p))

    def positions(self):
        out = np.empty(self.np * 2 * self.nw)
        self.api("get_positions", self.h, ptr(out))
        return out

    def evaluate(self):
        self.api("evaluate", self.h)

    def log_psi(self):
        out = np.empty(self.nw)
        self.api("get_log_psi", self.h, ptr(out))
        return out

    def gradients(self):
        g = np.empty(self.np * 2 * self.nw)
        l = np.empty_like(g)
        self.api("get_gradients", self.h, ptr(g), ptr(l))
        return g, l

    def phase_gradients(self):
        g = np.empty(self.np * 2 * self.nw)
        l = np.empty_like(g)
        self.api("get_phase_gradients", self.h, ptr(g), ptr(l))
        return g, l

    def local_energy(self):
        re = np.empty(self.nw)
        im = np.empty(self.nw)
        self.api("local_energy", self.h, ptr(re), ptr(im))
        return re, im

    # ── per-call path ──────────────────────────────────────────────────────
    def ratio(self, p, new_pos):
        npos = f64(new_pos)
        lr = np.empty(self.nw)
        ph = np.empty(self.nw)
        ng = np.empty(2 * self.nw)
        nl = np.empty(2 * self.nw)
        self.api("ratio", self.h, p, ptr(npos), ptr(lr), ptr(ph), ptr(ng), ptr(nl))
        return lr, ph, ng, nl

    def accept(self, p, mask):
        m = u8(mask)
        self.api("accept", self.h, p, ptr(m, _P_U8))

    # ── per-call path over a walker range (crowds) ─────────────────────────
    def ratio_range(self, p, w_begin, w_end, new_pos, out=None):
        """ratio_batch(ws, p, w_begin, w_end): the range's entries of out
        (log_ratio, phase, new_grad, new_lap; batch stride) are written."""
        npos = f64(new_pos)
        if out is None:
            out = (np.zeros(self.nw), np.zeros(self.nw), np.zeros(2 * self.nw), np.zeros(2 * self.nw))
        lr, ph, ng, nl = out
        self.api("ratio_range", self.h, p, w_begin, w_end, ptr(npos), ptr(lr), ptr(ph), ptr(ng), ptr(nl))
        return out

    @staticmethod
    def _param_array(params):
        from ._capi import AleaParam
        n = len(params)
        arr = (AleaParam * max(n, 1))()
        keep = []
        for i, q in enumerate(params):
            arr[i].factor, arr[i].index = q["factor"], q["index"]
            t = q.get("tables")
            if t:
                w, sp = f64(t["w"]), f64(t["spline"])
                keep += [w, sp]
                arr[i].weights, arr[i].spline = ptr(w), ptr(sp)
                arr[i].n_knots, arr[i].r_cut, arr[i].inv_delta, arr[i].limit = (
                    t["n_knots"], t["r_cut"], t["inv_delta"], t["u_inf"])
        return arr, keep

    def parameter_derivatives(self, params):
        """TrialWavefunction::compute_parameter_derivatives for the free
        parameters `params` (system.free_parameters): [n][nw] of
        d log|psi| / d p."""
        n = len(params)
        arr, keep = self._param_array(params)
        out = np.zeros(max(n, 1) * self.nw)
        self.api("parameter_derivatives", self.h, arr, n, ptr(out))
        return out[:n * self.nw].reshape(n, self.nw)

    def local_energy_parameter_derivatives(self, params):
        """Hamiltonian::local_energy_parameter_derivatives over
        TrialWavefunction::compute_parameter_gradient_derivatives for the free
        parameters `params`: (re, im), each [n][nw], of dE_L / d p (im zero
        for a real trial)."""
        n = len(params)
        arr, keep = self._param_array(params)
        re = np.zeros(max(n, 1) * self.nw)
        im = np.zeros(max(n, 1) * self.nw)
        self.api("local_energy_parameter_derivatives", self.h, arr, n, ptr(re), ptr(im))
        return re[:n * self.nw].reshape(n, self.nw), im[:n * self.nw].reshape(n, self.nw)

    def delta_potential(self, p, new_pos, w_begin=0, w_end=-1):
        """Hamiltonian::delta_potential_particle for every walker of the range:
        V with particle p at new_pos [2][nw] minus V now, nothing committed."""
        npos = np.ascontiguousarray(new_pos, dtype=np.float64).reshape(-1)
        out = np.zeros(self.nw)
        self.api("delta_potential", self.h, p, w_begin, w_end, ptr(npos), ptr(out))
        return out

    def accept_range(self, p, w_begin, w_end, mask):
        m = u8(mask)
        self.api("accept_range", self.h, p, w_begin, w_end, ptr(m, _P_U8))

    # ── drivers ────────────────────────────────────────────────────────────
    def vmc_sweeps(self, n_sweeps, tau, tape=None, drift=False):
        """VMC::sweep_walkers (vmc.cpp:82-155); drift=True is its use_drift
        form, the Langevin proposal with the Metropolis-Hastings correction."""
        acc = C.c_longlong()
        t = None if tape is None else f64(tape)
        self.api("vmc_drift_sweeps" if drift else "vmc_sweeps", self.h, n_sweeps, tau, ptr(t), C.byref(acc))
        return acc.value

    def dmc_sweeps(self, n_sweeps, tau, tape=None, stride=0, rr_p=None, rr_a=None):
        """rr_p / rr_a None: the rr sums accumulate where the walkers live."""
        acc = C.c_longlong()
        t = None if tape is None else f64(tape)
        self.api("dmc_sweeps", self.h, n_sweeps, tau, ptr(t), stride or self.nw, ptr(rr_p), ptr(rr_a),
                 C.byref(acc))
        return acc.value

    def refresh_inverse(self, w_begin=0, w_end=-1):
        """SlaterDeterminant::rebuild_inverse of every block for the walkers
        [w_begin, w_end) (w_end < 0: to nw)."""
        self.api("refresh_inverse", self.h, w_begin, w_end)

    def check_status(self):
        """Raises if a walker's determinant went singular in a rebuild."""
        self.api("check_status", self.h)

    def set_cancellation_factor(self, factor):
        self.api("set_cancellation_factor", self.h, float(factor))

    def set_refresh_updates(self, updates):
        self.api("set_refresh_updates", self.h, int(updates))

    def set_refresh_kernel(self, kind):
        """1: blocks of 33-64 orbitals rebuild on the matrix cores; 0: vector units."""
        self.api("set_refresh_kernel", self.h, int(kind))

    def rebuild_counts(self, reset=True):
        """Rebuild triggers since the last reset: (counts[4], per-walker totals)."""
        c = (C.c_longlong * 4)()
        per = np.zeros(self.nw, dtype=np.int32)
        self.api("rebuild_counts", self.h, c, ptr(per, _P_I), 1 if reset else 0)
        return list(c), per

    def draw_tape(self, n_sweeps):
        """The next n_sweeps sweeps' device-stream numbers [sweep][np][3][nw],
        advancing the streams as the sweeps would (alea_draw_tape)."""
        out = np.empty(n_sweeps * self.np * 3 * self.nw)
        self.api("draw_tape", self.h, n_sweeps, ptr(out))
        return out

    def sweep_kernel_ms(self):
        """Sweep-kernel time of every sweep of the last eager sweep call."""
        n = C.c_int()
        self.api("sweep_kernel_ms", self.h, None, 0, C.byref(n))
        out = np.zeros(n.value)
        self.api("sweep_kernel_ms", self.h, ptr(out), n.value, C.byref(n))
        return out

    def permute(self, src):
        s = i32(src)
        self.api("permute", self.h, len(s), ptr(s, _P_I))
        self.nw = len(s)

    def num_walkers(self):
        n = C.c_int()
        self.api("get_num_walkers", self.h, C.byref(n))
        return n.value

    # ── DMC population step ────────────────────────────────────────────────
    def dmc_begin(self):
        self.api("dmc_begin", self.h)

    def dmc_weights(self, e_trial, e_ref, e_cut, e_max, tau):
        self.api("dmc_weights", self.h, e_trial, e_ref, e_cut, e_max, tau)

    def dmc_sums(self):
        e, w, c = C.c_double(), C.c_double(), C.c_longlong()
        self.api("dmc_sums", self.h, C.byref(e), C.byref(w), C.byref(c))
        return e.value, w.value, c.value

    def weights(self):
        out = np.empty(self.nw)
        self.api("get_weights", self.h, ptr(out))
        return out

    def set_weights(self, w):
        a = f64(w)
        assert a.size == self.nw
        self.api("set_weights", self.h, ptr(a))

    def branch(self, rng_state, target):
        """WalkerSet::branch: rng_state (uint64[4]) advances in place; returns src."""
        assert rng_state.dtype == np.uint64 and rng_state.size == 4 and rng_state.flags.c_contiguous
        src = np.empty(2 * target + 1, dtype=np.int32)
        n = C.c_int()
        self.api("branch", self.h, rng_state.ctypes.data_as(_P_U64), target, C.byref(n), ptr(src, _P_I))
        self.nw = n.value
        return src[:n.value].copy()

    def close(self):
        if self.h:
            self.api("destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpu_walkers(spec: Spec, nw: int, precision: int = FP64) -> Walkers:
    """A walker batch on the GPU (libalea_amd.so)."""
    from . import load
    return Walkers(load(), spec, nw, precision)


def seed(w: Walkers, seed_value: int, stream: int = 0x766d632d74687264):
    w.api("seed", w.h, seed_value, stream)


def set_sweep_graphs(w: Walkers, on: bool = True):
    """Device-stream sweeps of `w` replay one captured hipGraph per block of
    sweeps (alea_set_sweep_graphs); results are bit-identical to the eager
    launches."""
    w.api("set_sweep_graphs", w.h, 1 if on else 0)


def vmc_sweeps_fused(batches, n_sweeps: int, taus):
    """n_sweeps VMC sweeps of several walker batches on their device streams
    as fused launches (alea_vmc_sweeps_fused): one sweep kernel per sweep over
    the walkers of all of them.  The batches share the precision, the complex
    form and the block sizes; queued asynchronously on the first batch's
    stream, which alea_query on that batch then covers."""
    import ctypes as C
    if not batches:
        return
    api = batches[0].api
    hs = (C.c_void_p * len(batches))(*[b.h for b in batches])
    ts = (C.c_double * len(batches))(*[float(t) for t in taus])
    api("vmc_sweeps_fused", hs, len(batches), n_sweeps, ts)
