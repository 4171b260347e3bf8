"""The C++20 host interface (include/alea_amd.hpp) through tests/cpp/hpp_driver.

CPU: the header compiles under -std=c++20 -Wall -Wextra -Werror; refused
inputs surface as std::invalid_argument before any device work; its host
restatements (xoshiro seeding, robust_spread, PopulationControl) print the
same values as alea_amd.dmc, which the population tests pin against the
reference.

GPU: a walker batch built from the same spec through WalkerSet /
TrialWavefunction / Hamiltonian gives the same log psi, derivatives, ratios,
local energies, VMC trajectory and DMC populations as the Python binding on
the same inputs and tapes.  Both drive the same C-ABI calls in the same order
on deterministic kernels, so the comparison is exact.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from alea_amd import dmc
from alea_amd.system import heg_spec, dipole_spec, uniform_positions
from alea_amd.walkers import FP64, MIXED, gpu_walkers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "hpp_driver")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/hpp_driver is not built: run __graft_entry__.build()")
    return DRIVER


def test_header_compiles_and_driver_builds():
    """Build from source here (host compiler only), so the header is checked
    on every CPU run."""
    if not os.path.exists(os.path.join(ROOT, "alea_amd", "libalea_amd.so")):
        pytest.fail("alea_amd/libalea_amd.so is not built: run __graft_entry__.build()")
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "tests", "cpp")], check=True, timeout=300)
    assert os.access(DRIVER, os.X_OK)


def _selftest():
    r = subprocess.run([_driver(), "selftest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines()}


def test_refusals_and_host_restatements():
    out = _selftest()
    st = dmc.xoshiro_state(dmc.stream_seed(7, dmc.K_DMC_BRANCH_STREAM, 3))
    assert [int(x) for x in out["xoshiro"]] == [int(x) for x in st]
    assert float(out["robust_odd"][0]) == dmc.robust_spread([3.0, -1.0, 4.0, 1.5, 9.0, -2.5, 6.0])
    assert float(out["robust_even"][0]) == dmc.robust_spread([0.25, -0.5, 2.0, 8.0, 1.0, -3.0])
    pc = dmc.PopulationControl(0.02, 2, 8.0, 100, "universal", 2.0, 26)
    pc.start(-1.25)
    pc.set_energy_spread(0.4)
    assert [float(x) for x in out["popctl_cut"]] == [pc.e_cut, pc.e_max]
    w = [pc.branch_weight(-1.1, -1.32), pc.branch_weight(-1.25 - 1.1 * pc.e_cut, -1.2),
         pc.branch_weight(-1.25 - 2.0 * pc.e_cut, -1.2)]
    assert [float(x) for x in out["popctl_w"]] == w
    # closed forms: x = 1.1 e_cut tapers tau/2 by 1 - 0.1/0.5; x = 2 e_cut > e_max cuts it to zero
    assert math.isclose(math.log(w[1]), 0.8 * 0.01 * (2 * -1.25 + 1.25 + 1.1 * pc.e_cut + 1.2), rel_tol=1e-12)
    assert w[2] == 1.0
    assert math.isclose(math.log(w[0]), 0.01 * (2 * -1.25 + 1.1 + 1.32), rel_tol=1e-12)
    pc.update(-1.3, 97.5, False)
    assert [float(x) for x in out["popctl_u0"]] == [pc.eref, pc.etrial]
    pc.update(-1.28, 103.0, True)
    pc.update(-1.31, 99.0, True)
    assert [float(x) for x in out["popctl_u2"]] == [pc.eref, pc.etrial]
    assert float(out["popctl_drv"][0]) == dmc.PopulationControl(0.04, 1, 10.0, 50, "drv").e_cut


# ── spec and inputs as text ──────────────────────────────────────────────────
def _arr(a):
    a = np.asarray(a).reshape(-1)
    if a.dtype.kind in "iub":
        return f"{a.size} " + " ".join(str(int(x)) for x in a)
    return f"{a.size} " + " ".join(repr(float(x)) for x in a)


def _table(t):
    return " ".join([_arr(t["n1"]), _arr(t["n2"]), _arr(t["kx"]), _arr(t["ky"]), _arr(t["w"]),
                     _arr(t["spline"]), repr(float(t["r_cut"])), repr(float(t["inv_delta"]))])


def spec_text(s, nw, precision):
    lines = [f"system {nw} {s.n_particles} {precision} " + (_arr(s.lattice) if s.lattice is not None else "0")]
    for first, last, q, im in s.groups:
        lines.append(f"group {first} {last} {q!r} {1.0 / im!r}")
    for j in s.pade:
        t = j.get("image")
        tail = f"1 {_table(t)} {float(t['u_inf'])!r}" if t else "0"
        lines.append(f"pade {j['ga']} {j['gb']} {j['num']} {j['den']} {_arr(j['coeffs'])} {tail}")
    for d in s.drummond:
        lines.append(f"drummond {d['ga']} {d['gb']} {d['C']} {float(d['gamma'])!r} {float(d['L'])!r} {_arr(d['coeffs'])}")
    for n in s.nosanow:
        lines.append(f"nosanow {n['group']} {float(n['alpha'])!r} {_arr(n['sites'])}")
    for first, last, k, sn in s.slater_pw:
        lines.append(f"slater_pw {first} {last} {_arr(np.asarray(k, dtype=np.float64))} {_arr(np.asarray(sn, dtype=np.int64))}")
    for first, last, q in s.slater_twisted:
        lines.append(f"slater_twisted {first} {last} {_arr(np.asarray(q, dtype=np.float64))}")
    for first, last, cen, al in s.slater_gauss:
        lines.append(f"slater_gauss {first} {last} {_arr(np.asarray(cen, dtype=np.float64))} {_arr(np.asarray(al, dtype=np.float64))}")
    # the factors in the order Walkers registers them, so that both count them alike
    for ga, gb, b, n in s.mcmillan:
        lines.append(f"mcmillan {ga} {gb} {float(b)!r} {n}")
    if s.ewald:
        e = s.ewald
        lines.append(f"ewald {_table(e)} {float(e['v_long_zero'])!r} {float(e['sr_k0'])!r}")
    for ga, gb, eps in s.coulomb:
        lines.append(f"coulomb {ga} {gb} {float(eps)!r}")
    for ga, gb, c, p in s.inverse_power:
        lines.append(f"inverse_power {ga} {gb} {float(c)!r} {float(p)!r}")
    lines.append("end")
    return "\n".join(lines) + "\n"


def _read_out(path):
    out = {}
    with open(path) as f:
        for ln in f:
            tok = ln.split()
            out[tok[0]] = np.array([float(x) for x in tok[2:2 + int(tok[1])]])
    return out


def _case(spec, cap, seed_):
    """Inputs of one run: positions, per-particle proposals and masks, a VMC
    tape, DMC settings and per-substep tapes of [np][3][cap]."""
    rng = np.random.default_rng(seed_)
    box = 3.0 if spec.lattice is None else 1.0
    pos = uniform_positions(spec, cap, rng, box=box)
    scale = 0.5 if spec.lattice is None else spec.lattice[0] / 10
    P = pos.reshape(spec.n_particles, 2, cap)
    moves = [((P[p] + rng.normal(scale=scale, size=(2, cap))).reshape(-1), rng.random(cap) < 0.6)
             for p in range(spec.n_particles)]
    tau = 0.3 if spec.lattice is None else (spec.lattice[0] / 6) ** 2
    n_sw = 3
    tape = np.stack([np.concatenate([rng.normal(size=(spec.n_particles, 2, cap)),
                                     rng.random((spec.n_particles, 1, cap))], axis=1) for _ in range(n_sw)])
    dmc_cfg = dict(target=cap // 3, tau=0.05 if spec.lattice is None else (spec.lattice[0] / 40) ** 2,
                   n_sub=2, n_steps=4, measure_from=2, seed=9)
    dtapes = [[np.concatenate([rng.normal(size=(spec.n_particles, 2, cap)), rng.random((spec.n_particles, 1, cap))],
                              axis=1).reshape(-1) for _ in range(dmc_cfg["n_sub"])] for _ in range(dmc_cfg["n_steps"])]
    return pos, moves, tau, n_sw, tape.reshape(-1), dmc_cfg, dtapes


def _input_text(case):
    pos, moves, tau, n_sw, tape, c, dtapes = case
    parts = [_arr(pos)]
    for newp, mask in moves:
        parts += [_arr(newp), _arr(mask.astype(np.float64))]
    parts += [repr(tau), str(n_sw), _arr(tape)]
    parts += [str(c["target"]), repr(c["tau"]), str(c["n_sub"]), str(c["n_steps"]), str(c["measure_from"]),
              str(c["seed"])]
    for step in dtapes:
        parts += [_arr(t) for t in step]
    return "\n".join(parts) + "\n"


def _python_run(make, case):
    pos, moves, tau, n_sw, tape, c, dtapes = case
    w = make()
    cap = w.capacity
    out = {}
    try:
        w.set_positions(pos)
        w.evaluate()
        out["log_psi0"] = w.log_psi()
        out["grad0"], out["lap0"] = w.gradients()
        out["e0"], out["e0_im"] = w.local_energy()
        for p, (newp, mask) in enumerate(moves):
            lr, ph, ng, nl = w.ratio(p, newp)
            out[f"lr{p}"], out[f"ph{p}"], out[f"ng{p}"], out[f"nl{p}"] = lr, ph, ng, nl
            w.accept(p, mask)
        out["log_psi1"] = w.log_psi()
        out["e1"] = w.local_energy()[0]
        free = _table_free_parameters(w.spec)
        if free:
            out["dlogpsi"] = w.parameter_derivatives(free).reshape(-1)
            de, de_im = w.local_energy_parameter_derivatives(free)
            out["de"], out["de_im"] = de.reshape(-1), de_im.reshape(-1)
        out["vmc_accepted"] = np.array([float(w.vmc_sweeps(n_sw, tau, tape))])
        out["pos2"] = w.positions()
        out["log_psi2"] = w.log_psi()
        out["e2"] = w.local_energy()[0]
        w.permute(np.arange(c["target"]))
        d = dmc.DMC(w, c["tau"], c["target"], n_substeps=c["n_sub"], seed=c["seed"],
                    tapes=lambda step, sub, n: dtapes[step][sub], stride=cap)
        for step in range(c["n_steps"]):
            e, ws = d.step(measure=step >= c["measure_from"])
            h = d.history[-1]
            out[f"dmc_e{step}"], out[f"dmc_w{step}"] = np.array([e]), np.array([ws])
            out[f"dmc_nw{step}"] = np.array([float(h["nw"])])
            out[f"dmc_etrial{step}"] = np.array([h["etrial"]])
            out[f"dmc_src{step}"] = np.asarray(h["src"], dtype=np.float64)
        out["pos3"] = w.positions()
        out["weights3"] = w.weights()
        out["langevin_accepted"] = np.array([float(w.vmc_sweeps(n_sw, tau, tape, drift=True))])
        out["pos_l"] = w.positions()
        out["log_psi_l"] = w.log_psi()
    finally:
        w.close()
    return out


def _table_free_parameters(spec):
    """The parameters the driver frees: those that need no breakup tables
    (minimum-image Padé coefficients, Drummond coefficients and L, Nosanow
    alpha, a uniform Gaussian block's width), factors in registration order."""
    free, f = [], 0
    for j in spec.pade:
        free += [] if j.get("image") else [(f, i) for i in range(j["num"] + j["den"])]
        f += 1
    for d in spec.drummond:
        free += [(f, i) for i in range(len(d["coeffs"]) + 1)]
        f += 1
    for _ in spec.nosanow:
        free.append((f, 0))
        f += 1
    f += len(spec.slater_pw) + len(spec.slater_twisted)
    for first, last, cen, al in spec.slater_gauss:
        if np.all(np.asarray(al) == np.asarray(al)[0]):
            free.append((f, 0))
        f += 1
    return [{"factor": a, "index": b} for a, b in free]


def _specs():
    a = heg_spec(5, 5, 2.0, twist=(0.13, 0.27))
    a.drummond.append(dict(ga=0, gb=1, C=3, gamma=1.0, L=a.lattice[0] / 2, coeffs=[1e-3, -2e-4, 3e-5]))
    a.name = "twisted N=10"
    b = dipole_spec(16, density=4.0)
    b.name = "dipoles N=16"
    return [a, b]


def _run_driver(driver, spec, cap, precision, case, tmp_path):
    sp, inp, outp = tmp_path / "spec.txt", tmp_path / "in.txt", tmp_path / "out.txt"
    sp.write_text(spec_text(spec, cap, precision))
    inp.write_text(_input_text(case))
    print(f"{os.path.basename(driver)} run: {spec.name} precision={precision}", flush=True)
    r = subprocess.run([driver, "run", str(sp), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return _read_out(outp)


def _compare(got, want, n_steps):
    assert set(got) == set(want)
    for k, v in want.items():
        v = np.asarray(v, dtype=np.float64)
        assert got[k].shape == v.shape, k
        assert np.array_equal(got[k], v), k
    assert len({float(want[f"dmc_nw{s}"][0]) for s in range(n_steps)}) > 1     # the population moved


@pytest.mark.parametrize("spec", _specs(), ids=lambda s: s.name)
def test_cpp_interface_on_the_oracle(spec, tmp_path):
    """The same driver linked against the oracle library, on the CPU."""
    from backends import oracle_walkers
    cap = 48
    case = _case(spec, cap, 5)
    driver = DRIVER + "_oracle"
    if not os.path.exists(driver):
        pytest.fail("tests/cpp/hpp_driver_oracle is not built: run __graft_entry__.build()")
    got = _run_driver(driver, spec, cap, FP64, case, tmp_path)
    _compare(got, _python_run(lambda: oracle_walkers(spec, cap), case), case[5]["n_steps"])


@pytest.mark.gpu
@pytest.mark.parametrize("spec,precision", [pytest.param(s, p, id=f"{s.name}-{'fp64' if p == FP64 else 'mixed'}")
                                             for s in _specs() for p in (FP64, MIXED)])
def test_cpp_interface_matches_python_binding(spec, precision, tmp_path):
    cap = 48
    case = _case(spec, cap, 5)
    got = _run_driver(_driver(), spec, cap, precision, case, tmp_path)
    _compare(got, _python_run(lambda: gpu_walkers(spec, cap, precision), case), case[5]["n_steps"])


# ── Estimators through the C++ class ───────────────────────────────────────
def _est_cases():
    from test_estimators import _spec
    from alea_amd.system import hexagonal_lattice
    return [pytest.param(_spec(hexagonal_lattice(10, 1.5), (6, 4), "hex"), (3.0, 24, 9.0, 0.0, 12), id="hex 6+4"),
            pytest.param(_spec(None, (3, 2), "open"), (2.0, 20, 0.0, 1.2, 10), id="open 3+2")]


def _est_compare(driver, make, spec, args, tmp_path, weighted):
    from test_estimators import _inputs, _run
    nw, blocks, steps = 11, 3, 2
    rmax, nb, kmax, d_rmax, d_bins = args
    pos, wts = _inputs(spec, nw, blocks, steps, weighted, 21)
    sp, inp, outp = tmp_path / "spec.txt", tmp_path / "in.txt", tmp_path / "out.txt"
    lines = [f"system {nw} {spec.n_particles} 0 " + (_arr(spec.lattice) if spec.lattice is not None else "0")]
    for first, last, q, im in spec.groups:
        lines.append(f"group {first} {last} {q!r} {1.0 / im!r}")
    sp.write_text("\n".join(lines + ["end"]) + "\n")
    parts = [repr(rmax), str(nb), repr(kmax), str(blocks), str(steps), "1" if weighted else "0", repr(d_rmax),
             str(d_bins)]
    for m in range(blocks * steps):
        parts += [_arr(pos[m]), _arr(wts[m] if weighted else np.ones(nw))]
    inp.write_text("\n".join(parts) + "\n")
    print(f"{os.path.basename(driver)} estimators: {spec.name} weighted={weighted}", flush=True)
    r = subprocess.run([driver, "estimators", str(sp), str(inp), str(outp)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = _read_out(outp)
    kw = dict(gofr=(rmax, nb), density=(d_rmax, d_bins))
    if kmax > 0:
        kw["sofk"] = kmax
    w = make(spec, nw)
    try:
        want = _run(w, spec, kw, pos, wts, blocks, steps)
    finally:
        w.close()
    assert set(got) == set(want)
    for k, v in want.items():
        assert np.array_equal(got[k], np.asarray(v, dtype=np.float64)), k


@pytest.mark.parametrize("weighted", [False, True], ids=["vmc", "dmc"])
@pytest.mark.parametrize("spec,args", _est_cases())
def test_cpp_estimators_on_the_oracle(spec, args, weighted, tmp_path):
    """The C++ Estimators class gives the Python module's series bit for
    bit on the same library (here the oracle, on the CPU)."""
    from backends import oracle_walkers
    driver = DRIVER + "_oracle"
    if not os.path.exists(driver):
        pytest.fail("tests/cpp/hpp_driver_oracle is not built: run __graft_entry__.build()")
    _est_compare(driver, oracle_walkers, spec, args, tmp_path, weighted)


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [False, True], ids=["vmc", "dmc"])
@pytest.mark.parametrize("spec,args", _est_cases())
def test_cpp_estimators_match_python_binding(spec, args, weighted, tmp_path):
    _est_compare(_driver(), lambda s, n: gpu_walkers(s, n), spec, args, tmp_path, weighted)
