"""End-to-end config runs (alea_amd.run): the stage loop, the carry-over of
walkers between stages and the refusals on the CPU; on the GPU, the
reference's exact-energy checks run as whole configs (tests/test_vmc_physics.cpp:
86 for the 2-D exciton, E_L = -1 with zero variance; the plane-wave
eigenstate of tests/configs/periodic_2d_pw_free.yaml, E_L = 8 pi sqrt(3) /
(N rs^2) = 6.218741 at N = 7, rs = 1)."""
import math

import numpy as np
import pytest

from alea_amd import run

EXCITON = {
    "project": {"name": "exciton_2d_cusp_m1", "seed": 11},
    "system": {"lattice": {"type": "open", "dimension": 2, "L": 1.0},
               "particles": [{"name": "e", "group": "electron_group", "type": "fermions", "mass": 1.0,
                              "charge": -1, "up": 1, "down": 0, "initialization": "uniform"},
                             {"name": "h", "group": "hole_group", "type": "fermions", "mass": 1.0,
                              "charge": 1, "up": 1, "down": 0, "initialization": "uniform"}]},
    "wavefunction": [{"jastrow": {"type": "padé", "pair": ["electron_group", "hole_group"]}}],
    "hamiltonian": {"terms": [{"type": "coulomb", "pair": ["electron_group", "hole_group"], "epsilon": 1.0}]},
    "qmc": [{"method": "vmc", "random_walk": "metropolis", "samples": 256, "equilibration_blocks": 2,
             "accumulation_blocks": 8, "steps_per_block": 5, "sub_steps": 2, "time_step": 0.5},
            {"method": "dmc", "random_walk": "drift_diffusion", "samples": 256, "equilibration_blocks": 2,
             "accumulation_blocks": 8, "steps_per_block": 5, "sub_steps": 2, "time_step": 0.02}],
}

PW_FREE = {
    "project": {"name": "periodic_2d_pw_free", "seed": 42},
    "system": {"lattice": {"type": "periodic", "dimension": 2, "shape": "hexagonal", "rs": 1.0},
               "particles": [{"name": "e", "group": "electron_group", "type": "fermions", "mass": 1.0,
                              "charge": -1, "up": 7, "down": 0, "initialization": "uniform"}]},
    "wavefunction": [{"slater": {"group": "electron_group", "orbitals": {"type": "plane_wave"}}}],
    "hamiltonian": {"terms": []},
    "qmc": [{"method": "vmc", "random_walk": "metropolis", "samples": 64, "equilibration_blocks": 4,
             "accumulation_blocks": 8, "steps_per_block": 10, "sub_steps": 2, "time_step": 0.5}],
}


def test_rows_round_trip_and_seeding():
    n, nw = 3, 5
    soa = np.arange(n * 2 * nw, dtype=float)
    rows = run._rows(soa, n, nw)
    assert rows.shape == (nw, n * 2)
    assert rows[2, 1 * 2 + 1] == soa[1 * 2 * nw + 1 * nw + 2]
    assert np.array_equal(run._soa(rows, n), soa)
    # seed_walker_rows: walker w takes row w mod n_rows
    seeded = run._seed_rows(rows, 12)
    assert np.array_equal(seeded[7], rows[2]) and np.array_equal(seeded[11], rows[1])


@pytest.mark.parametrize("edit,msg", [
    (lambda c: c["qmc"][0].update(random_walk="heat_bath"), "random_walk"),
    (lambda c: c["qmc"][1].update(random_walk="metropolis"), "random_walk"),
    (lambda c: c["qmc"][1].update(reweighting="suppressed"), "reweighting"),
    (lambda c: c["qmc"][0].update(estimators=[{"type": "bond_order"}]), "bond_order"),
    (lambda c: c["qmc"][1].update(estimators=[{"type": "gofr", "rmax": 1, "n_bins": 4, "name": "x"}]), "named"),
    (lambda c: c["qmc"].insert(0, {"method": "optimizer"}), "optimizer"),
    (lambda c: c["qmc"][0].update(time_step=0), "time_step"),
    (lambda c: c["qmc"][0].update(forward_walking={"lags": [0.1], "interval": 1}), "belongs to a dmc stage"),
    (lambda c: c["qmc"][1].update(forward_walking={"lags": [0.03], "interval": 1}), "whole number of steps"),
    (lambda c: c["qmc"][1].update(forward_walking={"lags": [0.08, 0.04], "interval": 1}), "strictly increasing"),
    (lambda c: c["qmc"][1].update(forward_walking={"lags": [0.04], "interval": 2}), "divide steps_per_block"),
    (lambda c: c["qmc"][1].update(forward_walking={"lags": [0.44], "interval": 1}), "exceeds the equilibration"),
    (lambda c: c["qmc"][1].update(forward_walking={"lags": [0.04]}), "must be a map"),
])
def test_refusals(edit, msg, monkeypatch):
    import copy
    cfg = copy.deepcopy(EXCITON)
    edit(cfg)
    # nothing reaches the device: the walker factory fails the test if called
    monkeypatch.setattr(run, "gpu_walkers", lambda *a, **k: pytest.fail("reached the device"))
    with pytest.raises(ValueError, match=msg):
        run.run(cfg)


@pytest.mark.gpu
def test_exciton_config_runs_exactly():
    print("run: exciton VMC then DMC", flush=True)
    vmc, dmc_ = run.run(EXCITON)
    assert vmc["method"] == "vmc" and dmc_["method"] == "dmc"
    assert all(abs(e + 1.0) < 1e-12 for e in vmc["block_energies"])
    assert abs(vmc["variance"]) < 1e-12
    assert 0.0 < vmc["acceptance"] < 1.0
    assert all(abs(e + 1.0) < 1e-12 for e in dmc_["block_energies"])
    assert abs(dmc_["e_trial"] + 1.0) < 1e-6
    assert 0.9 < dmc_["acceptance"] <= 1.0


@pytest.mark.gpu
def test_exciton_langevin_vmc_runs_exactly():
    """random_walk: langevin (app.cpp:162) drifts the exact exciton along its
    own gradient: E_L stays -1 with zero variance, and the drift raises the
    acceptance above the Metropolis walk's at the same time step."""
    import copy
    cfg = copy.deepcopy(EXCITON)
    cfg["qmc"] = [cfg["qmc"][0]]
    print("run: exciton Metropolis VMC", flush=True)
    (plain,) = run.run(cfg)
    cfg["qmc"][0]["random_walk"] = "langevin"
    print("run: exciton Langevin VMC", flush=True)
    (vmc,) = run.run(cfg)
    assert all(abs(e + 1.0) < 1e-12 for e in vmc["block_energies"])
    assert abs(vmc["variance"]) < 1e-12
    assert plain["acceptance"] < vmc["acceptance"] < 1.0


@pytest.mark.gpu
def test_exciton_dmc_with_forward_walking():
    """forward_walking {lags, interval} on the exact exciton's DMC stage: the
    pure potential at every lag, with the lag's g(r) labelled fw<i>, and the
    virial theorem's <V> = 2E = -2 within the reblocked errors."""
    import copy
    cfg = copy.deepcopy(EXCITON)
    st = cfg["qmc"][1]
    st.update(forward_walking={"lags": [0.0, 0.2, 0.4], "interval": 5}, equilibration_blocks=4,
              accumulation_blocks=20, estimators=[{"type": "gofr", "rmax": 3.0, "n_bins": 12}])
    print("run: exciton DMC with forward walking", flush=True)
    _, d = run.run(cfg)
    fw = d["forward_walking"]
    assert fw["fw_lag_times"] == pytest.approx([0.0, 0.2, 0.4])
    assert [fw[f"fw{i}_lag_steps"] for i in range(3)] == [0.0, 5.0, 10.0]
    for i in range(3):
        assert fw[f"fw{i}_measurements"] == 20.0        # one snapshot per 5 steps, 5 steps per block
        v, e = fw[f"potential_fw{i}"], fw[f"potential_fw{i}_error"]
        assert e > 0.0 and abs(v + 2.0) < 5.0 * e + 2e-2, (i, v, e)
        assert len(fw[f"gofr_fw{i}_electron_group_hole_group"]) == 12


@pytest.mark.gpu
def test_config_estimators_run_on_the_device():
    """A config's gofr, sofk and density estimators: VMC of the free
    plane-wave eigenstate, g(r) and S(k) finite with block errors and a
    density that integrates to the particle count; the DMC stage of the
    exciton measures g_eh(r) with its weights."""
    import copy
    cfg = copy.deepcopy(PW_FREE)
    cfg["qmc"][0]["estimators"] = [{"type": "gofr", "rmax": 2.0, "n_bins": 20}, {"type": "sofk", "kmax": 12.0},
                                   {"type": "density", "n_bins": 8}]
    print("run: plane-wave VMC with gofr and sofk", flush=True)
    (vmc,) = run.run(cfg)
    est = vmc["estimators"]
    g = np.array(est["gofr_electron_group_electron_group"])
    assert g.shape == (20,) and np.all(np.isfinite(g))
    # Pauli exclusion: same-spin fermions avoid each other at short range
    assert g[0] < 0.2 * g[-5:].mean()
    s = np.array(est["sofk_electron_group_electron_group"])
    assert s.size == len(est["sofk_k"]) and np.all(s >= 0.0)
    rho = np.array(est["density_electron_group"])
    assert rho.shape == (64,) and len(est["density_edges"]) == 9 and np.all(rho >= 0.0)
    # the cell's area is N pi rs^2 = 7 pi: the density integrates to N = 7
    assert abs(rho.sum() * (7.0 * math.pi / 64.0) - 7.0) < 1e-9
    ex = copy.deepcopy(EXCITON)
    ex["qmc"][1]["estimators"] = [{"type": "gofr", "rmax": 3.0, "n_bins": 30}]
    print("run: exciton DMC with gofr", flush=True)
    _, d = run.run(ex)
    geh = np.array(d["estimators"]["gofr_electron_group_hole_group"])
    assert np.all(np.isfinite(geh)) and geh.sum() > 0.0


@pytest.mark.gpu
def test_plane_wave_eigenstate_config_runs_exactly():
    print("run: periodic plane-wave VMC", flush=True)
    (vmc,) = run.run(PW_FREE)
    exact = 8.0 * math.pi * math.sqrt(3.0) / 7.0
    assert abs(vmc["energy"] - exact) < 1e-9 * exact
    assert abs(vmc["variance"]) < 1e-9
