"""The device factors as components of the reference's own TrialWavefunction
and Hamiltonian (include/alea_amd_component.hpp), through
tests/cpp/component_check.

The driver builds one trial twice from the reference's classes (compiled from
its sources into oracle/_ref/libalea_ref.so): once from the reference's native
PadeJastrow, DrummondJastrow, NosanowJastrow and EwaldTerm, once with
Drummond + Nosanow replaced by one DeviceComponent and the Ewald term by a
DevicePotential, and drives both through the component contract as
VMC::sweep_walkers and the DMC crowds do (vmc.cpp:82-155, dmc.cpp:300-340):
evaluate_batch, per-move ratio_batch / apply_move_batch over crowd ranges,
the deferred commit with refresh_derivatives (at the sweep's end, and per
particle before its move), branching with copies, a shrink
and a growth, clone, d log psi / dp, rebuild after a parameter change, local
energies and delta_potential_particle.  Determinant blocks (the reference's
SlaterDeterminant needs Eigen, absent here) sit in a DeviceComponent beside the
native Padé and are held to one C-ABI system holding all three factors, real
and twisted.  Every comparison holds to 1e-10 relative.

CPU: the C-ABI is the oracle's (tests/cpp/oracle_names.h); GPU: the HIP
library in FP64.  The driver is built only where the reference tree exists;
the GPU binary travels with the tree.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPU_DRIVER = os.path.join(ROOT, "tests", "cpp", "component_check")
GPU_DRIVER = os.path.join(ROOT, "tests", "cpp", "component_check_gpu")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libalea_ref.so")


def _run(driver):
    r = subprocess.run([driver], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    last = out.strip().splitlines()[-1]
    assert last.endswith(" 0 failures"), out[-4000:]
    n_checks = int(last.split()[0])
    assert n_checks > 1200, last


def test_device_components_inside_the_reference_trial_oracle_backed():
    if os.path.isdir("/root/reference/src") and os.path.exists(REF_LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"),
                        os.path.join(ROOT, "tests", "cpp", "component_check")], check=True, timeout=300)
    if not os.path.exists(CPU_DRIVER):
        pytest.skip("tests/cpp/component_check needs the reference's headers and oracle/_ref to build")
    _run(CPU_DRIVER)


@pytest.mark.gpu
def test_device_components_inside_the_reference_trial_on_the_gpu():
    if not os.path.exists(GPU_DRIVER):
        pytest.skip("tests/cpp/component_check_gpu is built only where the reference tree exists")
    _run(GPU_DRIVER)
