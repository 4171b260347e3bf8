"""The sweep kernels' bounded-argument sincos (alea_amd/csrc/trig.h) against
libm on the host: at most 4 ulp (values above 1e-3) and 4e-16 absolute; the
single-precision form within 2e-7 absolute (under two float ulp of 1)."""
import os
import subprocess

from backends import ROOT


def test_sincos_bounded_against_libm(tmp_path):
    exe = tmp_path / "trig_check"
    subprocess.run(["g++", "-O2", "-std=c++20", "-mfma", "-ffp-contract=fast", "-o", str(exe),
                    os.path.join(ROOT, "tests", "trig_check.cpp")], check=True)
    for lim in ("1e3", "1e5"):
        out = subprocess.run([str(exe), lim], check=True, capture_output=True, text=True).stdout.split()
        ulp_s, ulp_c, abs_err, abs_f32 = map(float, out)
        assert ulp_s <= 4 and ulp_c <= 4 and abs_err <= 4e-16, (lim, out)
        assert abs_f32 <= 2e-7, (lim, out)
