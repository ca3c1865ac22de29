"""The FP32 exp(iQ) / Lambda coefficient series of the device code, run on the
host: the device headers are host-callable, so hipcc builds a CPU program that
checks the series against the FP64 closed forms (hep-lat/0311018) and against
e^{iq} on a diagonal Q."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_series_coefficients_match_closed_forms(tmp_path):
    exe = tmp_path / "ch_series_check"
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip",
                    os.path.join(HERE, "cpp", "ch_series_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
