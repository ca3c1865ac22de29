"""The compile-time switches of the HIP sources still compile for gfx950 in
their non-default settings (device code only, no GPU needed): the LDS-tiled
FP32 Iwasaki top kernel (MAG_TOP_LDS = 12, 20), the FP64 18-real stout
kernels with full 18-real reads or with two-row reads in the top kernel and
the thin force too (MAG_FP64_READ12 = 0, 63), and the local
Metropolis and checkerboard demon with full reads (MAG_LOCAL_READ12,
MAG_DEMON_READ12 = 0). About 45 s of hipcc."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mag_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CASES = [
    ("inst_f32.hip", "MAG_TOP_LDS=12"),
    ("inst_f32.hip", "MAG_TOP_LDS=20"),
    ("inst_f64.hip", "MAG_FP64_READ12=0"),
    ("inst_f64.hip", "MAG_FP64_READ12=63"),
    ("local.hip", "MAG_LOCAL_READ12=0"),
    ("demon.hip", "MAG_DEMON_READ12=0"),
]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("src,define", CASES)
def test_switch_compiles(src, define, tmp_path):
    out = tmp_path / "dev.o"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--offload-device-only", "-c", "-D" + define, "-o", str(out),
           os.path.join(CSRC, src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert out.stat().st_size > 0
