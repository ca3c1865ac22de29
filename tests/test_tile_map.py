"""The stencil kernels' tile map, run on the host: hipcc builds a CPU program
from the device header that checks, over 16 lattice shapes and both walk
orders, that each (direction, site block) tile is launched exactly once and
that the XCD slab walk keeps each XCD's tiles inside its own x3 slab."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_tile_map_covers_every_tile_once(tmp_path):
    exe = tmp_path / "tile_map_check"
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip",
                    os.path.join(HERE, "cpp", "tile_map_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == 32 and all(ln.startswith("ok ") for ln in lines), r.stdout
    # the shapes whose x3 slice (not their (x0, x1) plane) is a whole number of blocks take the slab walk
    assert "ok 24x24x24x24 X2=0 xcd-slab" in lines and "ok 8x8x8x8 X2=0 xcd-slab" in lines
    assert "ok 8x8x4x16 X2=0 xcd-slab" in lines and "ok 8x8x4x16 X2=1 lexicographic" in lines
    assert "ok 24x24x24x24 X2=1 lexicographic" in lines and "ok 32x32x32x32 X2=1 xcd-slab" in lines
