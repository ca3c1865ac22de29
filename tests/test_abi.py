"""The C-ABI library loads and exports every entry point include/mag_amd.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "mag_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"^\s*(?:int|const char\*)\s+(mag_\w+)\s*\(", src, flags=re.M)


def test_header_declares_the_boundary():
    names = header_functions()
    assert len(names) >= 40
    for must in ["mag_gauge_force", "mag_evolve_gauge", "mag_fg_integrate", "mag_hamiltonian", "mag_chain_run",
                 "mag_stout_smear", "mag_avg_wilson_loop", "mag_demon_microcanonical", "mag_expiQ_batch"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from mag_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from mag_amd import _lib
    assert sorted(header_functions()) == sorted(_lib.exported_names())


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from mag_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libmag_amd.so"))
    with pytest.raises(ImportError):
        _lib.lib()


def test_cpp_host_api_compiles():
    """The C++ mirror of the reference entry points (include/mag_amd.hpp) builds against the C ABI."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("no g++")
    out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                          os.path.join(ROOT, "tests", "cpp", "chmc_driver.cpp")], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
