"""The C-ABI library loads and exports every symbol include/phx_elph.h declares."""
import ctypes
import os
import re

import numpy as np

import phoebe_amd
from phoebe_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "phx_elph.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(phoebe_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(phoebe_amd.EXPORTS), set(names) ^ set(phoebe_amd.EXPORTS)


def test_version_and_error_strings():
    assert b"gfx950" in phoebe_amd.lib().phx_version()
    assert phoebe_amd.lib().phx_last_error() is not None


def test_error_path_without_gpu_work():
    # argument validation fails before any device call
    rc = phoebe_amd.lib().phx_zheev_batched(1, 100, None, None, None)
    assert rc < 0
    assert b"N <= 44" in phoebe_amd.lib().phx_last_error()


def test_synthetic_gw_host_twin_matches_numpy():
    p = synth.make_problem(1, 2, 2, 2, 2, seed=11)
    a = phoebe_amd.synth_gw_host(p)
    b = p.gw_host().reshape(-1).view(np.float64)
    assert np.array_equal(a, b)


def test_cpp_adaptor_compiles():
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "adaptor_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
