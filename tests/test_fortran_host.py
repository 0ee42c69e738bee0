"""The Fortran host (sfincs_lib entry points + BMI names over the C ABI).

CPU: the host library exports the reference's driver and BMI names. GPU: the host
executable, run on an input deck written in the reference's formats, reproduces the
Python-driven run of the same scenario exactly (same kernels, same inputs)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from sfincs_amd import deck
from sfincs_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "sfincs_amd", "fortran")
HOSTLIB = os.path.join(FDIR, "libsfincs_host.so")
EXE = os.path.join(FDIR, "sfincs_mi355x")

needs_host = pytest.mark.skipif(not os.path.exists(HOSTLIB), reason="Fortran host not built (amdflang)")


@needs_host
def test_host_exports_driver_and_bmi_names():
    lib = ctypes.CDLL(HOSTLIB)
    for name in ("initialize", "update", "update_until", "finalize", "get_value_ptr", "get_current_time",
                 "get_start_time", "get_end_time", "get_time_step"):
        assert hasattr(lib, name), name
    syms = subprocess.run(["nm", "-D", HOSTLIB], capture_output=True, text=True).stdout
    for name in ("sfincs_initialize", "sfincs_update", "sfincs_finalize"):
        assert name in syms, name


def test_deck_round_trip_of_regular_mesh(tmp_path):
    sc = W.regular_case(12, 16, dx=50.0, seed=4, t_end=600.0)
    d = deck.write(sc, str(tmp_path))
    ind = np.fromfile(os.path.join(d, "sfincs.ind"), np.int32)
    assert ind[0] == 12 * 16 and np.array_equal(ind[1:], np.arange(1, 12 * 16 + 1))
    dep = np.fromfile(os.path.join(d, "sfincs.dep"), np.float32)
    assert np.array_equal(dep, sc.zb_grid.reshape(-1))


@needs_host
@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["regular", "quadtree"])
def test_fortran_host_matches_python_path(kind, tmp_path):
    print(f"case fortran host {kind}", flush=True)
    from sfincs_amd import _lib
    if kind == "regular":
        sc = W.regular_case(28, 36, dx=50.0, seed=41, t_end=900.0)
    else:
        sc = W.quadtree_case(8, 10, nlev=3, dx0=100.0, seed=42, t_end=900.0)
    d = deck.write(sc, str(tmp_path))
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    zs_f = np.fromfile(os.path.join(d, "zs_final.dat"), np.float64)
    inst = _lib.Instance(sc.build_model())
    inst.update(-1000.0)
    zs_p = inst.get("zs")
    assert zs_f.shape == zs_p.shape
    assert np.array_equal(zs_f, zs_p)
