"""The Fortran ISO_C_BINDING interface (wabbit_amd/fortran/wabbit_hip.f90)
compiles with the image's flang, links against libwabbit_hip.so and calls
the entry points that need no GPU (tests/fortran/abi_probe.f90); struct
sizes agree with the C header."""
import ctypes
import os
import shutil
import subprocess

import pytest

from wabbit_amd import lib as L
from wabbit_amd import wavelets as W

FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "wabbit_amd", "fortran")
DRIVERS = os.path.join(HERE, "fortran")


def build_driver(tmp_path, driver):
    """Compile the bind(C) module and a driver program of tests/fortran,
    linked against libwabbit_hip.so; returns the executable's path."""
    obj = tmp_path / "wh.o"
    subprocess.run([FLANG, "-c", os.path.join(FDIR, "wabbit_hip.f90"), "-module-dir", str(tmp_path),
                    "-o", str(obj)], check=True, cwd=tmp_path)
    exe = tmp_path / os.path.splitext(driver)[0]
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run([FLANG, os.path.join(DRIVERS, driver), str(obj), "-I", str(tmp_path),
                    f"-L{libdir}", "-lwabbit_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True, cwd=tmp_path)
    return exe


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not in this image")
def test_fortran_interface_links_and_runs(tmp_path):
    exe = build_driver(tmp_path, "abi_probe.f90")
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert "wh_version=100" in out
    assert f"sizeof(wh_pool_desc)={ctypes.sizeof(L.PoolDesc)}" in out
    assert f"sizeof(wh_wavelet)={ctypes.sizeof(W.WaveletC)}" in out
    assert f"sizeof(wh_adapt_params)={ctypes.sizeof(W.AdaptParamsC)}" in out
    assert f"sizeof(wh_mask_params)={ctypes.sizeof(W.MaskParamsC)}" in out
    assert f"sizeof(wh_time_params)={ctypes.sizeof(L.TimeParamsC)}" in out
    assert "CDF44 g=6 g_rhs=2 lifted=1" in out
    assert "forest_n=4" in out
