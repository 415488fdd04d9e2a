"""The C-ABI library loads and exports every symbol include/wabbit_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import subprocess

from wabbit_amd import lib as L


def test_header_declares_entry_points():
    syms = L.declared_symbols()
    for s in ("wh_pool_create", "wh_sync_ghosts", "wh_rhs_acm", "wh_rk_step", "wh_dt_reduce",
              "wh_run_steps"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [s for s in L.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(L.declared_symbols()) <= exported


def test_gfx950_code_object_present():
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o",
                          f"--input={L.LIB_PATH}"], capture_output=True, text=True)
    # the bundler lists the offload targets embedded in the host object
    if out.returncode == 0 and out.stdout:
        assert "gfx950" in out.stdout
