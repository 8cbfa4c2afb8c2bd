"""The C-ABI library loads, exports every entry point that include/pm_api.h
declares, and rejects bad arguments before touching the GPU.  CPU only."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

import pmhip

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "pm_api.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert pmhip.LIB_PATH.exists(), "run __graft_entry__.build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", str(pmhip.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (pm_[a-z0-9_]+)$", out, flags=re.M))
    decl = declared_functions()
    assert decl, "no declarations parsed"
    missing = [f for f in decl if f not in exported]
    assert not missing, f"declared but not exported: {missing}"
    # the ctypes binding covers the same set
    assert set(pmhip.SIGNATURES) == set(decl)


def test_library_loads_and_reports_version():
    lib = pmhip.lib()
    assert b"gfx950" in lib.pm_version()
    assert abs(lib.pm_function_a(0.9) - 1.0 / (0.3 / 0.9 + 0.7 * 0.81) ** 0.5) < 1e-15


@pytest.mark.parametrize("ng", [0, 7, 12, 2048])
def test_plan_rejects_bad_mesh(ng):
    lib = pmhip.lib()
    h = ctypes.c_void_p()
    rc = lib.pm_plan_create(ctypes.byref(h), ng, 100, pmhip.PM_FP32, pmhip.PM_GREEN_REF_KH, 1.0, None)
    assert rc == pmhip.PM_ERR_ARG
    assert h.value is None
    assert "power of two" in pmhip.last_error()


def test_plan_rejects_bad_modes():
    lib = pmhip.lib()
    h = ctypes.c_void_p()
    assert lib.pm_plan_create(ctypes.byref(h), 16, 0, pmhip.PM_FP32, 0, 1.0, None) == pmhip.PM_ERR_ARG
    assert lib.pm_plan_create(ctypes.byref(h), 16, 10, 7, 0, 1.0, None) == pmhip.PM_ERR_ARG
    assert lib.pm_plan_create(ctypes.byref(h), 16, 10, 0, 9, 1.0, None) == pmhip.PM_ERR_ARG
    assert lib.pm_plan_create(None, 16, 10, 0, 0, 1.0, None) == pmhip.PM_ERR_ARG
    assert lib.pm_fft_create(ctypes.byref(h), 24, 1, None) == pmhip.PM_ERR_ARG


def test_null_arguments_rejected():
    lib = pmhip.lib()
    assert lib.pm_deposit(None, None, None, None, 0) == pmhip.PM_ERR_ARG
    assert lib.pm_step(None, None, None, None, None, None, None, 0.0, 0.0, 0) == pmhip.PM_ERR_ARG
    assert lib.pm_reorder_slab(None, None, 16, 16, None) == pmhip.PM_ERR_ARG
    assert lib.pm_reorder_slab_back(None, None, 16, 8, None) == pmhip.PM_ERR_ARG
    assert lib.pm_centered_derivatives(None, None, None, None, 5, 5, 5, None) == pmhip.PM_ERR_ARG
    assert lib.pm_update_particles(None, None, None, None, None, None, None, None, None, 0.9, 10, 5, 5, 5, 0, 0, 0,
                                   None) == pmhip.PM_ERR_ARG
    assert lib.pm_init_buf(None, 5, 5, 5, None) == pmhip.PM_ERR_ARG
