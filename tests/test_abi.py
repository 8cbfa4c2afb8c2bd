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
    assert lib.pm_plan_create(ctypes.byref(h), 16, 1 << 32, pmhip.PM_FP32, 0, 1.0, None) == pmhip.PM_ERR_ARG
    assert "2^32" in pmhip.last_error()
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


def test_tuning_entry_points_reject_bad_arguments():
    """The FFT shape / tuning cache entry points check their arguments before
    touching a GPU."""
    lib = pmhip.lib()
    shape = (ctypes.c_int * pmhip.PM_FFT_SLOTS)()
    ms = ctypes.c_double()
    assert lib.pm_plan_tune(None, 5, ctypes.byref(ms)) == pmhip.PM_ERR_ARG
    assert lib.pm_plan_set_fft_shape(None, shape) == pmhip.PM_ERR_ARG
    assert lib.pm_plan_get_fft_shape(None, shape) == pmhip.PM_ERR_ARG
    assert lib.pm_fft_tune(None, None, 5, None) == pmhip.PM_ERR_ARG
    assert lib.pm_fft_get_shape(None, shape) == pmhip.PM_ERR_ARG
    assert lib.pm_wisdom_import(None) == pmhip.PM_ERR_ARG
    assert lib.pm_wisdom_export(None) == pmhip.PM_ERR_ARG
    assert lib.pm_wisdom_lookup(512, pmhip.PM_FP32, shape, None) == pmhip.PM_ERR_ARG
    found = ctypes.c_int(7)
    pmhip.wisdom_forget()
    assert lib.pm_wisdom_lookup(512, pmhip.PM_FP32, None, ctypes.byref(found)) == pmhip.PM_OK
    assert found.value == 0
    gbs = ctypes.c_double()
    assert lib.pm_copy_rate(1 << 30, 5, None) == pmhip.PM_ERR_ARG
    assert lib.pm_copy_rate(1 << 20, 5, ctypes.byref(gbs)) == pmhip.PM_ERR_ARG  # below 16 MiB
    assert lib.pm_copy_rate((1 << 30) + 8, 5, ctypes.byref(gbs)) == pmhip.PM_ERR_ARG  # not whole 16-byte units
    assert lib.pm_copy_rate(1 << 30, 0, ctypes.byref(gbs)) == pmhip.PM_ERR_ARG


def _write_snapshot_numpy(path, n, ng, step, a, magic=b"PMSNAP01", version=1):
    import numpy as np

    head = bytearray(64)
    head[0:8] = magic
    head[8:16] = np.array([version, 64], "<u4").tobytes()
    head[16:24] = np.array([n], "<i8").tobytes()
    head[24:32] = np.array([ng, step], "<i4").tobytes()
    head[32:40] = np.array([a], "<f8").tobytes()
    data = np.arange(6 * n, dtype="<f8")
    with open(path, "wb") as f:
        f.write(bytes(head))
        f.write(data.tobytes())
    return data.reshape(6, n)


def test_snapshot_header_parsed_by_library_and_numpy(tmp_path):
    """pm_snapshot_info reads only the file (no GPU): it must decode a header
    written independently with numpy, agree with the numpy loader, and
    reject a wrong magic or version."""
    p = tmp_path / "snap.bin"
    data = _write_snapshot_numpy(p, 1000, 64, 7, 0.625)
    assert pmhip.snapshot_info(p) == (1000, 64, 7, 0.625)
    head, pos, vel = pmhip.load_snapshot_numpy(p)
    assert head == {"np": 1000, "ng": 64, "step": 7, "a": 0.625}
    assert (pos == data[:3]).all() and (vel == data[3:]).all()
    for kw in ({"magic": b"PMSNAP02"}, {"version": 2}):
        q = tmp_path / "bad.bin"
        _write_snapshot_numpy(q, 10, 8, 0, 1.0, **kw)
        with pytest.raises(pmhip.PMError) as e:
            pmhip.snapshot_info(q)
        assert e.value.code == pmhip.PM_ERR_ARG
    with pytest.raises(pmhip.PMError):
        pmhip.snapshot_info(tmp_path / "missing.bin")


def test_driver_rejects_more_than_one_rank():
    """The drop-in driver keeps the reference's MPI set-up (MPIEnv,
    MPI_COMM_WORLD, FFT::comm()) and runs as one rank owning the whole box:
    under two ranks it stops before touching the GPU."""
    exe = ROOT / "particle-mesh_amd" / "bin" / "pm_main"
    mpirun = Path("/opt/conda/bin/mpirun")
    if not exe.exists() or not mpirun.exists():
        pytest.skip("driver or mpirun not present")
    r = subprocess.run([str(mpirun), "-n", "2", str(exe), "16", "16"], capture_output=True, text=True, timeout=60,
                       cwd="/tmp")
    assert r.returncode != 0
    assert "one MPI rank" in r.stderr


def test_drop_in_headers_keep_the_reference_communicators():
    """The signatures a reference caller binds: FFT::comm(), the MPI_Comm of
    TransposeSlab / TransposeSlabBack (transpose_slab.h:6,
    transpose_slab_back.h:10) and report_max_time (utils.h:4)."""
    inc = ROOT / "include"
    fft = (inc / "fft" / "hip" / "fft.h").read_text()
    assert re.search(r"MPI_Comm\s+comm\(\)\s+const;", fft)
    ts = (inc / "decomp" / "slab" / "transpose_slab.h").read_text()
    assert re.search(r"const std::vector<ptrdiff_t>& all_n0, MPI_Comm comm\);", ts)
    tsb = (inc / "decomp" / "slab" / "transpose_slab_back.h").read_text()
    assert "MPI_Comm comm = MPI_COMM_WORLD" in tsb
    ut = (inc / "utils" / "utils.h").read_text()
    assert re.search(r"report_max_time\(double local_sec, MPI_Comm comm, const int i, const int worm_up, "
                     r"double& sum_time\);", ut)
    assert (inc / "utils" / "mpi_env.h").exists()
