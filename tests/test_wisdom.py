"""The FFT shape cache as a file (pm_wisdom_import / pm_wisdom_export /
pm_wisdom_lookup / pm_wisdom_forget): host code only, no GPU.  The format
is the library's own (the reference's FFTW wisdom, fftw/fft.cpp:28-49, is
FFTW's): a "pm-wisdom 1" line, then "ng fp32|fp64 v0..v4 solve_ms" lines."""
import pytest

import pmhip


@pytest.fixture(autouse=True)
def clean_cache():
    pmhip.wisdom_forget()
    yield
    pmhip.wisdom_forget()


def test_import_lookup_export_round_trip(tmp_path):
    f = tmp_path / "w.dat"
    f.write_text("pm-wisdom 1\n# comment\n\n512 fp32 1 2 3 4 5 1.25\n64 fp64 0 5 0 1 0 0.5\n")
    pmhip.wisdom_import(f)
    assert pmhip.wisdom_lookup(512, pmhip.PM_FP32) == [1, 2, 3, 4, 5]
    assert pmhip.wisdom_lookup(64, pmhip.PM_FP64) == [0, 5, 0, 1, 0]
    assert pmhip.wisdom_lookup(512, pmhip.PM_FP64) is None
    out = tmp_path / "o.dat"
    pmhip.wisdom_export(out)
    lines = out.read_text().splitlines()
    assert lines[0] == "pm-wisdom 1"
    assert [ln.split()[:7] for ln in lines if ln and ln[0].isdigit()] == [
        ["64", "fp64", "0", "5", "0", "1", "0"], ["512", "fp32", "1", "2", "3", "4", "5"]]
    pmhip.wisdom_forget()
    assert pmhip.wisdom_lookup(512, pmhip.PM_FP32) is None
    pmhip.wisdom_import(out)
    assert pmhip.wisdom_lookup(512, pmhip.PM_FP32) == [1, 2, 3, 4, 5]


def test_import_replaces_same_key_and_keeps_others(tmp_path):
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    a.write_text("pm-wisdom 1\n256 fp32 1 1 1 1 1 0.3\n128 fp32 2 2 2 2 2 0.1\n")
    b.write_text("pm-wisdom 1\n256 fp32 4 0 4 0 4 0.2\n")
    pmhip.wisdom_import(a)
    pmhip.wisdom_import(b)
    assert pmhip.wisdom_lookup(256, pmhip.PM_FP32) == [4, 0, 4, 0, 4]
    assert pmhip.wisdom_lookup(128, pmhip.PM_FP32) == [2, 2, 2, 2, 2]


def test_small_meshes_always_take_the_default_shape(tmp_path):
    f = tmp_path / "w.dat"
    f.write_text("pm-wisdom 1\n32 fp32 3 3 3 3 3 0.01\n")
    pmhip.wisdom_import(f)
    assert pmhip.wisdom_lookup(32, pmhip.PM_FP32) == [0, 0, 0, 0, 0]


@pytest.mark.parametrize("line", [
    "512 fp16 0 0 0 0 0 1.0",        # unknown precision
    "500 fp32 0 0 0 0 0 1.0",        # not a power of two
    "2048 fp32 0 0 0 0 0 1.0",       # above the largest mesh
    "512 fp32 0 0 9 0 0 1.0",        # variant out of range
    "512 fp32 0 0 -1 0 0 1.0",
    "512 fp32 0 0 0 0 1.0",          # a field missing
    "512 fp32 0 0 0 0 0 1.0 extra",  # trailing garbage
    "512 fp32 0 0 0 0 0 nan",        # not a time
    "pm-wisdom 1",                   # a second header
])
def test_a_bad_entry_imports_nothing(tmp_path, line):
    """A valid entry before the bad one is not imported either."""
    good = tmp_path / "good.dat"
    good.write_text("pm-wisdom 1\n128 fp64 1 0 0 0 0 0.2\n")
    bad = tmp_path / "bad.dat"
    bad.write_text("pm-wisdom 1\n256 fp64 1 1 1 1 1 0.2\n" + line + "\n")
    pmhip.wisdom_import(good)
    with pytest.raises(pmhip.PMError) as e:
        pmhip.wisdom_import(bad)
    assert e.value.code == pmhip.PM_ERR_ARG
    assert "line 3" in str(e.value)
    assert pmhip.wisdom_lookup(256, pmhip.PM_FP64) is None
    assert pmhip.wisdom_lookup(128, pmhip.PM_FP64) == [1, 0, 0, 0, 0]


@pytest.mark.parametrize("body", [
    "pm-wisdom 2\n512 fp32 0 0 0 0 0 1.0\n",   # unknown version
    "wisdom 1\n512 fp32 0 0 0 0 0 1.0\n",      # wrong tag
    "512 fp32 0 0 0 0 0 1.0\n",                 # no header
    "",                                         # empty
])
def test_a_bad_header_imports_nothing(tmp_path, body):
    bad = tmp_path / "bad.dat"
    bad.write_text(body)
    with pytest.raises(pmhip.PMError) as e:
        pmhip.wisdom_import(bad)
    assert e.value.code == pmhip.PM_ERR_ARG
    assert pmhip.wisdom_lookup(512, pmhip.PM_FP32) is None


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(pmhip.PMError):
        pmhip.wisdom_import(tmp_path / "absent.dat")
