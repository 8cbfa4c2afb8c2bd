"""FFT pass shapes and the tuning cache on the GPU.  Every shape variant of
every pass gives the default shape's potential (plan) and spectrum /
inverse (FFT handle) bit for bit: only the split of the work over
workgroups differs.  pm_plan_tune / pm_fft_tune pick valid shapes, cache
them for handles created later, and the cache survives an export / import
round trip; the drop-in FFT class of the C++ driver writes and then reads
its wisdom file as the reference's FFTW class does (fftw/fft.cpp:28-49,
63-70)."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pmhip

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "particle-mesh_amd" / "bin" / "pm_main"


@pytest.fixture(autouse=True)
def clean_cache():
    pmhip.wisdom_forget()
    yield
    pmhip.wisdom_forget()


def _particles(ng, dev, seed):
    import torch

    nrow = ng // 2
    noise = torch.tensor(np.random.default_rng(seed).standard_normal(nrow ** 3), device=dev)
    return pmhip.zeldovich_ic(nrow, ng, noise, 1.0, 1.0)


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
@pytest.mark.parametrize("ng", [64, 256])
def test_every_plan_shape_solves_bit_identically(dev, ng, prec):
    import torch

    print(f"case plan shapes ng={ng} prec={prec}", flush=True)
    pos, _ = _particles(ng, dev, 51)
    with pmhip.Plan(ng, pos[0].numel(), prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        assert plan.fft_shape == [0] * pmhip.PM_FFT_SLOTS
        plan.deposit(pos, pmhip.PM_DEPOSIT_SORTED)
        plan.solve()
        torch.cuda.synchronize()
        ref = plan.potential().clone()
        assert float(ref.abs().max()) > 0
        shapes = [[v if s == slot else 0 for s in range(pmhip.PM_FFT_SLOTS)]
                  for slot in range(pmhip.PM_FFT_SLOTS) for v in range(1, pmhip.PM_FFT_VARIANTS)]
        rng = np.random.default_rng(52)
        shapes += [list(rng.integers(0, pmhip.PM_FFT_VARIANTS, pmhip.PM_FFT_SLOTS)) for _ in range(4)]
        for shape in shapes:
            plan.fft_shape = shape
            assert plan.fft_shape == [int(v) for v in shape]
            plan.solve()
            torch.cuda.synchronize()
            assert torch.equal(plan.potential(), ref), shape


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_every_fft_handle_shape_is_bit_identical(dev, tmp_path, prec):
    """The FFT handle runs the forward-only and inverse-only x passes;
    shapes reach a handle through the cache when it is created."""
    import torch

    ng = 128
    print(f"case fft handle shapes prec={prec}", flush=True)
    dt = torch.float32 if prec == pmhip.PM_FP32 else torch.float64
    g = torch.Generator(device=dev)
    g.manual_seed(53)
    slab = torch.rand(ng ** 3, dtype=dt, device=dev, generator=g)

    def run(f):
        out = torch.empty_like(slab)
        f.forward(slab)
        spec = f.spectrum().clone()
        f.apply_green(0.9)
        f.backward(out)
        torch.cuda.synchronize()
        return spec, out

    f0 = pmhip.FFT(ng, prec)
    spec0, out0 = run(f0)
    f0.close()
    tag = "fp32" if prec == pmhip.PM_FP32 else "fp64"
    for v in range(1, pmhip.PM_FFT_VARIANTS):
        w = tmp_path / f"w{v}.dat"
        w.write_text(f"pm-wisdom 1\n{ng} {tag} {v} {v} {v} {v} {v} 1.0\n")
        pmhip.wisdom_import(w)
        f = pmhip.FFT(ng, prec)
        assert f.shape == [v] * pmhip.PM_FFT_SLOTS
        spec, out = run(f)
        f.close()
        assert torch.equal(spec, spec0), v
        assert torch.equal(out, out0), v


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_plan_tune_caches_and_round_trips(dev, tmp_path, prec):
    import torch

    ng = 128
    print(f"case plan tune prec={prec}", flush=True)
    pos, vel = _particles(ng, dev, 54)
    npart = pos[0].numel()
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        ms = plan.tune(repeats=3)
        shape = plan.fft_shape
        assert ms > 0
        assert all(0 <= v < pmhip.PM_FFT_VARIANTS for v in shape)
        assert pmhip.wisdom_lookup(ng, prec) == shape
        with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as later:
            assert later.fft_shape == shape
        # a tuned plan steps exactly as a default-shape one
        with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as dflt:
            dflt.fft_shape = [0] * pmhip.PM_FFT_SLOTS
            a = ([p.clone() for p in pos], [v.clone() for v in vel])
            b = ([p.clone() for p in pos], [v.clone() for v in vel])
            for s in range(3):
                flag = pmhip.PM_DEPOSIT_SORTED | (pmhip.PM_STEP_REUSE_BINS if s else 0)
                plan.step(*a, 0.01, 0.02, flag)
                dflt.step(*b, 0.01, 0.02, flag)
            torch.cuda.synchronize()
            for x, y in zip(a[0] + a[1], b[0] + b[1]):
                assert torch.equal(x, y)
    w = tmp_path / "wisdom.dat"
    pmhip.wisdom_export(w)
    pmhip.wisdom_forget()
    assert pmhip.wisdom_lookup(ng, prec) is None
    pmhip.wisdom_import(w)
    assert pmhip.wisdom_lookup(ng, prec) == shape


def test_fft_tune(dev):
    import torch

    ng = 128
    print("case fft tune", flush=True)
    slab = torch.zeros(ng ** 3, dtype=torch.float64, device=dev)
    f = pmhip.FFT(ng, pmhip.PM_FP64)
    ms = f.tune(slab, repeats=3)
    shape = f.shape
    f.close()
    assert ms > 0
    assert all(0 <= v < pmhip.PM_FFT_VARIANTS for v in shape)
    assert pmhip.wisdom_lookup(ng, pmhip.PM_FP64) == shape


def test_pm_main_writes_then_reads_its_wisdom_file(dev, tmp_path):
    """First run of bin/pm_main: no wisdom_64_hip.dat, the FFT class tunes
    and exports it; the second run imports it (the file is left as it was);
    both runs give the same momenta up to the order of pm_main's FP64 atomic
    deposit (the reference-layout deposit sums with float atomics)."""
    assert EXE.exists(), "build first"
    print("case driver wisdom", flush=True)
    env = dict(os.environ, PM_WISDOM_DIR=str(tmp_path))
    dumps = []
    for k in range(2):
        out = tmp_path / f"p{k}.bin"
        r = subprocess.run([str(EXE), "64", "32", "--warmup", "1", "--loop", "1", "--dump", str(out)],
                           capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode == 0, r.stderr
        w = tmp_path / "wisdom_64_hip.dat"
        assert w.exists()
        text = w.read_text()
        if k == 0:
            first = (text, w.stat().st_mtime_ns)
        else:
            assert (text, w.stat().st_mtime_ns) == first
        lines = text.splitlines()
        assert lines[0] == "pm-wisdom 1"
        assert any(ln.split()[:2] == ["64", "fp64"] for ln in lines[1:])
        dumps.append(np.fromfile(out, dtype=np.float64))
    assert np.max(np.abs(dumps[0] - dumps[1])) < 1e-11 * np.abs(dumps[0]).max()


def test_copy_rate_is_a_plausible_hbm_rate(dev):
    """pm_copy_rate times a plain 1 GiB device copy: on an MI355X (8 TB/s
    nameplate) read + write lands between 2 and 8 TB/s."""
    print("case copy rate", flush=True)
    gbs = pmhip.copy_rate(1 << 30, 10)
    print(f"copy rate {gbs:.0f} GB/s", flush=True)
    assert 2000.0 < gbs < 8000.0
