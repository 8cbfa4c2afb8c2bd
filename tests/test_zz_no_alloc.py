"""pm_deposit_sorted_noalloc -- the sorted CIC deposit whose work buffers live
in the library's fixed device array instead of device allocations -- against
pm_deposit(plan, ..., PM_DEPOSIT_SORTED), through the C ABI: bit for bit on two
calls in a row, the second with fewer particles on a smaller mesh; sizes
beyond the array are rejected without touching the output."""
import ctypes

import numpy as np
import pytest

import pmhip

pytestmark = pytest.mark.gpu


def _positions(ng, npart, seed):
    """Half uniform, half in a few dense clumps (many particles per cell),
    plus particles exactly on the box faces 0 and nextafter(ng, 0)."""
    rng = np.random.default_rng(seed)
    nu = npart // 2
    centres = rng.uniform(0, ng, (8, 3))
    pick = rng.integers(0, 8, npart - nu)
    pos = []
    for a in range(3):
        u = rng.uniform(0, ng, nu)
        c = np.mod(centres[pick, a] + rng.standard_normal(npart - nu) * 0.7, ng)
        v = np.concatenate([u, c])
        v[v >= ng] = 0.0
        v[:4] = [0.0, np.nextafter(ng, 0), 0.0, np.nextafter(ng, 0)]
        pos.append(v)
    return pos


def _noalloc(pos, ng, prec, out):
    return pmhip.lib().pm_deposit_sorted_noalloc(*(pmhip.dptr(t) for t in pos), pos[0].numel(), ng, prec,
                                                 pmhip.dptr(out), pmhip.current_stream())


def _plan_density(torch, pos, ng, prec):
    with pmhip.Plan(ng, pos[0].numel(), prec) as plan:
        plan.deposit(pos, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
        return plan.density().clone()


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_noalloc_deposit_matches_plan_deposit_on_two_calls(dev, prec):
    import torch

    dt = pmhip.mesh_dtype(prec)
    cases = [(128, 2_000_000, 31), (64, 70_001, 32)]  # the second call is the smaller one
    inputs, refs, outs = [], [], []
    for ng, npart, seed in cases:
        pos = [torch.tensor(v, device=dev) for v in _positions(ng, npart, seed)]
        inputs.append(pos)
        refs.append(_plan_density(torch, pos, ng, prec))
        # NaN everywhere: the deposit must write every node
        outs.append(torch.full((ng, ng, ng), float("nan"), dtype=dt, device=dev))
    assert pmhip.range_errors() == 0
    torch.cuda.synchronize()
    free = []
    for (ng, npart, seed), pos, out in zip(cases, inputs, outs):
        print(f"case noalloc deposit prec={prec} ng={ng} np={npart}", flush=True)
        pmhip.check(_noalloc(pos, ng, prec, out))
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
    # the second call allocated no device memory
    assert free[1] == free[0]
    for (ng, npart, _), out, ref in zip(cases, outs, refs):
        assert torch.equal(out, ref), f"ng={ng} np={npart}"
        assert abs(float(out.double().sum()) - ng ** 3) < 1e-6 * ng ** 3
    assert pmhip.range_errors() == 0


def test_noalloc_deposit_rejects_sizes_beyond_the_array(dev):
    import torch

    ng = 256
    npart = 4 << 20  # 12 B per particle alone is 48 MiB
    pos = [torch.zeros(npart, dtype=torch.float64, device=dev) for _ in range(3)]
    out = torch.full((ng, ng, ng), -7.0, dtype=torch.float32, device=dev)
    print(f"case noalloc deposit reject ng={ng} np={npart}", flush=True)
    rc = _noalloc(pos, ng, pmhip.PM_FP32, out)
    assert rc == pmhip.PM_ERR_ARG
    assert "scratch" in pmhip.last_error()
    torch.cuda.synchronize()
    assert bool((out == -7.0).all())
    # the largest mesh whose shells alone exceed the array
    out2 = torch.zeros(1, dtype=torch.float32, device=dev)
    assert _noalloc([p[:10] for p in pos], 512, pmhip.PM_FP32, out2) == pmhip.PM_ERR_ARG
    assert pmhip.lib().pm_deposit_sorted_noalloc(None, None, None, 10, 64, 0, None, None) == pmhip.PM_ERR_ARG
