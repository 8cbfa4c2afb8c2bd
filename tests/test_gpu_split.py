"""The split-precision particle state (pm_split_*): positions as an integer
cell plus a 2^-S-cell fixed-point fraction (S = 32 - log2 ng), FP32
velocities, resident in the plan.  Checked against the FP64-position path
(pm_step, the same deposit, solve and force) and the FP64 oracle."""
import math

import numpy as np
import pytest

import pmhip
import pyoracle

pytestmark = pytest.mark.gpu


def _shift(ng):
    return 32 - int(round(math.log2(ng)))


def _on_grid(pos, ng):
    """Positions rounded to the nearest 2^-S cells, as pm_split_load does
    (x * 2^S is exact; 2^32 wraps to 0)."""
    import torch

    s = 2.0 ** _shift(ng)
    return [torch.remainder(torch.round(p * s), 2.0 ** 32) / s for p in pos]


def _ics(nrow, ng, rms, seed, dev, vel_factor=math.sqrt(4.0 * math.pi)):
    import torch

    noise = np.random.default_rng(seed).standard_normal(nrow ** 3)
    return pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), rms, vel_factor)


def _periodic_diff(a, b, ng):
    d = a - b
    return d - ng * (d / ng).round()


def test_split_load_store_round_trip(dev):
    """pm_split_store gives the positions rounded to 2^-S cells exactly (a
    periodic image of them at the upper box edge) and the velocities rounded
    to FP32; positions outside the box are counted and wrapped in."""
    import torch

    print("case split round trip ng=64", flush=True)
    ng, npart = 64, 4096 + 512
    rng = np.random.default_rng(41)
    x = rng.uniform(0.0, ng, (3, npart))
    x[0, :4] = [0.0, np.nextafter(ng, 0.0), ng - 2.0 ** -27, 0.5 * 2.0 ** -26]  # edges and a tie
    v = rng.standard_normal((3, npart)) * 10.0
    pos = [torch.tensor(a, device=dev) for a in x]
    vel = [torch.tensor(a, device=dev) for a in v]
    out_p = [torch.empty_like(p) for p in pos]
    out_v = [torch.empty_like(p) for p in vel]
    pmhip.range_errors()
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos, vel)
        plan.split_store(out_p, out_v)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    s = 2.0 ** _shift(ng)
    for a, b in zip(out_p, pos):
        q = np.round(b.cpu().numpy() * s)
        expect = np.where(q >= 2.0 ** 32, q - 2.0 ** 32, q) / s
        assert np.array_equal(a.cpu().numpy(), expect)
    assert out_p[0][1].item() == 0.0  # just below the edge rounds to the periodic 0
    for a, b in zip(out_v, vel):
        assert torch.equal(a, b.float().double())
    # outside the box: counted, wrapped
    pos[0][5] = -1.25
    pos[1][6] = float("nan")
    pos[2][7] = ng + 3.5
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos, vel)
        plan.split_store(out_p, out_v)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 3
    assert out_p[0][5].item() == ng - 1.25 and out_p[1][6].item() == 0.0 and out_p[2][7].item() == 3.5


def test_split_rejects_unloaded_state(dev):
    import torch

    pos = [torch.zeros(64, dtype=torch.float64, device=dev) for _ in range(3)]
    with pmhip.Plan(16, 64, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        with pytest.raises(pmhip.PMError, match="no particles loaded"):
            plan.split_step(0.1, 0.1)
        with pytest.raises(pmhip.PMError, match="no particles loaded"):
            plan.split_store(pos, pos)


@pytest.mark.parametrize("ng,nrow", [(16, 16), (64, 64), (256, 128)])
def test_split_step_matches_fp64_position_step(dev, ng, nrow):
    """From positions on the 2^-S grid, one split step deposits, solves and
    interpolates exactly as pm_step with FP64 positions (density, potential
    and the kicked velocities before their FP32 rounding agree); the kick in
    FP32 differs by the FP32 rounding of v, the drift by the rounding of
    the displacement to 2^-S cells."""
    import torch

    print(f"case split step vs pm_step ng={ng} nrow={nrow}", flush=True)
    npart = nrow ** 3
    pos0, vel0 = _ics(nrow, ng, 1.5, 43, dev, vel_factor=3.0)
    pos0 = _on_grid(pos0, ng)
    vel0 = [v.float().double() for v in vel0]
    kick, drift = 0.02, 0.05
    pmhip.range_errors()
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as a, \
            pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as b:
        pa, va = [p.clone() for p in pos0], [v.clone() for v in vel0]
        a.step(pa, va, kick, drift, pmhip.PM_DEPOSIT_SORTED)
        b.split_load(pos0, vel0)
        b.split_step(kick, drift)
        pb, vb = [torch.empty_like(p) for p in pos0], [torch.empty_like(v) for v in vel0]
        b.split_store(pb, vb)
        torch.cuda.synchronize()
        assert torch.equal(a.density(), b.density())
        assert torch.equal(a.potential(), b.potential())
    assert pmhip.range_errors() == 0
    vmax = max(float(v.abs().max()) for v in va)
    dv = max(float((x - y).abs().max()) for x, y in zip(va, vb))
    dx = max(float(_periodic_diff(x, y, ng).abs().max()) for x, y in zip(pa, pb))
    print(f"max |dv| {dv:.3e} (max |v| {vmax:.2f}), max |dx| {dx:.3e} cells", flush=True)
    assert dv <= 2.0 ** -23 * vmax
    # rounding to 2^-S cells, the velocity difference, and the FP32 product
    # v * (drift * 2^S) (two roundings of 2^-24 relative)
    assert dx <= 2.0 ** -(_shift(ng) + 1) + drift * dv + 2.0 ** -22 * vmax * drift


def test_split_restaged_equals_staged(dev):
    """Continuing from the records the interpolation staged and restaging
    from the positions (a store and reload between steps) give the same
    particles bit for bit; so does a drift that packs every particle into
    one eighth of the box, which overflows the buckets and makes the next
    step restage behind the overflow flag."""
    import torch

    print("case split staged vs restaged", flush=True)
    for ng, nrow, dt, pack in [(64, 64, 0.01, False), (64, 32, 0.1, True)]:
        npart = nrow ** 3
        pos0, vel0 = _ics(nrow, ng, 1.0, 45, dev, vel_factor=2.0)
        if pack:
            c = ng / 2.0
            vel0 = [(c - p) * (0.5 / dt) for p in pos0]  # x -> (x + c) / 2 in one drift
        sched = [(0.5 * dt, dt), (dt, dt), (dt, dt), (0.5 * dt, 0.0)]
        with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as a, \
                pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as b:
            a.split_load(pos0, vel0)
            b.split_load(pos0, vel0)
            p, v = [torch.empty_like(t) for t in pos0], [torch.empty_like(t) for t in vel0]
            for k, d in sched:
                a.split_step(k, d)
                b.split_step(k, d)
                b.split_store(p, v)
                b.split_load(p, v)
            pa, va = [torch.empty_like(t) for t in pos0], [torch.empty_like(t) for t in vel0]
            a.split_store(pa, va)
            b.split_store(p, v)
            torch.cuda.synchronize()
        assert pmhip.range_errors() == 0
        for x, y in zip(pa + va, p + v):
            assert torch.equal(x, y), (ng, nrow, pack)
        with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as c_:
            c_.split_load(pos0, vel0)
            c_.split_run_kdk(3, dt)
            a_run = [torch.empty_like(t) for t in pos0 + vel0]
            c_.split_store(a_run[:3], a_run[3:])
            torch.cuda.synchronize()
        for x, y in zip(pa + va, a_run):
            assert torch.equal(x, y), "split_run_kdk is the step schedule above"


def test_split_zero_steps_is_identity(dev):
    import torch

    ng, nrow = 32, 16
    pos0, vel0 = _ics(nrow, ng, 0.5, 47, dev)
    pos0 = _on_grid(pos0, ng)
    vel0 = [v.float().double() for v in vel0]
    with pmhip.Plan(ng, nrow ** 3, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos0, vel0)
        plan.split_run_kdk(0, 0.01)
        p, v = [torch.empty_like(t) for t in pos0], [torch.empty_like(t) for t in vel0]
        plan.split_store(p, v)
        torch.cuda.synchronize()
    for x, y in zip(p + v, pos0 + vel0):
        assert torch.equal(x, y)


@pytest.mark.parametrize("ng,nrow", [(64, 64), (128, 128)])
def test_split_kdk_vs_fp64_oracle(dev, ng, nrow):
    """10 KDK steps of clustered ICs: split state against the FP64 oracle on
    the same inputs, final positions within 1e-5 cells rms (the north-star
    budget is 1e-4) and 1e-4 max."""
    import torch

    print(f"case split kdk vs oracle ng={ng}", flush=True)
    nsteps, dt = 10, 0.005
    npart = nrow ** 3
    pos0, vel0 = _ics(nrow, ng, 1.5, 49, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos0, vel0)
        plan.split_run_kdk(nsteps, dt)
        p = [torch.empty_like(t) for t in pos0]
        v = [torch.empty_like(t) for t in vel0]
        plan.split_store(p, v)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    ph = [t.cpu().numpy().copy() for t in pos0]
    vh = [t.cpu().numpy().copy() for t in vel0]
    pyoracle.set_threads(min(16, pyoracle.max_threads()))
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, 1.0)
    d2 = sum(_periodic_diff(a.cpu().numpy(), b, ng) ** 2 for a, b in zip(p, ph))
    rms, mx = float(np.sqrt(d2.mean())), float(np.sqrt(d2.max()))
    print(f"split vs oracle after {nsteps} steps: rms {rms:.3e} max {mx:.3e} cells", flush=True)
    assert rms < 1e-5
    assert mx < 1e-4


@pytest.mark.parametrize("prec,tol_rms,tol_max", [(pmhip.PM_FP32, 1e-4, 1e-3), (pmhip.PM_FP64, 1e-9, 1e-8)])
def test_split_long_high_mobility_vs_fp64_oracle(dev, prec, tol_rms, tol_max):
    """100 KDK steps in which up to a fifth of the particles change tile every
    step (velocities up to 40 cells per unit time, dt = 0.01; gravity weakened
    by a Green's scale of 1e-3 so that FP32 and FP64 trajectories stay
    comparable): the outboxes, the hole lists, the arrivals into holes and onto
    bucket ends and the rebuilds behind full buckets all run many times.
    Every particle comes back exactly once (outputs start as NaN), and the
    positions match the FP64 oracle of the same run."""
    import torch

    print(f"case split long high-mobility run ng=64 np=64^3 100 steps prec={prec}", flush=True)
    ng, nrow, nsteps, dt, scale = 64, 64, 100, 0.01, 1e-3
    npart = nrow ** 3
    rng = np.random.default_rng(53)
    pos0 = [torch.tensor(rng.uniform(0.0, ng, npart), device=dev) for _ in range(3)]
    pos0 = _on_grid(pos0, ng)
    vel0 = [torch.tensor(rng.uniform(-40.0, 40.0, npart), device=dev) for _ in range(3)]
    if prec == pmhip.PM_FP32:
        vel0 = [v.float().double() for v in vel0]
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, scale) as plan:
        plan.split_load(pos0, vel0)
        plan.split_run_kdk(nsteps, dt)
        p = [torch.full_like(t, float("nan")) for t in pos0]
        v = [torch.full_like(t, float("nan")) for t in vel0]
        plan.split_store(p, v)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    for t in p + v:
        assert not bool(torch.isnan(t).any()), "a particle was lost"
    ph = [t.cpu().numpy().copy() for t in pos0]
    vh = [t.cpu().numpy().copy() for t in vel0]
    pyoracle.set_threads(min(16, pyoracle.max_threads()))
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, scale)
    d2 = sum(_periodic_diff(a.cpu().numpy(), b, ng) ** 2 for a, b in zip(p, ph))
    rms, mx = float(np.sqrt(d2.mean())), float(np.sqrt(d2.max()))
    moved = float(np.sqrt(sum(_periodic_diff(a.cpu().numpy(), b.cpu().numpy(), ng) ** 2
                              for a, b in zip(p, pos0)).mean()))
    print(f"after {nsteps} steps: rms displacement {moved:.1f} cells; vs oracle rms {rms:.3e} max {mx:.3e} cells",
          flush=True)
    assert moved > 5.0  # the particles did cross many tiles
    assert rms < tol_rms
    assert mx < tol_max


# ------------------------------------------------------------ FP64 mesh
def _shift64(ng):
    return 64 - int(round(math.log2(ng)))


def test_split64_load_store_round_trip(dev):
    """With an FP64 mesh the split state holds u64 positions at 2^-(64 - log2 ng)
    cells and FP64 velocities: any FP64 position in the box and any velocity
    come back bit for bit."""
    import torch

    print("case split64 round trip ng=64", flush=True)
    ng, npart = 64, 20000
    rng = np.random.default_rng(61)
    pos = [torch.tensor(rng.uniform(0.0, ng, npart), device=dev) for _ in range(3)]
    vel = [torch.tensor(rng.normal(0.0, 3.0, npart), device=dev) for _ in range(3)]
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos, vel)
        p = [torch.full_like(t, float("nan")) for t in pos]
        v = [torch.full_like(t, float("nan")) for t in vel]
        plan.split_store(p, v)
        torch.cuda.synchronize()
    for x, y in zip(p + v, pos + vel):
        assert torch.equal(x, y)


@pytest.mark.parametrize("ng,nrow", [(16, 16), (64, 64), (256, 128)])
def test_split64_step_matches_fp64_position_step(dev, ng, nrow):
    """One split step with an FP64 mesh deposits and solves exactly as pm_step
    with FP64 positions (density and potential bit for bit: the in-tile
    offsets are truncated to 2^-38 cells in both), and kicks and drifts to
    the rounding of the positions (2^-(64 - log2 ng) cells against an FP64
    add)."""
    import torch

    print(f"case split64 step vs pm_step ng={ng} nrow={nrow}", flush=True)
    npart = nrow ** 3
    pos0, vel0 = _ics(nrow, ng, 1.5, 43, dev, vel_factor=3.0)
    kick, drift = 0.02, 0.05
    pmhip.range_errors()
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as a, \
            pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as b:
        pa, va = [p.clone() for p in pos0], [v.clone() for v in vel0]
        a.step(pa, va, kick, drift, pmhip.PM_DEPOSIT_SORTED)
        b.split_load(pos0, vel0)
        b.split_step(kick, drift)
        pb, vb = [torch.empty_like(p) for p in pos0], [torch.empty_like(v) for v in vel0]
        b.split_store(pb, vb)
        torch.cuda.synchronize()
        assert torch.equal(a.density(), b.density())
        assert torch.equal(a.potential(), b.potential())
    assert pmhip.range_errors() == 0
    vmax = max(float(v.abs().max()) for v in va)
    dv = max(float((x - y).abs().max()) for x, y in zip(va, vb))
    dx = max(float(_periodic_diff(x, y, ng).abs().max()) for x, y in zip(pa, pb))
    print(f"max |dv| {dv:.3e} (max |v| {vmax:.2f}), max |dx| {dx:.3e} cells", flush=True)
    assert dv <= 1e-14 * vmax
    assert dx <= 2.0 ** -(_shift64(ng) + 1) + 2.0 ** -52 * ng + drift * dv


@pytest.mark.parametrize("ng,nrow", [(64, 64), (128, 128)])
def test_split64_kdk_vs_fp64_oracle(dev, ng, nrow):
    """10 KDK steps of clustered ICs with an FP64 mesh: the split state against
    the FP64 oracle on the same inputs, final positions within 1e-10 cells
    rms and 1e-9 max."""
    import torch

    print(f"case split64 kdk vs oracle ng={ng}", flush=True)
    nsteps, dt = 10, 0.005
    npart = nrow ** 3
    pos0, vel0 = _ics(nrow, ng, 1.5, 49, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos0, vel0)
        plan.split_run_kdk(nsteps, dt)
        p = [torch.empty_like(t) for t in pos0]
        v = [torch.empty_like(t) for t in vel0]
        plan.split_store(p, v)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    ph = [t.cpu().numpy().copy() for t in pos0]
    vh = [t.cpu().numpy().copy() for t in vel0]
    pyoracle.set_threads(min(16, pyoracle.max_threads()))
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, 1.0)
    d2 = sum(_periodic_diff(a.cpu().numpy(), b, ng) ** 2 for a, b in zip(p, ph))
    rms, mx = float(np.sqrt(d2.mean())), float(np.sqrt(d2.max()))
    print(f"split64 vs oracle after {nsteps} steps: rms {rms:.3e} max {mx:.3e} cells", flush=True)
    assert rms < 1e-10
    assert mx < 1e-9
