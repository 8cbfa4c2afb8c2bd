"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle and
the reference's golden fixtures.

Tolerances: FP64 mesh paths agree with the oracle to ~1e-12 relative (only
summation order and FMA contraction differ); FP32 mesh paths are bounded by
FP32 rounding of the mesh, FFT and force (relative 1e-5 on the potential,
1e-4 on per-particle forces relative to the largest force).
"""
import numpy as np
import pytest

import pmhip
import pyoracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def to_dev(arrs, dev):
    import torch

    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev) for a in arrs]


def host(t):
    return t.detach().cpu().numpy()


# -------------------------------------------------------------------- FFT
@pytest.mark.parametrize("prec,tol", [(pmhip.PM_FP64, 1e-12), (pmhip.PM_FP32, 2e-6)])
@pytest.mark.parametrize("ng", [8, 16, 32, 64, 128])
def test_fft_forward_backward_vs_numpy(dev, ng, prec, tol):
    import torch

    print(f"case fft ng={ng} prec={prec}", flush=True)
    rng = np.random.default_rng(ng)
    r = rng.standard_normal((ng, ng, ng))
    dt = pmhip.mesh_dtype(prec)
    slab = torch.tensor(r, dtype=dt, device=dev)
    f = pmhip.FFT(ng, prec)
    f.forward(slab)
    torch.cuda.synchronize()
    spec = host(f.spectrum())  # [kz][ky][kx]
    ref = np.fft.rfftn(r).transpose(2, 1, 0)
    assert rel(spec, ref) < tol
    out = torch.empty_like(slab)
    f.backward(out)
    torch.cuda.synchronize()
    assert rel(host(out), r) < tol
    f.close()


def test_fft_apply_green_matches_oracle(dev):
    import torch

    ng, a = 32, 0.9
    rng = np.random.default_rng(3)
    d = rng.standard_normal((ng, ng, ng))
    d -= d.mean()
    slab = torch.tensor(d, device=dev)
    f = pmhip.FFT(ng, pmhip.PM_FP64)
    f.forward(slab)
    f.apply_green(a)
    out = torch.empty_like(slab)
    f.backward(out)
    torch.cuda.synchronize()
    assert rel(host(out), pyoracle.solve(d, pyoracle.GREEN_REF_KH, 1.0 / a)) < 1e-12
    f.close()


# -------------------------------------------------- reference-layout operators
def test_reference_layout_step_vs_golden(dev, golden_dir):
    """The drop-in operators in the reference's order reproduce the
    reference's own one-kick fixture."""
    import torch

    lib = pmhip.lib()
    for name in ("ref_kick_ng16_np4096", "ref_kick_ng16_np1728", "ref_kick_ng32_np4096"):
        print(f"case reference layout {name}", flush=True)
        f = np.load(golden_dir / f"{name}.npz")
        ng, a = int(f["ng"]), float(f["a"])
        pos = to_dev((f["posx"], f["posy"], f["posz"]), dev)
        npart = pos[0].numel()
        n1, n3 = ng + 1, ng + 3
        cube = torch.empty(n1 ** 3, dtype=torch.float64, device=dev)
        recv = torch.empty(n1 ** 3, dtype=torch.float64, device=dev)
        slab = torch.empty(ng ** 3, dtype=torch.float64, device=dev)
        yz = torch.empty(ng * n3 * n3, dtype=torch.float64, device=dev)
        halo = torch.empty(n3 ** 3, dtype=torch.float64, device=dev)
        grads = [torch.empty(n1 ** 3, dtype=torch.float64, device=dev) for _ in range(3)]
        mom = [torch.zeros(npart, dtype=torch.float64, device=dev) for _ in range(3)]
        st = pmhip.current_stream()
        pmhip.check(lib.pm_init_buf(cube.data_ptr(), n1, n1, n1, st))
        pmhip.check(lib.pm_deposit_particles(*(p.data_ptr() for p in pos), npart, cube.data_ptr(), n1, n1, n1, 0, 0, 0,
                                             ng ** 3 / npart, st))
        pmhip.check(lib.pm_transpose_slab(cube.data_ptr(), recv.data_ptr(), ng, st))
        pmhip.check(lib.pm_reorder_slab(recv.data_ptr(), slab.data_ptr(), ng, ng, st))
        fft = pmhip.FFT(ng, pmhip.PM_FP64)
        fft.forward(slab)
        fft.apply_green(a)
        fft.backward(slab)
        pmhip.check(lib.pm_reorder_slab_back(slab.data_ptr(), yz.data_ptr(), ng, ng, st))
        pmhip.check(lib.pm_transpose_slab_back(yz.data_ptr(), halo.data_ptr(), ng, st))
        pmhip.check(lib.pm_centered_derivatives(halo.data_ptr(), *(g.data_ptr() for g in grads), n3, n3, n3, st))
        pmhip.check(lib.pm_update_particles(*(p.data_ptr() for p in pos), *(g.data_ptr() for g in grads),
                                            *(m.data_ptr() for m in mom), a, npart, n1, n1, n1, 0, 0, 0, st))
        torch.cuda.synchronize()
        assert pmhip.range_errors() == 0
        if "slab" in f:
            assert rel(host(slab), f["phi"].ravel()) < 1e-12
        for m, key in zip(mom, ("px", "py", "pz")):
            assert rel(host(m), f[key]) < 1e-11
        fft.close()


def test_reference_layout_operators_vs_oracle(dev):
    import torch

    lib = pmhip.lib()
    ng, npart = 16, 2500
    rng = np.random.default_rng(9)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    pos = to_dev(ph, dev)
    p_ref = [np.zeros(npart) for _ in range(3)]
    ref = pyoracle.reference_step(ng, ph, p_ref)
    n1, n3 = ng + 1, ng + 3
    st = pmhip.current_stream()
    cube = torch.empty(n1 ** 3, dtype=torch.float64, device=dev)
    pmhip.check(lib.pm_init_buf(cube.data_ptr(), n1, n1, n1, st))
    pmhip.check(lib.pm_deposit_particles(*(p.data_ptr() for p in pos), npart, cube.data_ptr(), n1, n1, n1, 0, 0, 0,
                                         ng ** 3 / npart, st))
    torch.cuda.synchronize()
    assert rel(host(cube), ref["cube"]) < 1e-13
    slab = torch.full((ng * ng * (ng + 2),), 7.0, dtype=torch.float64, device=dev)  # padded stride
    pmhip.check(lib.pm_reorder_slab(cube.data_ptr(), slab.data_ptr(), ng, ng + 2, st))
    torch.cuda.synchronize()
    s = host(slab).reshape(ng, ng, ng + 2)
    assert rel(s[:, :, :ng].ravel(), ref["slab"]) < 1e-13
    assert np.all(s[:, :, ng:] == 7.0)  # padding untouched
    phi = torch.tensor(ref["phi"], device=dev)
    halo = torch.empty(n3 ** 3, dtype=torch.float64, device=dev)
    yz = torch.empty(ng * n3 * n3, dtype=torch.float64, device=dev)
    pmhip.check(lib.pm_reorder_slab_back(phi.data_ptr(), yz.data_ptr(), ng, ng, st))
    torch.cuda.synchronize()
    # the y/z halo block alone is what reorder_slab_back.cpp:11-21 builds
    assert np.array_equal(host(yz), ref["halo"].reshape(n3, n3, n3)[1:ng + 1].ravel())
    pmhip.check(lib.pm_transpose_slab_back(yz.data_ptr(), halo.data_ptr(), ng, st))
    grads = [torch.empty(n1 ** 3, dtype=torch.float64, device=dev) for _ in range(3)]
    pmhip.check(lib.pm_centered_derivatives(halo.data_ptr(), *(g.data_ptr() for g in grads), n3, n3, n3, st))
    torch.cuda.synchronize()
    assert np.array_equal(host(halo), ref["halo"])
    for g, k in zip(grads, ("dx", "dy", "dz")):
        assert np.array_equal(host(g), ref[k])


def test_range_errors_counted(dev):
    import torch

    lib = pmhip.lib()
    pmhip.range_errors()
    ng = 8
    n1 = ng + 1
    pos = to_dev([np.array([1.0, 8.0, -0.5, 3.0]), np.array([1.0, 1.0, 1.0, np.nan]), np.array([1.0] * 4)], dev)
    cube = torch.zeros(n1 ** 3, dtype=torch.float64, device=dev)
    pmhip.check(lib.pm_deposit_particles(*(p.data_ptr() for p in pos), 4, cube.data_ptr(), n1, n1, n1, 0, 0, 0, 1.0,
                                         pmhip.current_stream()))
    assert pmhip.range_errors() == 3
    assert abs(float(cube.sum()) - 1.0) < 1e-15


# ----------------------------------------------------------- fused plan path
@pytest.mark.parametrize("ng,npart", [(16, 4096), (32, 20000), (64, 100000)])
def test_plan_deposit_solve_force_fp64(dev, ng, npart):
    import torch

    print(f"case plan fp64 ng={ng} np={npart}", flush=True)
    rng = np.random.default_rng(ng + 1)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    pos = to_dev(ph, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_REF_KH, 1.0 / 0.9) as plan:
        plan.deposit(pos)
        plan.solve()
        g = [torch.empty(npart, dtype=torch.float64, device=dev) for _ in range(3)]
        plan.interp_force(pos, g)
        torch.cuda.synchronize()
        delta_ref = pyoracle.deposit_periodic(ph, ng)
        assert rel(host(plan.density()) - 1.0, delta_ref) < 1e-12
        phi_ref = pyoracle.solve(delta_ref, pyoracle.GREEN_REF_KH, 1.0 / 0.9)
        assert rel(host(plan.potential()), phi_ref) < 1e-11
        g_ref = pyoracle.interp_force(ph, phi_ref)
        for a, b in zip(g, g_ref):
            assert rel(host(a), b) < 1e-10


@pytest.mark.parametrize("ng,npart", [(16, 4096), (64, 100000)])
def test_plan_fp32_within_fp32_bounds(dev, ng, npart):
    import torch

    print(f"case plan fp32 ng={ng} np={npart}", flush=True)
    rng = np.random.default_rng(ng + 2)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    pos = to_dev(ph, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.deposit(pos)
        plan.solve()
        g = [torch.empty(npart, dtype=torch.float64, device=dev) for _ in range(3)]
        plan.interp_force(pos, g)
        torch.cuda.synchronize()
        delta_ref = pyoracle.deposit_periodic(ph, ng)
        rho = host(plan.density()).astype(np.float64)
        assert np.max(np.abs(rho - 1.0 - delta_ref)) < 2e-6 * max(1.0, np.abs(delta_ref).max())
        phi_ref = pyoracle.solve(delta_ref, pyoracle.GREEN_CONT_K2, 1.0)
        assert rel(host(plan.potential()), phi_ref) < 1e-5
        g_ref = pyoracle.interp_force(ph, phi_ref)
        gmax = max(np.abs(v).max() for v in g_ref)
        for a, b in zip(g, g_ref):
            assert np.max(np.abs(host(a) - b)) < 1e-4 * gmax


def test_kdk_config1_fp64_matches_oracle(dev):
    """64^3 particles from the reference's own random stream on a 64^3 mesh,
    10 KDK steps, dt = 0.01, continuous -1/k^2 Green's function (the
    reference's CPU-runnable case, in unit-box units)."""
    import torch

    ng, nrow, nsteps, dt = 64, 64, 10, 0.01
    npart = nrow ** 3
    ph = list(pyoracle.init_particles(npart, ng))
    vh = [np.zeros(npart) for _ in range(3)]
    pos = to_dev(ph, dev)
    vel = to_dev(vh, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(pos, vel, nsteps, dt)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, 1.0)
    for a, b in zip(pos, ph):
        d = host(a) - b
        d -= ng * np.round(d / ng)  # periodic distance
        assert np.max(np.abs(d)) < 1e-8
    vmax = max(np.abs(v).max() for v in vh)
    for a, b in zip(vel, vh):
        assert np.max(np.abs(host(a) - b)) < 1e-8 * vmax


@pytest.mark.parametrize("nrow,ng", [(16, 16), (32, 64)])
def test_zeldovich_ic_vs_numpy(dev, nrow, ng):
    import torch

    print(f"case zeldovich nrow={nrow} ng={ng}", flush=True)
    noise = np.random.default_rng(nrow).standard_normal((nrow, nrow, nrow))
    pos, vel = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise.ravel(), device=dev), 0.7, 2.0)
    ref_pos, ref_vel = pyoracle.zeldovich_ic(noise, ng, 0.7, 2.0)
    for a, b in zip(pos, ref_pos):
        d = host(a) - b
        d -= ng * np.round(d / ng)
        assert np.max(np.abs(d)) < 1e-10
    for a, b in zip(vel, ref_vel):
        assert np.max(np.abs(host(a) - b)) < 1e-10
    disp = np.sqrt(np.mean(sum(host(v) ** 2 for v in vel))) / 2.0
    assert abs(disp - 0.7) < 1e-9


def test_zeldovich_ic_bitwise_reproducible(dev):
    """The rms normalisation is summed in a fixed order (per-workgroup
    partials, then one ordered reduction): the same noise gives the same
    particles bit for bit, also at a size with thousands of partials."""
    import torch

    nrow, ng = 128, 128
    print(f"case zeldovich reproducible nrow={nrow}", flush=True)
    noise = torch.tensor(np.random.default_rng(5).standard_normal(nrow ** 3), device=dev)
    a = pmhip.zeldovich_ic(nrow, ng, noise, 1.5, 3.0)
    b = pmhip.zeldovich_ic(nrow, ng, noise, 1.5, 3.0)
    for u, v in zip(a[0] + a[1], b[0] + b[1]):
        assert torch.equal(u, v)


# --------------------------------------------------------- sorted deposit
def _plan_density(dev, ng, ph, prec, variant):
    import torch

    pos = to_dev(ph, dev)
    with pmhip.Plan(ng, ph[0].size, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.deposit(pos, variant)
        torch.cuda.synchronize()
        return host(plan.density()).astype(np.float64)


@pytest.mark.parametrize("ng,npart", [(8, 3000), (16, 4096), (32, 30000), (64, 250000)])
def test_sorted_deposit_fp64_vs_oracle(dev, ng, npart):
    print(f"case sorted deposit fp64 ng={ng} np={npart}", flush=True)
    rng = np.random.default_rng(100 + ng)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    rho = _plan_density(dev, ng, ph, pmhip.PM_FP64, pmhip.PM_DEPOSIT_SORTED)
    ref = pyoracle.deposit_periodic(ph, ng) + 1.0
    # contributions are rounded to 2^-32 mass units before the exact sum
    assert np.max(np.abs(rho - ref)) < 1e-8
    assert pmhip.range_errors() == 0


def test_sorted_deposit_deterministic_and_order_free(dev):
    ng, npart = 32, 50000
    rng = np.random.default_rng(7)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    a = _plan_density(dev, ng, ph, pmhip.PM_FP32, pmhip.PM_DEPOSIT_SORTED)
    b = _plan_density(dev, ng, ph, pmhip.PM_FP32, pmhip.PM_DEPOSIT_SORTED)
    perm = rng.permutation(npart)
    c = _plan_density(dev, ng, [v[perm] for v in ph], pmhip.PM_FP32, pmhip.PM_DEPOSIT_SORTED)
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)  # fixed-point sums do not depend on the particle order
    atomic = _plan_density(dev, ng, ph, pmhip.PM_FP32, pmhip.PM_DEPOSIT_ATOMIC)
    assert np.max(np.abs(a - atomic)) < 1e-5 * max(1.0, np.abs(a).max())


def test_sorted_deposit_clustered_and_edges(dev):
    """Dense cells (many particles in one cell), particles on tile faces and
    at the periodic box edge, empty tiles."""
    ng = 32
    rng = np.random.default_rng(3)
    n1 = 20000
    clump = [np.full(n1, 15.999) + rng.uniform(0, 1e-3, n1) for _ in range(3)]
    edge = [np.array([0.0, 31.999999, 16.0, 15.5, 31.0]), np.array([0.0, 31.999999, 16.0, 0.5, 15.9999999]),
            np.array([0.0, 31.999999, 16.0, 31.5, 7.0])]
    ph = [np.concatenate([c, e]) for c, e in zip(clump, edge)]
    rho = _plan_density(dev, ng, ph, pmhip.PM_FP64, pmhip.PM_DEPOSIT_SORTED)
    ref = pyoracle.deposit_periodic(ph, ng, mass=ng ** 3 / ph[0].size) + 1.0
    assert np.max(np.abs(rho - ref)) < 1e-8 * max(1.0, np.abs(ref).max())
    assert abs(rho.sum() - ng ** 3) < 1e-6


@pytest.mark.parametrize("prec", [pmhip.PM_FP64, pmhip.PM_FP32])
def test_sorted_deposit_across_segment_seams(dev, prec):
    """ng = 256: a tile column of 16 tiles is walked in two segments of 8,
    so z = 128 is a seam between segments and z = 0 the periodic one; the
    particles sit on and beside those planes, on the tile faces x, y = 16 k,
    at the box edge, over a uniform background.  Against the oracle's
    deposit: 1e-8 of the particle mass (FP64: each weight is rounded to
    2^-32 of the mass, a node sums a few of them) and 1e-6 of
    the largest node value (FP32 mesh: a node sums up to 8 FP32-rounded
    parts, its own and its neighbours' shells and seams; a lost or doubled
    shell or seam value is ~1e-3 of it)."""
    ng = 256
    print(f"case sorted deposit seams prec={prec}", flush=True)
    rng = np.random.default_rng(61)
    n_bg, n_s = 300000, 40000
    bg = [rng.uniform(0, ng, n_bg) for _ in range(3)]
    planes = np.array([127.9999, 128.0, 128.5, 255.99999, 0.0, 0.3, 15.9, 16.0, 143.99, 144.0, 240.0, 239.5])
    zs = rng.choice(planes, n_s) + rng.uniform(0, 1e-6, n_s) * (rng.random(n_s) < 0.5)
    zs = np.minimum(zs, np.nextafter(ng, 0))
    face = lambda: np.where(rng.random(n_s) < 0.4, 16.0 * rng.integers(0, 16, n_s), rng.uniform(0, ng, n_s))
    ph = [np.concatenate([bg[0], face()]), np.concatenate([bg[1], face()]), np.concatenate([bg[2], zs])]
    rho = _plan_density(dev, ng, ph, prec, pmhip.PM_DEPOSIT_SORTED)
    ref = pyoracle.deposit_periodic(ph, ng) + 1.0
    mass = ng ** 3 / ph[0].size
    tol = 1e-8 * mass if prec == pmhip.PM_FP64 else 1e-6 * np.abs(ref).max()
    assert np.max(np.abs(rho - ref)) < tol
    assert pmhip.range_errors() == 0


def test_sorted_deposit_skips_out_of_box(dev):
    ng = 16
    pmhip.range_errors()
    ph = [np.array([1.5, 16.0, -1e-9, 3.0, np.nan]), np.array([1.5] * 5), np.array([1.5] * 5)]
    rho = _plan_density(dev, ng, ph, pmhip.PM_FP64, pmhip.PM_DEPOSIT_SORTED)
    assert pmhip.range_errors() == 3
    # the two in-box particles carry mass ng^3/5 each
    assert abs(rho.sum() - 2 * ng ** 3 / 5) < 1e-6


@pytest.mark.parametrize("variant", [pmhip.PM_DEPOSIT_ATOMIC, pmhip.PM_DEPOSIT_SORTED])
def test_kdk_fp64_variants_match_oracle(dev, variant):
    import torch

    ng, npart, nsteps, dt = 32, 32 ** 3, 5, 0.01
    ph = list(pyoracle.init_particles(npart, ng))
    vh = [np.zeros(npart) for _ in range(3)]
    pos, vel = to_dev(ph, dev), to_dev(vh, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(pos, vel, nsteps, dt, variant)
        torch.cuda.synchronize()
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, 1.0)
    for a, b in zip(pos, ph):
        d = host(a) - b
        d -= ng * np.round(d / ng)
        assert np.max(np.abs(d)) < 1e-7


def test_step_interp_matches_standalone_interp(dev):
    """pm_step with the sorted deposit interpolates while staging the
    drifted positions for the next deposit; the standalone
    pm_interp_kick_drift does the interpolation alone.  Same arithmetic in
    the same order; only the compiler's FMA contraction may differ between
    the two kernels, so the bound is a few ulps of the mesh type."""
    import torch

    ng, nrow = 32, 32
    npart = nrow ** 3
    noise = np.random.default_rng(21).standard_normal(npart)
    pos0, vel0 = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 1.0, 3.0)
    for prec in (pmhip.PM_FP32, pmhip.PM_FP64):
        a_pos, a_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
        b_pos, b_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
        with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
            plan.step(a_pos, a_vel, 0.01, 0.02, pmhip.PM_DEPOSIT_SORTED)
            plan.deposit(b_pos, pmhip.PM_DEPOSIT_SORTED)
            plan.solve()
            plan.interp_kick_drift(b_pos, b_vel, 0.01, 0.02)
            torch.cuda.synchronize()
        eps = 1e-6 if prec == pmhip.PM_FP32 else 1e-14
        for a, b in zip(a_vel, b_vel):
            assert float((a - b).abs().max()) <= eps * float(b.abs().max())
        for a, b in zip(a_pos, b_pos):
            d = a - b
            d = d - ng * torch.round(d / ng)
            assert float(d.abs().max()) <= eps * 0.02 * max(float(v.abs().max()) for v in b_vel) + 1e-12


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
@pytest.mark.parametrize("ng,nrow", [(8, 8), (64, 32), (256, 128)])
def test_step_field_bitwise_equals_deposit_and_solve(dev, prec, ng, nrow):
    """pm_step's density and potential must equal, bit for bit, those of
    pm_deposit followed by pm_solve on the same positions: on the first step
    (binned) and on the second (deposited from the records that the first
    step's interpolation staged), and again after a re-solve of the step's
    density.  ng = 8 is one tile folded onto itself."""
    import torch

    npart = nrow ** 3
    noise = np.random.default_rng(ng + nrow).standard_normal(npart)
    pos, vel = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 1.5, 3.0)
    dt = 0.02
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as pa, \
            pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as pb:
        for k in range(2):
            x_before = [t.clone() for t in pos]
            pa.step(pos, vel, dt, dt, pmhip.PM_DEPOSIT_SORTED | (pmhip.PM_STEP_REUSE_BINS if k else 0))
            pb.deposit(x_before, pmhip.PM_DEPOSIT_SORTED)
            pb.solve()
            torch.cuda.synchronize()
            assert torch.equal(pa.potential(), pb.potential()), f"step {k}: potential"
            rho_a = pa.density().clone()
            torch.cuda.synchronize()
            assert torch.equal(rho_a, pb.density()), f"step {k}: density"
            pa.solve()
            torch.cuda.synchronize()
            assert torch.equal(pa.potential(), pb.potential()), f"step {k}: re-solve of the folded density"


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
@pytest.mark.parametrize("ng,nrow,shuffle", [(32, 32, False), (128, 64, True)])
def test_reused_binning_bitwise_equals_fresh_binning(dev, prec, ng, nrow, shuffle):
    """pm_run_kdk bins every step after the first from the keys and tile
    counts its interpolation wrote; stepping with a fresh binning each time
    must give bit-identical trajectories (the fixed-point deposit is
    order-free).  Shuffled particles on 512 tiles give short key runs and
    more distinct tiles per workgroup than its count hash holds."""
    import torch

    nsteps, dt = 6, 0.02
    npart = nrow ** 3
    noise = np.random.default_rng(22).standard_normal(npart)
    pos0, vel0 = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 1.5, 3.0)
    if shuffle:
        perm = torch.tensor(np.random.default_rng(23).permutation(npart), device=dev)
        pos0 = [t[perm].contiguous() for t in pos0]
        vel0 = [t[perm].contiguous() for t in vel0]
    a_pos, a_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
    b_pos, b_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(a_pos, a_vel, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
        for s in range(nsteps + 1):
            kick = 0.5 * dt if s in (0, nsteps) else dt
            plan.step(b_pos, b_vel, kick, 0.0 if s == nsteps else dt, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    for a, b in zip(a_pos + a_vel, b_pos + b_vel):
        assert torch.equal(a, b)


# ------------------------------------------------ benchmark-size meshes
def _plane_wave(ng, m, dev):
    """theta . i on the mesh (FP64, [x][y][z]) for integer mode m."""
    import torch

    th = [2.0 * np.pi * mi / ng for mi in m]
    i = torch.arange(ng, dtype=torch.float64, device=dev)
    return (th[0] * i).view(-1, 1, 1) + (th[1] * i).view(1, -1, 1) + (th[2] * i).view(1, 1, -1), th


@pytest.mark.parametrize("ng,prec", [(512, pmhip.PM_FP32), (1024, pmhip.PM_FP32), (512, pmhip.PM_FP64),
                                     (1024, pmhip.PM_FP64)])
def test_large_mesh_plane_wave_solve_and_force(dev, ng, prec):
    """Size-independent check at the benchmark meshes.  For delta =
    A cos(theta . i) (theta = 2 pi m / N) the CONT_K2 potential is exactly
    A G(m) cos(theta . i), G = -N^2 / (pi |m|^2); the centred difference of
    B cos(theta . i) along e is B sin(theta_e) sin(theta . i), so the force
    at a particle is the trilinear blend of that at its 8 corners.  Particles
    near the top corner of the box exercise the largest mesh offsets."""
    import torch

    print(f"case plane wave ng={ng} prec={prec}", flush=True)
    m, amp = (3, 5, 7), 0.5
    npart = 1 << 20
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        ph, th = _plane_wave(ng, m, dev)
        plan.density().copy_(1.0 + amp * torch.cos(ph))
        plan.solve()
        torch.cuda.synchronize()
        b = amp * (-(ng * ng) / (np.pi * sum(mi * mi for mi in m)))
        want = torch.cos(ph)
        want *= b
        err = float((plan.potential().double() - want).abs().max()) / abs(b)
        del want, ph
        tol_phi, tol_f = (1e-5, 1e-4) if prec == pmhip.PM_FP32 else (1e-11, 1e-10)
        assert err < tol_phi, err
        g = torch.Generator(device=dev)
        g.manual_seed(ng)
        pos = [torch.rand(npart, dtype=torch.float64, device=dev, generator=g) * ng for _ in range(3)]
        corner = npart // 8  # the last eighth sits in the top cell of every axis
        for p in pos:
            p[-corner:] = (ng - 1) + torch.rand(corner, dtype=torch.float64, device=dev, generator=g) * 0.999
        f = [torch.empty(npart, dtype=torch.float64, device=dev) for _ in range(3)]
        plan.interp_force(pos, f)
        torch.cuda.synchronize()
    i0 = [torch.floor(p) for p in pos]
    t = [p - i for p, i in zip(pos, i0)]
    want = [torch.zeros(npart, dtype=torch.float64, device=dev) for _ in range(3)]
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2]))
                s = torch.sin(th[0] * (i0[0] + dx) + th[1] * (i0[1] + dy) + th[2] * (i0[2] + dz))
                for e in range(3):
                    want[e] += w * (b * np.sin(th[e])) * s
    fmax = max(float(v.abs().max()) for v in want)
    for a, w in zip(f, want):
        assert float((a - w).abs().max()) < tol_f * fmax


@pytest.mark.parametrize("ng", [512, 1024])
def test_large_mesh_sorted_deposit_vs_atomic_fp64(dev, ng):
    """Sorted FP32 deposit (262144 tiles at 1024^3) against the FP64 atomic
    deposit: same CIC weights, so cells and the total mass differ by FP32
    rounding of the mesh only."""
    import torch

    print(f"case large deposit ng={ng}", flush=True)
    npart = 1 << 21
    g = torch.Generator(device=dev)
    g.manual_seed(ng + 5)
    pos = [torch.rand(npart, dtype=torch.float64, device=dev, generator=g) * ng for _ in range(3)]
    for p in pos:  # a dense clump at the top corner and particles on the seams
        p[:4096] = ng - 0.5 + 0.49 * torch.rand(4096, dtype=torch.float64, device=dev, generator=g)
        p[4096:8192] = 0.0
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as p64:
        p64.deposit(pos, pmhip.PM_DEPOSIT_ATOMIC)
        torch.cuda.synchronize()
        ref = p64.density().clone()
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as p32:
        p32.deposit(pos, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
        rho = p32.density().double()
        assert pmhip.range_errors() == 0
        assert float((rho - ref).abs().max()) < 2e-6 * float(ref.abs().max())
        assert abs(float(rho.sum()) - float(ref.sum())) < 1e-3 * ng ** 3 * 1e-6 + 1e-2


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_large_mesh_reused_binning_1024(dev, prec):
    """Key reuse across steps at 1024^3 (262144 tiles; FP64: 42-bit staged
    offsets): run_kdk against fresh binning every step, bit-identical."""
    import torch

    print(f"case reused binning 1024 prec={prec}", flush=True)

    ng, nrow, nsteps, dt = 1024, 128, 3, 0.05
    npart = nrow ** 3
    noise = np.random.default_rng(24).standard_normal(npart)
    pos0, vel0 = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 2.0, 3.0)
    a_pos, a_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
    b_pos, b_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(a_pos, a_vel, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
        for s in range(nsteps + 1):
            kick = 0.5 * dt if s in (0, nsteps) else dt
            plan.step(b_pos, b_vel, kick, 0.0 if s == nsteps else dt, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    for a, b in zip(a_pos + a_vel, b_pos + b_vel):
        assert torch.equal(a, b)


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_reused_binning_after_bucket_overflow(dev, prec):
    """A drift that packs every particle into the central eighth of the box
    overflows the tile buckets that the FP32 interpolation stages into (each
    holds its previous count + 1/8 + 32): the next deposit must fall back to
    a full binning and still match fresh binning bit for bit."""
    import torch

    ng, nrow, dt = 64, 32, 0.1
    npart = nrow ** 3
    noise = np.random.default_rng(25).standard_normal(npart)
    pos0, vel0 = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 0.5, 0.0)
    c = ng / 2.0
    vel0 = [(c - p) * (0.5 / dt) for p in pos0]  # x -> (x + c) / 2 in one drift
    a_pos, a_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
    b_pos, b_vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(a_pos, a_vel, 2, dt, pmhip.PM_DEPOSIT_SORTED)
        for s, (k, d) in enumerate([(0.5 * dt, dt), (dt, dt), (0.5 * dt, 0.0)]):
            plan.step(b_pos, b_vel, k, d, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
    assert pmhip.range_errors() == 0
    for a, b in zip(a_pos + a_vel, b_pos + b_vel):
        assert torch.equal(a, b)


@pytest.mark.parametrize("ng,nrow", [(64, 32), (256, 128)])
@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_step_density_equals_standalone_deposit(dev, prec, ng, nrow):
    """pm_step deposits through the staging of the previous step's
    interpolation (column walk over the staged records, then the shell
    fold): after a step (binned first step, then a staged one) the plan's
    density must be bit for bit pm_deposit's density of the same positions."""
    import torch

    npart = nrow ** 3
    noise = np.random.default_rng(27).standard_normal(npart)
    pos, vel = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 1.5, 3.0)
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan, \
            pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as ref:
        for s in range(2):
            before = [p.clone() for p in pos]
            plan.step(pos, vel, 0.01, 0.05, pmhip.PM_DEPOSIT_SORTED | (pmhip.PM_STEP_REUSE_BINS if s else 0))
            ref.deposit(before, pmhip.PM_DEPOSIT_SORTED)
            torch.cuda.synchronize()
            assert torch.equal(plan.density(), ref.density()), s


def test_run_cosmo_matches_oracle_and_reference_kick(dev, golden_dir):
    """pm_run_cosmo (FP64 mesh, sorted staged deposit) against the oracle's
    cosmological KDK over 6 steps in a; zero steps from Start_a with
    da = 2 Da reproduce the reference's kick fixture.  Bound 1e-10: the
    sorted deposit rounds each weight to 2^-32 mass units, ~1e-11 relative
    on forces at this mass per particle, as in the plan tests."""
    import torch

    ng, npart, n, a0, da = 32, 32 ** 3, 6, 0.5, 0.02
    rng = np.random.default_rng(41)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    mh = [0.3 * rng.standard_normal(npart) for _ in range(3)]
    pos, mom = to_dev(ph, dev), to_dev(mh, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_REF_KH, 1.0) as plan:
        plan.run_cosmo(pos, mom, a0, da, n, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
    pyoracle.run_cosmo(ph, mh, ng, n, a0, da, pyoracle.GREEN_REF_KH, 1.0)
    for a, b in zip(pos, ph):
        d = host(a) - b
        d -= ng * np.round(d / ng)
        assert np.max(np.abs(d)) < 1e-10
    for a, b in zip(mom, mh):
        assert rel(host(a), b) < 1e-10

    f = np.load(golden_dir / "ref_kick_ng32_np4096.npz")
    pos = to_dev([f["posx"], f["posy"], f["posz"]], dev)
    mom = to_dev([np.zeros(f["posx"].size)] * 3, dev)
    with pmhip.Plan(int(f["ng"]), f["posx"].size, pmhip.PM_FP64, pmhip.PM_GREEN_REF_KH, 1.0) as plan:
        plan.run_cosmo(pos, mom, float(f["a"]), 2 * pmhip.DA, 0, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
    for a, key in zip(mom, ("px", "py", "pz")):
        assert rel(host(a), f[key]) < 1e-10


def test_snapshot_round_trip(dev, tmp_path):
    """pm_snapshot_write of device arrays, read back by pm_snapshot_read and
    by the numpy loader: bit-identical, header fields preserved."""
    import torch

    npart, ng = 100003, 64  # not a multiple of anything the copy loop uses
    g = torch.Generator(device=dev)
    g.manual_seed(51)
    pos = [torch.rand(npart, dtype=torch.float64, device=dev, generator=g) * ng for _ in range(3)]
    vel = [torch.randn(npart, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    path = tmp_path / "snap.bin"
    pmhip.write_snapshot(path, pos, vel, ng, step=12, a=0.75)
    assert pmhip.snapshot_info(path) == (npart, ng, 12, 0.75)
    pos2 = [torch.empty_like(p) for p in pos]
    vel2 = [torch.empty_like(v) for v in vel]
    pmhip.read_snapshot(path, pos2, vel2)
    torch.cuda.synchronize()
    for a, b in zip(pos + vel, pos2 + vel2):
        assert torch.equal(a, b)
    head, hp, hv = pmhip.load_snapshot_numpy(path)
    assert head["np"] == npart and head["step"] == 12
    for a, b in zip(pos + vel, list(hp) + list(hv)):
        assert np.array_equal(host(a), b)
    with pytest.raises(pmhip.PMError):
        pmhip.read_snapshot(path, [p[:10] for p in pos2], [v[:10] for v in vel2])


def test_phase_timing_records_reserved_steps(dev):
    """pm_plan_enable_timing(k) records the phases of the next k steps (no
    synchronisation inside the loop); pm_plan_phase_ms returns the sums and
    counts; reset clears them; k = 0 turns recording off."""
    import torch

    ng, nrow = 64, 32
    noise = torch.tensor(np.random.default_rng(71).standard_normal(nrow ** 3), device=dev)
    pos, vel = pmhip.zeldovich_ic(nrow, ng, noise, 1.0, 1.0)
    with pmhip.Plan(ng, nrow ** 3, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.enable_timing(3)
        for s in range(5):
            plan.step(pos, vel, 0.01, 0.02, pmhip.PM_DEPOSIT_SORTED | (pmhip.PM_STEP_REUSE_BINS if s else 0))
        ms, cnt = plan.phase_ms()
        assert cnt == [3] * pmhip.PM_NPHASES
        assert all(m > 0 for m in ms)
        plan.reset_timing()
        ms, cnt = plan.phase_ms()
        assert cnt == [0] * pmhip.PM_NPHASES and ms == [0.0] * pmhip.PM_NPHASES
        plan.enable_timing(0)
        plan.step(pos, vel, 0.01, 0.02, pmhip.PM_DEPOSIT_SORTED | pmhip.PM_STEP_REUSE_BINS)
        ms, cnt = plan.phase_ms()
        assert cnt == [0] * pmhip.PM_NPHASES


def test_plan_on_another_stream_matches_default_stream(dev):
    """pm_plan_set_stream moves a plan's work to another stream (default
    priority); the same steps give the same particles bit for bit."""
    import torch

    ng, nrow = 64, 32
    noise = torch.tensor(np.random.default_rng(72).standard_normal(nrow ** 3), device=dev)
    pos0, vel0 = pmhip.zeldovich_ic(nrow, ng, noise, 1.0, 1.0)
    torch.cuda.synchronize()
    out = []
    side = torch.cuda.Stream()
    for use_side in (False, True):
        pos, vel = [p.clone() for p in pos0], [v.clone() for v in vel0]
        torch.cuda.synchronize()
        with pmhip.Plan(ng, nrow ** 3, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
            if use_side:
                plan.set_stream(side.cuda_stream)
            for s in range(3):
                plan.step(pos, vel, 0.01, 0.02, pmhip.PM_DEPOSIT_SORTED | (pmhip.PM_STEP_REUSE_BINS if s else 0))
            (side if use_side else torch.cuda.current_stream()).synchronize()
            out.append([t.clone() for t in pos + vel])
    for a, b in zip(*out):
        assert torch.equal(a, b)
