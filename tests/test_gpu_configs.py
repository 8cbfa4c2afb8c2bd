"""The benchmark configurations that are parity cases rather than bench lines.

config 2: 256^3 Zel'dovich-lattice particles on a 256^3 mesh, 0.3-cell rms,
          20 KDK steps of dt = 0.005; the FP64 mesh path against the FP64
          oracle on the same inputs, and the FP32 mesh path against both
          (final-position error budget 1e-4 cells rms).
config 3: 512^3 particles on a 512^3 mesh, clustered Zel'dovich ICs at 1.5-cell
          rms, 20 KDK steps of dt = 0.005; the FP64 and FP32 mesh paths against
          the FP64 oracle, the split-precision state (the benchmark step)
          against it too, and the sorted deposit against the atomic one.
config 4: 128^3 particles on a 128^3 mesh, 10 KDK steps; the gradient of
          ||rho_final - rho_target||^2 with respect to the initial
          displacements through the adjoint kernels with per-step
          checkpoints, checked by central differences on 64 coordinates and
          against the float64 autograd gradient of tests/torch_pm.py.
config 5: 512^3 particles on a 1024^3 mesh, 1.0-cell rms, 30 KDK steps, then
          the reverse sweep (reverse-time leapfrog, no stored trajectory): the
          reversal drift, and the gradient of a linear loss on the final
          positions against central differences of full forward runs.
"""
import math

import numpy as np
import pytest

import pmhip
import pmhip.autograd as pag
import pyoracle

pytestmark = pytest.mark.gpu


def _zeldovich(nrow, ng, rms, seed, dev, vel_factor=math.sqrt(4.0 * math.pi)):
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    noise = torch.randn(nrow ** 3, dtype=torch.float64, device=dev, generator=g)
    return pmhip.zeldovich_ic(nrow, ng, noise, rms, vel_factor)


def _periodic_rms_max(a, b, ng):
    import torch

    d2 = None
    for p, q in zip(a, b):
        d = torch.as_tensor(p) - torch.as_tensor(q, device=torch.as_tensor(p).device)
        d = d - ng * torch.round(d / ng)
        d2 = d * d if d2 is None else d2 + d * d
    return float(torch.sqrt(d2.mean())), float(torch.sqrt(d2.max()))


def test_config2_fp32_error_budget(dev):
    import torch

    print("case config2 256^3 20 steps", flush=True)
    ng = nrow = 256
    nsteps, dt = 20, 0.005
    npart = nrow ** 3
    pos0, vel0 = _zeldovich(nrow, ng, 0.3, 2, dev)
    final = {}
    for prec in (pmhip.PM_FP64, pmhip.PM_FP32):
        pos, vel = [p.clone() for p in pos0], [v.clone() for v in vel0]
        with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
            plan.run_kdk(pos, vel, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
            torch.cuda.synchronize()
        final[prec] = pos
    assert pmhip.range_errors() == 0
    # the FP64 oracle on the same inputs (host, OpenMP)
    ph = [p.cpu().numpy().copy() for p in pos0]
    vh = [v.cpu().numpy().copy() for v in vel0]
    pyoracle.set_threads(min(16, pyoracle.max_threads()))
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, 1.0)
    rms64, max64 = _periodic_rms_max(final[pmhip.PM_FP64], [torch.tensor(p, device=dev) for p in ph], ng)
    rms32, max32 = _periodic_rms_max(final[pmhip.PM_FP32], [torch.tensor(p, device=dev) for p in ph], ng)
    print(f"FP64 vs oracle rms {rms64:.3e} max {max64:.3e}; FP32 vs oracle rms {rms32:.3e} max {max32:.3e}",
          flush=True)
    assert max64 < 1e-9
    assert rms32 < 1e-4
    assert max32 < 1e-3


def test_config4_gradient_128(dev):
    """Target: the density after a forward run of the seed-perturbed ICs.
    The FP32-force gradient (sorted deposit, per-step checkpoints) must agree
    with the float64 autograd gradient to 2e-3 of its largest component; the
    FP64 adjoint gradient (atomic deposit: smooth in x, see
    test_gpu_adjoint.py) to 1e-8, and with central differences on 64
    coordinates."""
    import torch
    import torch_pm

    print("case config4 128^3 10 steps", flush=True)
    ng = nrow = 128
    nsteps, dt, alpha = 10, 0.005, math.sqrt(4.0 * math.pi)
    npart = nrow ** 3
    # displacement field psi (unit velocity factor) and the lattice
    q0, psi = _zeldovich(nrow, ng, 0.3, 4, dev, vel_factor=1.0)
    Q = torch.stack([q - p for q, p in zip(q0, psi)])
    Q = Q - ng * torch.floor(Q / ng)
    psi0 = torch.stack(psi)
    _, dpsi = _zeldovich(nrow, ng, 0.1, 5, dev, vel_factor=1.0)
    psi_t = psi0 + torch.stack(dpsi)
    green = torch_pm.green_cont_k2(ng, dev)
    with torch.no_grad():
        xt, _ = torch_pm.run_kdk(torch.remainder(Q + psi_t, float(ng)), alpha * psi_t, ng, nsteps, dt, 1.0, green)
        target = torch_pm.deposit(xt, ng, 1.0)

    def loss(psi, plan, variant, checkpoint):
        x0 = torch.remainder(Q + psi, float(ng))
        x, _ = pag.evolve(x0, alpha * psi, plan, nsteps, dt, variant=variant, checkpoint=checkpoint)
        rho = pag.density(x, plan, variant)
        return ((rho.double() - target) ** 2).sum()

    # float64 autograd reference
    psi_r = psi0.clone().requires_grad_(True)
    x, _ = torch_pm.run_kdk(torch.remainder(Q + psi_r, float(ng)), alpha * psi_r, ng, nsteps, dt, 1.0, green)
    Lr = ((torch_pm.deposit(x, ng, 1.0) - target) ** 2).sum()
    (g_ref,) = torch.autograd.grad(Lr, psi_r)
    del x
    gmax = float(g_ref.abs().max())

    plan32 = pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0)
    p = psi0.clone().requires_grad_(True)
    (g32,) = torch.autograd.grad(loss(p, plan32, pmhip.PM_DEPOSIT_SORTED, True), p)
    plan32.close()
    err32 = float((g32 - g_ref).abs().max()) / gmax
    print(f"FP32 adjoint vs FP64 autograd: {err32:.3e}", flush=True)
    assert err32 < 2e-3

    plan64 = pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0)
    p = psi0.clone().requires_grad_(True)
    L64 = loss(p, plan64, pmhip.PM_DEPOSIT_ATOMIC, True)
    (g64,) = torch.autograd.grad(L64, p)
    err64 = float((g64 - g_ref).abs().max()) / gmax
    print(f"FP64 adjoint vs FP64 autograd: {err64:.3e}", flush=True)
    assert err64 < 1e-8

    rng = np.random.default_rng(6)
    coords = [(int(rng.integers(3)), int(rng.integers(npart))) for _ in range(64)]
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for a, i in coords:
            dp = psi0.clone()
            dp[a, i] += eps
            dm = psi0.clone()
            dm[a, i] -= eps
            fd = (float(loss(dp, plan64, pmhip.PM_DEPOSIT_ATOMIC, False)) -
                  float(loss(dm, plan64, pmhip.PM_DEPOSIT_ATOMIC, False))) / (2 * eps)
            an = float(g64[a, i])
            worst = max(worst, abs(fd - an) / (1e-4 * max(1.0, abs(an)) + 1e-5 * gmax))
    plan64.close()
    print(f"central differences: worst error / tolerance {worst:.3f}", flush=True)
    assert worst < 1.0


def _oracle_threads():
    # the host cores this GPU's process may use (16 on one GPU's share)
    pyoracle.set_threads(min(16, pyoracle.max_threads()))


def test_config3_512_fp64_fp32_vs_oracle_and_sorted_vs_atomic(dev):
    """FP64 mesh path vs the FP64 oracle after 20 steps: max position error
    below 1e-8 cells (the sorted deposit rounds each weight to 2^-32 of the
    mass before its exact sum; the oracle sums in FP64).  FP32 mesh path,
    with FP64 positions and with the split state, vs the oracle: rms below
    1e-4 cells (the error budget), max below 1e-3.
    Sorted vs atomic FP32 density after the first step: 1e-5 of the peak."""
    import torch

    print("case config3 512^3 20 steps", flush=True)
    ng = nrow = 512
    nsteps, dt = 20, 0.005
    npart = nrow ** 3
    pos0, vel0 = _zeldovich(nrow, ng, 1.5, 3, dev)
    pmhip.range_errors()
    final = {}
    for prec in (pmhip.PM_FP64, pmhip.PM_FP32):
        pos, vel = [p.clone() for p in pos0], [v.clone() for v in vel0]
        with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
            plan.run_kdk(pos, vel, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
            torch.cuda.synchronize()
        final[prec] = [p.cpu() for p in pos]
        del pos, vel
    # split-precision state (integer cell + 2^-23-cell fraction, FP32 velocities)
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.split_load(pos0, vel0)
        plan.split_run_kdk(nsteps, dt)
        pos = [torch.empty_like(p) for p in pos0]
        vel = [torch.empty_like(v) for v in vel0]
        plan.split_store(pos, vel)
        torch.cuda.synchronize()
    final["split"] = [p.cpu() for p in pos]
    del pos, vel
    assert pmhip.range_errors() == 0
    # sorted vs atomic deposit on the positions after one step
    pos1, vel1 = [p.clone() for p in pos0], [v.clone() for v in vel0]
    with pmhip.Plan(ng, npart, pmhip.PM_FP32, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(pos1, vel1, 1, dt, pmhip.PM_DEPOSIT_SORTED)
        plan.deposit(pos1, pmhip.PM_DEPOSIT_SORTED)
        rs = plan.density().clone()
        plan.deposit(pos1, pmhip.PM_DEPOSIT_ATOMIC)
        ra = plan.density().clone()
        torch.cuda.synchronize()
    da = float((rs - ra).abs().max()) / float(rs.abs().max())
    print(f"sorted vs atomic density after one step: max {da:.3e} of the peak "
          f"(peak {float(rs.max()):.1f}, min {float(rs.min()):.3f})", flush=True)
    assert da < 1e-5
    del pos1, vel1, rs, ra
    # the FP64 oracle on the same inputs
    ph = [p.cpu().numpy().copy() for p in pos0]
    vh = [v.cpu().numpy().copy() for v in vel0]
    del pos0, vel0
    _oracle_threads()
    pyoracle.run_kdk(ph, vh, ng, nsteps, dt, pyoracle.GREEN_CONT_K2, 1.0)
    ref = [torch.from_numpy(p) for p in ph]
    rms64, max64 = _periodic_rms_max(final[pmhip.PM_FP64], ref, ng)
    rms32, max32 = _periodic_rms_max(final[pmhip.PM_FP32], ref, ng)
    rmss, maxs = _periodic_rms_max(final["split"], ref, ng)
    print(f"FP64 vs oracle rms {rms64:.3e} max {max64:.3e}; FP32 vs oracle rms {rms32:.3e} max {max32:.3e}; "
          f"split state vs oracle rms {rmss:.3e} max {maxs:.3e}", flush=True)
    assert max64 < 1e-8
    assert rms32 < 1e-4
    assert max32 < 1e-3
    assert rmss < 1e-4
    assert maxs < 1e-3


def test_config5_1024_mesh_reversal_and_gradient(dev):
    """Forward 30 KDK steps of 512^3 particles on a 1024^3 mesh, then
    pm_run_adjoint with the cotangent w of L = sum w . x_final.
    FP32 mesh (sorted deposit, the benchmark path): the recomputed initial
    positions are within 2e-6 cells rms and 5e-5 max of the true ones.
    FP64 mesh (atomic deposit, smooth in x for differencing): reversal within
    1e-9 cells max, and dL/dx_0 at 8 coordinates against central
    differences (eps = 1e-4 cells) of two full forward runs each, to 1e-4
    relative plus 1e-6 of the largest gradient component."""
    import torch

    print("case config5 512^3 on 1024^3 30 steps + adjoint", flush=True)
    ng, nrow = 1024, 512
    nsteps, dt = 30, 0.005
    npart = nrow ** 3
    pos0, vel0 = _zeldovich(nrow, ng, 1.0, 7, dev)
    g = torch.Generator(device=dev)
    g.manual_seed(8)
    w = [torch.randn(npart, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    pmhip.range_errors()

    def forward(plan, pos, vel, variant):
        plan.run_kdk(pos, vel, nsteps, dt, variant)

    for prec, variant in ((pmhip.PM_FP32, pmhip.PM_DEPOSIT_SORTED), (pmhip.PM_FP64, pmhip.PM_DEPOSIT_ATOMIC)):
        with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
            pos, vel = [p.clone() for p in pos0], [v.clone() for v in vel0]
            forward(plan, pos, vel, variant)
            xfinal = [p.clone() for p in pos]
            xbar = [t.clone() for t in w]
            vbar = [torch.zeros_like(t) for t in w]
            plan.run_adjoint(pos, vel, xbar, vbar, nsteps, dt, variant)
            torch.cuda.synchronize()
            rms, mx = _periodic_rms_max(pos, pos0, ng)
            print(f"prec={prec}: reversal drift rms {rms:.3e} max {mx:.3e} cells", flush=True)
            if prec == pmhip.PM_FP32:
                assert rms < 2e-6 and mx < 5e-5
                continue
            assert mx < 1e-9
            gmax = max(float(t.abs().max()) for t in xbar)
            rng = np.random.default_rng(9)
            eps, worst, checked = 1e-4, 0.0, 0
            while checked < 8:
                a, i = int(rng.integers(3)), int(rng.integers(npart))
                # keep away from the periodic wrap of the loss and of the box
                x0, xf = float(pos0[a][i]), float(xfinal[a][i])
                if min(xf, ng - xf, x0, ng - x0) < 0.1:
                    continue
                runs = []
                for sgn in (1.0, -1.0):
                    p, v = [t.clone() for t in pos0], [t.clone() for t in vel0]
                    p[a][i] += sgn * eps
                    forward(plan, p, v, variant)
                    runs.append(p)
                d = 0.0
                for b in range(3):
                    dx = runs[0][b] - runs[1][b]
                    dx = dx - ng * torch.round(dx / ng)
                    d += float((w[b] * dx).sum())
                fd = d / (2 * eps)
                an = float(xbar[a][i])
                err = abs(fd - an) / (1e-4 * max(1.0, abs(an)) + 1e-6 * gmax)
                print(f"  dL/dx0[{a}][{i}]: adjoint {an:.8f} central difference {fd:.8f}", flush=True)
                worst = max(worst, err)
                checked += 1
                del runs, p, v
            print(f"central differences: worst error / tolerance {worst:.3f}", flush=True)
            assert worst < 1.0
    assert pmhip.range_errors() == 0
