"""GPU tests of the hand-written adjoint kernels (through the C ABI).

Pins: a numpy scatter restatement of the interpolation transpose, the
oracle's solve for the adjoint solve, central finite differences of the
oracle's force and deposit for the vector-Jacobian products, exact reversal
of a step, and torch float64 autograd through an independent restatement of
the forward step (tests/torch_pm.py) as the FP64 reference gradient.
Tolerances are stated per test (FP64 mesh unless noted).
"""
import numpy as np
import pytest

import pmhip
import pmhip.autograd as pag
import pyoracle

pytestmark = pytest.mark.gpu


def to_dev(arrs, dev):
    import torch

    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev) for a in arrs]


def host(t):
    return t.detach().cpu().numpy()


def cic_scatter(pos, w, ng):
    """numpy np.add.at CIC of per-particle weights w on the periodic mesh."""
    out = np.zeros((ng, ng, ng))
    i = [np.floor(p).astype(np.int64) for p in pos]
    t = [p - ii for p, ii in zip(pos, i)]
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                wt = (t[0] if a else 1 - t[0]) * (t[1] if b else 1 - t[1]) * (t[2] if c else 1 - t[2])
                np.add.at(out, ((i[0] + a) % ng, (i[1] + b) % ng, (i[2] + c) % ng), wt * w)
    return out


def _seam_positions(rng, ng, npart):
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    for p in ph:  # on the periodic seams
        p[:64] = 0.0
        p[64:128] = np.nextafter(float(ng), 0.0)
    return ph


@pytest.mark.parametrize("ng,npart", [(8, 3000), (32, 20000), (256, 300000)])
@pytest.mark.parametrize("prec,tol", [(pmhip.PM_FP64, 5e-11), (pmhip.PM_FP32, 2e-6)])
def test_interp_transpose_vs_numpy(dev, prec, tol, ng, npart):
    """The LDS accumulation is exact in units of 2^-40 of the largest
    cotangent component, so FP64 agrees to ~1e-11 relative; FP32 to its
    final rounding.  ng = 8 is one 8^3 tile folded onto itself, 256 has two
    walk segments per column (the z-seam fold of each of the three meshes)."""
    import torch

    print(f"case interp transpose ng={ng} np={npart} prec={prec}", flush=True)
    pmhip.range_errors()  # reset the counter
    rng = np.random.default_rng(1 + ng)
    ph = _seam_positions(rng, ng, npart)
    uh = [rng.standard_normal(npart) * s for s in (1.0, 3.0, 0.25)]
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.interp_transpose(to_dev(ph, dev), to_dev(uh, dev))
        torch.cuda.synchronize()
        for a in range(3):
            ref = cic_scatter(ph, uh[a], ng)
            got = host(plan.adjoint_mesh(a)).astype(np.float64)
            err = float(np.max(np.abs(got - ref)) / np.abs(ref).max())
            assert err < tol, (a, err)
    assert pmhip.range_errors() == 0


@pytest.mark.parametrize("ng,npart", [(8, 3000), (32, 8000), (256, 300000)])
def test_adjoint_solve_vs_oracle(dev, ng, npart):
    """lambda = S sum_a D_a^T U_a from the three interpolation-transpose
    meshes (k_div3 + the solve) against the oracle's solve of the numpy
    scatter's divergence."""
    import torch

    print(f"case adjoint solve ng={ng} np={npart}", flush=True)
    rng = np.random.default_rng(2 + ng)
    ph = _seam_positions(rng, ng, npart)
    uh = [rng.standard_normal(npart) for _ in range(3)]
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.interp_transpose(to_dev(ph, dev), to_dev(uh, dev))
        plan.adjoint_solve()
        torch.cuda.synchronize()
        U = [cic_scatter(ph, uh[a], ng) for a in range(3)]
        # (D_a^T U)[c] = (U[c + e_a] - U[c - e_a]) / 2
        mu = sum((np.roll(U[a], -1, a) - np.roll(U[a], 1, a)) * 0.5 for a in range(3))
        lam = pyoracle.solve(mu, pyoracle.GREEN_CONT_K2, 1.0)
        err = float(np.max(np.abs(host(plan.adjoint_mesh(3)) - lam)) / np.abs(lam).max())
        assert err < 1e-9, err


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_non_finite_cotangent_is_reported(dev, bad):
    """No fixed-point scale holds an inf or NaN cotangent: the deposits of u
    count it in the range-error counter instead of failing silently."""
    import torch

    ng, npart = 16, 4000
    rng = np.random.default_rng(9)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    uh = [rng.standard_normal(npart) for _ in range(3)]
    uh[1][17] = bad
    pmhip.range_errors()  # reset the counter
    xbar = to_dev([np.zeros(npart)] * 3, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.interp_transpose(to_dev(ph, dev), to_dev(uh, dev))
        torch.cuda.synchronize()
        assert pmhip.range_errors() >= 1
        plan.adjoint_force(to_dev(ph, dev), to_dev(uh, dev), xbar, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
        assert pmhip.range_errors() >= 1
        uh[1][17] = 0.5
        plan.interp_transpose(to_dev(ph, dev), to_dev(uh, dev))
        torch.cuda.synchronize()
        assert pmhip.range_errors() == 0


@pytest.mark.parametrize("ng,npart", [(8, 3000), (32, 20000), (256, 300000)])
@pytest.mark.parametrize("prec,tol", [(pmhip.PM_FP64, 5e-11), (pmhip.PM_FP32, 2e-6)])
def test_divergence_deposit_vs_numpy(dev, prec, tol, ng, npart):
    """The adjoint paths deposit mu = sum_a D_a^T interp^T(u_a) in one pass
    (32 nodes per particle, column walk with three halo planes, its own
    fold) into adjoint mesh 0; pinned against the numpy CIC scatter of each
    u_a and the centred-difference transpose.  ng = 8 is one tile folded onto
    itself, 256 has two segments per column (seam fold).  The accumulation is
    exact in units of 2^-40 of the largest cotangent: FP64 to ~1e-11, FP32 to
    its final rounding."""
    import torch

    print(f"case divergence deposit ng={ng} np={npart} prec={prec}", flush=True)
    pmhip.range_errors()  # reset the counter
    rng = np.random.default_rng(ng + npart)
    ph = _seam_positions(rng, ng, npart)
    uh = [rng.standard_normal(npart) * s for s in (1.0, 3.0, 0.25)]
    xbar = to_dev([np.zeros(npart)] * 3, dev)
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.adjoint_force(to_dev(ph, dev), to_dev(uh, dev), xbar, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
        got = host(plan.adjoint_mesh(0)).astype(np.float64)
    assert pmhip.range_errors() == 0
    U = [cic_scatter(ph, uh[a], ng) for a in range(3)]
    mu = sum((np.roll(U[a], -1, a) - np.roll(U[a], 1, a)) * 0.5 for a in range(3))
    err = float(np.max(np.abs(got - mu)) / np.abs(mu).max())
    assert err < tol, err


def test_adjoint_deposit_vs_finite_differences(dev):
    import torch

    ng, npart = 16, 300
    rng = np.random.default_rng(3)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    mbar = rng.standard_normal((ng, ng, ng))
    xbar = to_dev([np.zeros(npart)] * 3, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.adjoint_deposit(to_dev(ph, dev), torch.tensor(mbar, device=dev), xbar)
        torch.cuda.synchronize()
    e = [rng.standard_normal(npart) for _ in range(3)]
    eps = 1e-7
    lp = np.sum(pyoracle.deposit_periodic([p + eps * d for p, d in zip(ph, e)], ng) * mbar)
    lm = np.sum(pyoracle.deposit_periodic([p - eps * d for p, d in zip(ph, e)], ng) * mbar)
    fd = (lp - lm) / (2 * eps)
    an = sum(np.dot(host(b), d) for b, d in zip(xbar, e))
    assert abs(fd - an) < 1e-6 * max(1.0, abs(an))


@pytest.mark.parametrize("variant", [pmhip.PM_DEPOSIT_SORTED, pmhip.PM_DEPOSIT_ATOMIC])
def test_adjoint_force_vs_finite_differences(dev, variant):
    """<u, J_F e> by central differences of the oracle's force equals
    <J_F^T u, e> from the adjoint kernels."""
    import torch

    ng, npart = 16, 500
    rng = np.random.default_rng(4)
    ph = [rng.uniform(0, ng, npart) for _ in range(3)]
    uh = [rng.standard_normal(npart) for _ in range(3)]
    xbar = to_dev([np.zeros(npart)] * 3, dev)
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.adjoint_force(to_dev(ph, dev), to_dev(uh, dev), xbar, variant)
        torch.cuda.synchronize()

    def F(pos):
        d = pyoracle.deposit_periodic(pos, ng)
        return pyoracle.interp_force(pos, pyoracle.solve(d, pyoracle.GREEN_CONT_K2, 1.0))

    for trial in range(3):
        e = [rng.standard_normal(npart) for _ in range(3)]
        eps = 1e-6
        fp = F([p + eps * d for p, d in zip(ph, e)])
        fm = F([p - eps * d for p, d in zip(ph, e)])
        fd = sum(np.dot(u, (a - b) / (2 * eps)) for u, a, b in zip(uh, fp, fm))
        an = sum(np.dot(host(b), d) for b, d in zip(xbar, e))
        assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (trial, fd, an)


def test_adjoint_step_reverses_forward_step(dev):
    import torch

    ng, npart = 32, 32 ** 3
    ph = list(pyoracle.init_particles(npart, ng))
    pos = to_dev(ph, dev)
    vel = to_dev([np.zeros(npart)] * 3, dev)
    xb = to_dev([np.zeros(npart)] * 3, dev)
    vb = to_dev([np.zeros(npart)] * 3, dev)
    kick, drift = 0.01, 0.01
    with pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.step(pos, vel, kick, drift, pmhip.PM_DEPOSIT_SORTED)
        plan.adjoint_step(pos, vel, xb, vb, kick, drift, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
    for a, b in zip(pos, ph):
        d = host(a) - b
        d -= ng * np.round(d / ng)
        assert np.max(np.abs(d)) < 1e-11
    assert max(np.abs(host(v)).max() for v in vel) < 1e-11
    assert max(np.abs(host(v)).max() for v in xb + vb) == 0.0  # zero cotangent stays zero


def _density_loss_setup(dev, ng, nrow, seed):
    import torch

    rng = np.random.default_rng(seed)
    q = (np.arange(nrow) + 0.5) * (ng / nrow)
    Q = np.stack([a.ravel() for a in np.meshgrid(q, q, q, indexing="ij")])
    psi = 0.4 * rng.standard_normal(Q.shape)
    target = np.abs(rng.standard_normal((ng, ng, ng))) + 0.5
    return torch.tensor(Q, device=dev), torch.tensor(psi, device=dev), torch.tensor(target, device=dev)


@pytest.mark.parametrize("variant,checkpoint", [(pmhip.PM_DEPOSIT_SORTED, False), (pmhip.PM_DEPOSIT_ATOMIC, False),
                                                (pmhip.PM_DEPOSIT_ATOMIC, True)])
def test_gradient_through_kdk_vs_torch_autograd_and_fd(dev, variant, checkpoint):
    """L = ||rho(x_n) - target||^2 after n KDK steps, gradient w.r.t. the
    initial displacement psi (x0 = q + psi, v0 = alpha psi): adjoint kernels
    vs torch float64 autograd of tests/torch_pm.py vs central differences.
    Central differences use the FP64 atomic deposit only: the sorted
    deposit rounds each contribution to 2^-32 mass units, which a 1e-6
    position step resolves only to ~1e-4 relative."""
    import torch
    import torch_pm

    ng, nrow, nsteps, dt, alpha = 16, 16, 4, 0.01, 1.5
    npart = nrow ** 3
    Q, psi0, target = _density_loss_setup(dev, ng, nrow, 5)
    plan = pmhip.Plan(ng, npart, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0)

    def loss_gpu(psi):
        x0 = torch.remainder(Q + psi, float(ng))
        x, v = pag.evolve(x0, alpha * psi, plan, nsteps, dt, variant=variant, checkpoint=checkpoint)
        rho = pag.density(x, plan, variant)
        return ((rho - target) ** 2).sum()

    psi = psi0.clone().requires_grad_(True)
    L = loss_gpu(psi)
    (g_adj,) = torch.autograd.grad(L, psi)

    # torch float64 autograd reference (independent forward restatement)
    green = torch_pm.green_cont_k2(ng, dev)
    psi_r = psi0.clone().requires_grad_(True)
    x0 = torch.remainder(Q + psi_r, float(ng))
    x, v = torch_pm.run_kdk(x0, alpha * psi_r, ng, nsteps, dt, 1.0, green)
    Lr = ((torch_pm.deposit(x, ng, 1.0) - target) ** 2).sum()
    (g_ref,) = torch.autograd.grad(Lr, psi_r)
    assert abs(float(L.detach()) - float(Lr.detach())) < 1e-10 * abs(float(Lr.detach()))
    err = float((g_adj - g_ref).abs().max() / g_ref.abs().max())
    assert err < 1e-9, err
    if variant != pmhip.PM_DEPOSIT_ATOMIC:
        plan.close()
        return

    # central differences on 16 random coordinates
    rng = np.random.default_rng(6)
    idx = [(int(rng.integers(3)), int(rng.integers(npart))) for _ in range(16)]
    eps = 1e-6
    with torch.no_grad():
        for (a, p) in idx:
            dp = psi0.clone()
            dp[a, p] += eps
            dm = psi0.clone()
            dm[a, p] -= eps
            fd = (float(loss_gpu(dp)) - float(loss_gpu(dm))) / (2 * eps)
            an = float(g_adj[a, p])
            assert abs(fd - an) < 1e-5 * max(1.0, abs(an)) + 1e-6 * float(g_adj.abs().max()), (a, p, fd, an)
    plan.close()


def test_fp32_adjoint_close_to_fp64(dev):
    import torch

    ng, nrow, nsteps, dt = 32, 32, 5, 0.01
    npart = nrow ** 3
    Q, psi0, target = _density_loss_setup(dev, ng, nrow, 9)
    grads = []
    for prec in (pmhip.PM_FP64, pmhip.PM_FP32):
        plan = pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0)
        psi = psi0.clone().requires_grad_(True)
        x, v = pag.evolve(torch.remainder(Q + psi, float(ng)), 1.5 * psi, plan, nsteps, dt)
        L = ((pag.density(x, plan).double() - target) ** 2).sum()
        (g,) = torch.autograd.grad(L, psi)
        grads.append(g)
        plan.close()
    err = float((grads[0] - grads[1]).abs().max() / grads[0].abs().max())
    assert err < 1e-3, err


@pytest.mark.parametrize("checkpoint", [False, True])
def test_zero_steps_are_the_identity(dev, checkpoint):
    """pm_run_kdk and pm_run_adjoint of zero steps change nothing (as the
    oracle's run_kdk), and the autograd harness passes the cotangents of a
    zero-step evolution through unchanged."""
    import torch

    ng, nrow = 16, 16
    noise = torch.tensor(np.random.default_rng(8).standard_normal(nrow ** 3), device=dev)
    pos, vel = pmhip.zeldovich_ic(nrow, ng, noise, 1.0, 3.0)
    p0, v0 = [t.clone() for t in pos], [t.clone() for t in vel]
    xb = [torch.full_like(t, 0.25) for t in pos]
    vb = [torch.full_like(t, -0.5) for t in pos]
    with pmhip.Plan(ng, nrow ** 3, pmhip.PM_FP64, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        plan.run_kdk(pos, vel, 0, 0.05, pmhip.PM_DEPOSIT_SORTED)
        plan.run_adjoint(pos, vel, xb, vb, 0, 0.05, pmhip.PM_DEPOSIT_SORTED)
        torch.cuda.synchronize()
        for a, b in zip(pos + vel, p0 + v0):
            assert torch.equal(a, b)
        assert all(bool((t == 0.25).all()) for t in xb) and all(bool((t == -0.5).all()) for t in vb)
        x0 = torch.stack(p0).requires_grad_(True)
        v0s = torch.stack(v0).requires_grad_(True)
        x, v = pag.evolve(x0, v0s, plan, 0, 0.05, checkpoint=checkpoint)
        assert torch.equal(x.detach(), x0.detach()) and torch.equal(v.detach(), v0s.detach())
        (x.sum() * 2.0 + v.sum() * 3.0).backward()
        assert bool((x0.grad == 2.0).all()) and bool((v0s.grad == 3.0).all())


@pytest.mark.parametrize("ng,nrow", [(32, 32), (256, 64)])
@pytest.mark.parametrize("prec", [pmhip.PM_FP64, pmhip.PM_FP32])
def test_run_adjoint_equals_step_by_step(dev, prec, ng, nrow):
    """pm_run_adjoint folds each step's undrift and the max |u| of the
    divergence deposit into the previous reverse step's gather, and stages
    the positions for the next step's deposits; the same reverse sweep as
    pm_adjoint_step calls with their own undrift and binning must match it
    bit for bit (positions, velocities and both cotangents).  256^3: two
    walk segments per tile column."""
    import torch

    print(f"case run_adjoint vs steps ng={ng} nrow={nrow} prec={prec}", flush=True)
    nsteps, dt = 4, 0.02
    npart = nrow ** 3
    noise = np.random.default_rng(31).standard_normal(npart)
    pos0, vel0 = pmhip.zeldovich_ic(nrow, ng, torch.tensor(noise, device=dev), 1.0, 3.0)
    g = torch.Generator(device=dev)
    g.manual_seed(32)
    xb0 = [torch.randn(npart, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    vb0 = [torch.randn(npart, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    out = []
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        for fused in (True, False):
            pos, vel = [t.clone() for t in pos0], [t.clone() for t in vel0]
            xb, vb = [t.clone() for t in xb0], [t.clone() for t in vb0]
            plan.run_kdk(pos, vel, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
            if fused:
                plan.run_adjoint(pos, vel, xb, vb, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
            else:
                for s in range(nsteps, -1, -1):
                    kick = 0.5 * dt if s in (0, nsteps) else dt
                    plan.adjoint_step(pos, vel, xb, vb, kick, 0.0 if s == nsteps else dt, pmhip.PM_DEPOSIT_SORTED)
            torch.cuda.synchronize()
            out.append(pos + vel + xb + vb)
    assert pmhip.range_errors() == 0
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.parametrize("prec", [pmhip.PM_FP32, pmhip.PM_FP64])
def test_run_adjoint_staging_overflow_falls_back(dev, prec):
    """The reverse sweep's gather stages the undrifted positions in tile
    buckets sized from the current counts; velocities that undrift every
    particle towards one point overflow them, and the next reverse step
    must fall back to a full binning: bit for bit the step-by-step sweep."""
    import torch

    ng, nrow, nsteps, dt = 32, 32, 3, 0.05
    npart = nrow ** 3
    rng = np.random.default_rng(33)
    x0 = [torch.tensor(rng.uniform(0, ng, npart), device=dev) for _ in range(3)]
    c = 8.5
    v0 = [(x - c) / dt for x in x0]  # x - dt v = c: all particles into one tile
    g = torch.Generator(device=dev)
    g.manual_seed(34)
    xb0 = [torch.randn(npart, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    vb0 = [torch.randn(npart, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]
    out = []
    with pmhip.Plan(ng, npart, prec, pmhip.PM_GREEN_CONT_K2, 1.0) as plan:
        for fused in (True, False):
            pos, vel = [t.clone() for t in x0], [t.clone() for t in v0]
            xb, vb = [t.clone() for t in xb0], [t.clone() for t in vb0]
            if fused:
                plan.run_adjoint(pos, vel, xb, vb, nsteps, dt, pmhip.PM_DEPOSIT_SORTED)
            else:
                for s in range(nsteps, -1, -1):
                    kick = 0.5 * dt if s in (0, nsteps) else dt
                    plan.adjoint_step(pos, vel, xb, vb, kick, 0.0 if s == nsteps else dt, pmhip.PM_DEPOSIT_SORTED)
            torch.cuda.synchronize()
            out.append(pos + vel + xb + vb)
    assert pmhip.range_errors() == 0
    for a, b in zip(*out):
        assert torch.equal(a, b)
