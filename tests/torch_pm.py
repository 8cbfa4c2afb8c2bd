"""Differentiable float64 PM step in plain torch (test-side reference).

Used only by the tests, as the FP64 reference gradient for the hand-written
adjoint: torch's autograd differentiates this restatement of the forward
step (CIC deposit via index_add_, FFT solve via torch.fft, centred
difference via roll, trilinear interpolation via gathers), so it shares no
code with the adjoint kernels.
"""
import math

import torch


def _corners(x, ng):
    i = torch.floor(x)
    t = x - i
    i = i.long()
    return i, t


def deposit(x, ng, mass):
    """x (3, np) cell units -> mass * CIC on the periodic ng^3 mesh."""
    i, t = _corners(x, ng)
    rho = torch.zeros(ng ** 3, dtype=x.dtype, device=x.device)
    for a in (0, 1):
        wx = t[0] if a else 1 - t[0]
        ix = (i[0] + a) % ng
        for b in (0, 1):
            wy = t[1] if b else 1 - t[1]
            iy = (i[1] + b) % ng
            for c in (0, 1):
                wz = t[2] if c else 1 - t[2]
                iz = (i[2] + c) % ng
                rho = rho.index_add(0, (ix * ng + iy) * ng + iz, mass * wx * wy * wz)
    return rho.reshape(ng, ng, ng)


def green_cont_k2(ng, device):
    n = torch.fft.fftfreq(ng, d=1.0 / ng, dtype=torch.float64, device=device)
    nz = torch.arange(ng // 2 + 1, dtype=torch.float64, device=device)
    n2 = n[:, None, None] ** 2 + n[None, :, None] ** 2 + nz[None, None, :] ** 2
    g = -(ng * ng) / (math.pi * n2)
    g[0, 0, 0] = 0.0
    return g


def solve(delta, green):
    ng = delta.shape[0]
    return torch.fft.irfftn(torch.fft.rfftn(delta) * green, s=(ng, ng, ng))


def interp_force(x, phi):
    ng = phi.shape[0]
    g = [(torch.roll(phi, 1, d) - torch.roll(phi, -1, d)) * 0.5 for d in range(3)]
    i, t = _corners(x, ng)
    out = []
    for f in g:
        ff = f.reshape(-1)
        acc = 0.0
        for a in (0, 1):
            wx = t[0] if a else 1 - t[0]
            ix = (i[0] + a) % ng
            for b in (0, 1):
                wy = t[1] if b else 1 - t[1]
                iy = (i[1] + b) % ng
                for c in (0, 1):
                    wz = t[2] if c else 1 - t[2]
                    iz = (i[2] + c) % ng
                    acc = acc + wx * wy * wz * ff[(ix * ng + iy) * ng + iz]
        out.append(acc)
    return torch.stack(out)


def force(x, ng, mass, green):
    rho = deposit(x, ng, mass)
    return interp_force(x, solve(rho - 1.0, green))


def run_kdk(x, v, ng, nsteps, dt, mass, green):
    """Merged-kick KDK of pm_run_kdk (nsteps + 1 force evaluations; none
    for zero steps)."""
    if nsteps == 0:
        return x, v
    for s in range(nsteps + 1):
        kick = 0.5 * dt if (s == 0 or s == nsteps) else dt
        drift = 0.0 if s == nsteps else dt
        v = v + kick * force(x, ng, mass, green)
        if drift:
            x = torch.remainder(x + drift * v, float(ng))
    return x, v
