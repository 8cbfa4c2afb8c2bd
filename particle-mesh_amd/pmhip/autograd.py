"""torch.autograd wrappers of the PM step (test harness for the adjoint).

``evolve(x0, v0, plan, nsteps, dt)`` runs ``pm_run_kdk`` and differentiates
it with the hand-written adjoint kernels: by reverse-time leapfrog
recomputation from the final state (``checkpoint=False``, O(1) memory in the
number of steps, via ``pm_run_adjoint``) or from per-step checkpoints
(``checkpoint=True``, one state per step, exact reversal).
``density(x, plan)`` is the CIC density (``pm_deposit``) with the transposed
CIC (``pm_adjoint_deposit``) as its backward.

Positions and velocities are (3, np) float64 device tensors in cell units.
"""
from __future__ import annotations

import torch

from . import PM_DEPOSIT_SORTED, Plan, mesh_dtype


def _rows(t):
    return [t[0], t[1], t[2]]


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, plan: Plan, nsteps: int, dt: float, variant: int, checkpoint: bool):
        x = x0.detach().clone().contiguous()
        v = v0.detach().clone().contiguous()
        states = []
        if nsteps == 0:
            pass  # zero steps: the identity
        elif checkpoint:
            for s in range(nsteps + 1):
                states.append((x.clone(), v.clone()))
                kick = 0.5 * dt if (s == 0 or s == nsteps) else dt
                drift = 0.0 if s == nsteps else dt
                plan.step(_rows(x), _rows(v), kick, drift, variant)
        else:
            plan.run_kdk(_rows(x), _rows(v), nsteps, dt, variant)
        ctx.plan, ctx.nsteps, ctx.dt, ctx.variant, ctx.checkpoint = plan, nsteps, dt, variant, checkpoint
        ctx.states = states
        ctx.save_for_backward(x, v)
        return x.clone(), v.clone()

    @staticmethod
    def backward(ctx, xbar, vbar):
        xf, vf = ctx.saved_tensors
        plan, nsteps, dt, variant = ctx.plan, ctx.nsteps, ctx.dt, ctx.variant
        x, v = xf.clone(), vf.clone()
        xb = (xbar if xbar is not None else torch.zeros_like(xf)).detach().clone().contiguous()
        vb = (vbar if vbar is not None else torch.zeros_like(vf)).detach().clone().contiguous()
        if nsteps == 0:
            pass
        elif ctx.checkpoint:
            for s in range(nsteps, -1, -1):
                kick = 0.5 * dt if (s == 0 or s == nsteps) else dt
                drift = 0.0 if s == nsteps else dt
                # state after step s: checkpoint s+1 (or the final state)
                if s < nsteps:
                    x.copy_(ctx.states[s + 1][0])
                    v.copy_(ctx.states[s + 1][1])
                plan.adjoint_step(_rows(x), _rows(v), _rows(xb), _rows(vb), kick, drift, variant)
        else:
            plan.run_adjoint(_rows(x), _rows(v), _rows(xb), _rows(vb), nsteps, dt, variant)
        ctx.recovered = (x, v)
        return xb, vb, None, None, None, None, None


def evolve(x0, v0, plan: Plan, nsteps: int, dt: float, variant: int = PM_DEPOSIT_SORTED, checkpoint: bool = False):
    return _Evolve.apply(x0, v0, plan, nsteps, dt, variant, checkpoint)


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan: Plan, variant: int):
        plan.deposit(_rows(x.detach().contiguous()), variant)
        ctx.plan = plan
        ctx.save_for_backward(x)
        return plan.density().clone()

    @staticmethod
    def backward(ctx, rho_bar):
        (x,) = ctx.saved_tensors
        plan = ctx.plan
        xb = torch.zeros_like(x)
        mb = rho_bar.detach().to(mesh_dtype(plan.precision)).contiguous()
        plan.adjoint_deposit(_rows(x.detach().contiguous()), mb, _rows(xb))
        return xb, None, None


def density(x, plan: Plan, variant: int = PM_DEPOSIT_SORTED):
    """mass * CIC density (without the -1) of positions x (3, np)."""
    return _Density.apply(x, plan, variant)
