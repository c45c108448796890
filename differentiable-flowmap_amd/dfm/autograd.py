"""torch.autograd front end of the flow-map simulator and the 2D vortex
initializer.

* ``VortexInit2D``: theta (4K, the reference's control layout) -> MAC initial
  velocity (unprojected), with the analytic vector-Jacobian product kernel in
  backward (replaces calculate_dtheta's finite differences,
  2D/vortex/optimize_vortex.py:193-226).
* ``LFMRollout``: (s0, u0) -> (s_T, u_T) over ``steps`` forward steps of an
  ``LFM_Diff_Simulator``; backward runs the simulator's continuous adjoint
  sweep (LFM_Diff_Simulator.init_gradient / backtrack_step_midpoint,
  2D/vortex/lfm_midpoint_diff_simulator.py:126-503) seeded with the incoming
  gradients.  The simulator is switched to per-interval checkpointing, so the
  sweep recomputes each interval's midpoint velocities and forward map from
  the interval-start state and holds O(reinit interval) of them.

The gradient is the reference's: the continuous adjoint on flow maps, not the
exact derivative of the discrete forward map (the two differ by the
discretisation gap of the adjoint equations; on the 2D vortex problem the
whole-gradient cosine against central differences is ~0.999).  Each
``LFMRollout`` forward must be followed by at most one backward.
"""
from __future__ import annotations

import torch

from . import _lib as L
from .ops import MacField


class VortexInit2D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, n, dx):
        theta = theta.detach().contiguous()
        if theta.dtype != torch.float32 or not theta.is_cuda:
            raise TypeError("theta must be a float32 HIP tensor")
        if theta.numel() % 4:
            raise ValueError("theta holds 4 blocks of K values (w, cx, cy, rr)")
        K = theta.numel() // 4
        g = L.make_grid(n, dx)
        u = MacField.empty(tuple(n), theta.device)
        L.check(L.lib().dfm_vortex_velocity_2d(g, theta.data_ptr(), K, u.s, L.stream_handle()), "vortex_velocity_2d")
        ctx.save_for_backward(theta)
        ctx.n, ctx.dx, ctx.K = tuple(n), float(dx), K
        return tuple(u.c)

    @staticmethod
    def backward(ctx, gx, gy):
        (theta,) = ctx.saved_tensors
        zero = lambda like_shape: torch.zeros(like_shape, dtype=torch.float32, device=theta.device)  # noqa: E731
        n = ctx.n
        lam = MacField([
            gx.contiguous().float() if gx is not None else zero((n[0] + 1, n[1])),
            gy.contiguous().float() if gy is not None else zero((n[0], n[1] + 1)),
        ])
        dth = torch.empty_like(theta)
        g = L.make_grid(n, ctx.dx)
        L.check(L.lib().dfm_vortex_velocity_2d_vjp(g, theta.data_ptr(), ctx.K, lam.s, dth.data_ptr(),
                                                   L.stream_handle()), "vortex_velocity_2d_vjp")
        return dth, None, None


def vortex_velocity_2d(theta, n, dx):
    """Differentiable MAC initial velocity (u_x, u_y) of the vortex sum."""
    return VortexInit2D.apply(theta, tuple(n), float(dx))


class LFMRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, steps, s0, *u0):
        sim.checkpoint = True
        sim.record = True
        sim.total_steps = steps
        sim.init(u0=[c.detach() for c in u0], init_passive_np=[s0.detach()])
        for _ in range(steps):
            sim.step_midpoint()
        ctx.sim, ctx.steps = sim, steps
        return (sim.passive[0].clone(),) + tuple(c.clone() for c in sim.u.c)

    @staticmethod
    def backward(ctx, g_s, *g_u):
        sim = ctx.sim
        if sim is None:
            raise RuntimeError("LFMRollout supports one backward per forward")
        ctx.sim = None
        gs = g_s if g_s is not None else torch.zeros_like(sim.passive[0])
        gu = None
        if any(g is not None for g in g_u):
            gu = [g if g is not None else torch.zeros_like(c) for g, c in zip(g_u, sim.u.c)]
        sim.init_gradient(grad_passive=[gs.float().contiguous()],
                          grad_u=None if gu is None else [g.float().contiguous() for g in gu])
        for _ in range(ctx.steps):
            sim.backtrack_step_midpoint()
        return (None, None, sim.adj_passive[0].clone()) + tuple(c.clone() for c in sim.adj_u.c)


def rollout(sim, steps, s0, u0):
    """(s_T, [u_T...]) after `steps` forward steps from (s0, u0), differentiable
    in s0 and u0 (u0 is projected by the simulator's init, so its gradient is
    divergence-free)."""
    out = LFMRollout.apply(sim, int(steps), s0, *u0)
    return out[0], list(out[1:])


def density_loss(sim, steps, s0, u0, target):
    """Final-time density loss sum (s_T - s*)^2 (the reference's objective,
    2D/vortex/lfm_midpoint_diff_simulator.py:159-174)."""
    s_T, _ = rollout(sim, steps, s0, u0)
    return torch.sum((s_T.double() - target.double()) ** 2)


def calculate_dtheta(simulator, theta):
    """dL/dtheta from the adjoint velocity at t = 0 held by `simulator`
    (after its backtrack sweep): the reference's calculate_dtheta
    (2D/vortex/optimize_vortex.py:222-226) evaluated with the analytic
    Jacobian.  Returns a list of 4K floats."""
    th = torch.as_tensor(theta, dtype=torch.float32).reshape(-1).to(simulator.device).contiguous()
    g = L.make_grid(simulator.n, simulator.dx)
    out = torch.empty_like(th)
    L.check(L.lib().dfm_vortex_velocity_2d_vjp(g, th.data_ptr(), th.numel() // 4, simulator.adj_u.s, out.data_ptr(),
                                               L.stream_handle()), "vortex_velocity_2d_vjp")
    return out.cpu().tolist()
