"""torch.autograd front end of the flow-map simulator and the 2D vortex
initializer.

* ``VortexInit2D``: theta (4K, the reference's control layout) -> MAC initial
  velocity (unprojected), with the analytic vector-Jacobian product kernel in
  backward (replaces calculate_dtheta's finite differences,
  2D/vortex/optimize_vortex.py:193-226).
* Per-stage Functions with exact discrete vector-Jacobian products (HIP
  kernels, no counterpart in the reference): ``Project`` (the MGPCG
  projection; symmetric, so its VJP is itself), ``March`` (RK2/RK4 march of a
  flow map through a velocity, differentiable in the velocity and the map),
  ``PullbackU`` / ``PullbackS`` (T . u0(psi) and s0(psi), differentiable in
  the source and the map) and ``SLAdvect`` (semi-Lagrangian advection,
  differentiable in the sources and the advecting velocity).  Composing them
  gives the derivative of the discrete forward map itself; the
  scatter-based transposes use FP32 atomic adds, so gradients agree to
  rounding (not bitwise) from run to run.
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
from .ops import FlowMap, MacField


def _zeros_like_or(g, like):
    return (g if g is not None else torch.zeros_like(like)).float().contiguous()


class Project(torch.autograd.Function):
    """u -> P u, the projection of an MGPCG ``solver`` (2D/vortex/mgpcg.py:363-368)."""

    @staticmethod
    def forward(ctx, solver, *u):
        out = MacField([c.detach().float().contiguous().clone() for c in u])
        solver.Poisson(out)
        ctx.solver = solver
        ctx.like = [c.detach() for c in u]
        return tuple(out.c)

    @staticmethod
    def backward(ctx, *g):
        gg = MacField([_zeros_like_or(x, c).clone() for x, c in zip(g, ctx.like)])
        ctx.solver.Poisson_adj(gg)
        return (None,) + tuple(gg.c)


class March(torch.autograd.Function):
    """(u, psi_f, T_f[, psi]) -> the map marched one RK step of size dt
    through u (march_psi_grid / RK4 of 2D/vortex/flowmap.py:83-231)."""

    @staticmethod
    def forward(ctx, ops, dt, order, has_cells, *t):
        d = ops.dim
        u = MacField([x.detach().contiguous() for x in t[:d]])
        psi_f = [x.detach().contiguous() for x in t[d:2 * d]]
        T_f = [x.detach().contiguous() for x in t[2 * d:3 * d]]
        psi = t[3 * d].detach().contiguous() if has_cells else None
        out = FlowMap.wrap(psi.clone() if has_cells else None, [x.clone() for x in psi_f], [x.clone() for x in T_f])
        ops.march(u, out, dt, order)
        ctx.ops, ctx.dt, ctx.order, ctx.has_cells = ops, float(dt), int(order), has_cells
        ctx.save_for_backward(*u.c, *psi_f, *T_f, *([psi] if has_cells else []))
        return tuple(out.psi_f) + tuple(out.T_f) + ((out.psi,) if has_cells else ())

    @staticmethod
    def backward(ctx, *g):
        d = ctx.ops.dim
        sv = ctx.saved_tensors
        u = MacField(sv[:d])
        psi_f, T_f = list(sv[d:2 * d]), list(sv[2 * d:3 * d])
        psi = sv[3 * d] if ctx.has_cells else None
        fm0 = FlowMap.wrap(psi, psi_f, T_f)
        gfm = FlowMap.wrap(_zeros_like_or(g[2 * d], psi).clone() if ctx.has_cells else None,
                           [_zeros_like_or(x, c).clone() for x, c in zip(g[:d], psi_f)],
                           [_zeros_like_or(x, c).clone() for x, c in zip(g[d:2 * d], T_f)])
        gu = MacField([torch.zeros_like(c) for c in u.c])
        ctx.ops.march_vjp(u, fm0, ctx.dt, ctx.order, gfm, gu)
        return (None, None, None, None) + tuple(gu.c) + tuple(gfm.psi_f) + tuple(gfm.T_f) + \
            ((gfm.psi,) if ctx.has_cells else ())


class PullbackU(torch.autograd.Function):
    """(src, psi_f, T_f) -> alpha T_a . src(psi_a) per face grid a (the
    flow-map pullback of the impulse, 2D/vortex/flowmap.py:187-231)."""

    @staticmethod
    def forward(ctx, ops, alpha, kern, *t):
        d = ops.dim
        src = MacField([x.detach().contiguous() for x in t[:d]])
        psi_f = [x.detach().contiguous() for x in t[d:2 * d]]
        T_f = [x.detach().contiguous() for x in t[2 * d:3 * d]]
        out = MacField([torch.empty_like(c) for c in src.c])
        ops.pullback_u(src, FlowMap.wrap(None, psi_f, T_f), out, alpha, kern=kern)
        ctx.ops, ctx.alpha, ctx.kern = ops, float(alpha), int(kern)
        ctx.save_for_backward(*src.c, *psi_f, *T_f)
        return tuple(out.c)

    @staticmethod
    def backward(ctx, *g):
        d = ctx.ops.dim
        sv = ctx.saved_tensors
        src = MacField(sv[:d])
        fm = FlowMap.wrap(None, list(sv[d:2 * d]), list(sv[2 * d:3 * d]))
        gsrc = MacField([torch.zeros_like(c) for c in src.c])
        want_map = any(ctx.needs_input_grad[3 + d:])
        gfm = FlowMap.wrap(None, [torch.zeros_like(x) for x in fm.psi_f],
                           [torch.zeros_like(x) for x in fm.T_f]) if want_map else None
        gout = MacField([_zeros_like_or(x, c) for x, c in zip(g, src.c)])
        ctx.ops.pullback_u_vjp(gout, fm, gsrc, ctx.alpha, src=src if want_map else None, g_fm=gfm, kern=ctx.kern)
        mg = tuple(gfm.psi_f) + tuple(gfm.T_f) if want_map else (None,) * (2 * d)
        return (None, None, None) + tuple(gsrc.c) + mg


class PullbackS(torch.autograd.Function):
    """(src, psi) -> alpha src(psi) at the cell centres (the flow-map
    pullback of a passive scalar)."""

    @staticmethod
    def forward(ctx, ops, alpha, kern, src, psi):
        src, psi = src.detach().contiguous(), psi.detach().contiguous()
        out = torch.empty_like(src)
        ops.pullback_s(src, psi, out, alpha, kern=kern)
        ctx.ops, ctx.alpha, ctx.kern = ops, float(alpha), int(kern)
        ctx.save_for_backward(src, psi)
        return out

    @staticmethod
    def backward(ctx, g):
        src, psi = ctx.saved_tensors
        gsrc = torch.zeros_like(src)
        gpsi = torch.zeros_like(psi) if ctx.needs_input_grad[4] else None
        ctx.ops.pullback_s_vjp(_zeros_like_or(g, src), psi, gsrc, ctx.alpha, src=src if gpsi is not None else None,
                               g_psi=gpsi, kern=ctx.kern)
        return None, None, None, gsrc, gpsi


class SLAdvect(torch.autograd.Function):
    """(vel, nvec MAC sources, nscal cell sources) -> each source sampled at
    the RK back-trace of its nodes through vel over dt (advect_u_grid /
    advect_scalar_grid, 2D/vortex/flowmap.py:59-81)."""

    @staticmethod
    def forward(ctx, ops, dt, order, kern, nvec, nscal, *t):
        d = ops.dim
        vel = MacField([x.detach().contiguous() for x in t[:d]])
        vsrc = [MacField([x.detach().contiguous() for x in t[d + m * d:d + (m + 1) * d]]) for m in range(nvec)]
        ssrc = [x.detach().contiguous() for x in t[d + nvec * d:]]
        vdst = [MacField([torch.empty_like(c) for c in m.c]) for m in vsrc]
        sdst = [torch.empty_like(x) for x in ssrc]
        ops.sl_advect(vel, dt, order, vsrc, vdst, ssrc, sdst, kern=kern)
        ctx.ops, ctx.dt, ctx.order, ctx.nvec, ctx.nscal = ops, float(dt), int(order), nvec, nscal
        ctx.kern = int(kern)
        ctx.save_for_backward(*vel.c, *[c for m in vsrc for c in m.c], *ssrc)
        return tuple(c for m in vdst for c in m.c) + tuple(sdst)

    @staticmethod
    def backward(ctx, *g):
        d, nv, ns = ctx.ops.dim, ctx.nvec, ctx.nscal
        sv = ctx.saved_tensors
        vel = MacField(sv[:d])
        vsrc = [MacField(sv[d + m * d:d + (m + 1) * d]) for m in range(nv)]
        ssrc = list(sv[d + nv * d:])
        gvd = [MacField([_zeros_like_or(g[m * d + a], vsrc[m].c[a]) for a in range(d)]) for m in range(nv)]
        gsd = [_zeros_like_or(g[nv * d + k], ssrc[k]) for k in range(ns)]
        gvs = [MacField([torch.zeros_like(c) for c in m.c]) for m in vsrc]
        gss = [torch.zeros_like(x) for x in ssrc]
        want_vel = any(ctx.needs_input_grad[6:6 + d])
        gvel = MacField([torch.zeros_like(c) for c in vel.c]) if want_vel else None
        ctx.ops.sl_advect_vjp(vel, ctx.dt, ctx.order, gvd, gvs, gsd, gss,
                              vec_src=vsrc if want_vel else (), s_src=ssrc if want_vel else (), g_vel=gvel,
                              kern=ctx.kern)
        return (None,) * 6 + (tuple(gvel.c) if want_vel else (None,) * d) + \
            tuple(c for m in gvs for c in m.c) + tuple(gss)


def project(solver, u):
    return list(Project.apply(solver, *u))


def march(ops, u, psi_f, T_f, dt, order, psi=None):
    """Differentiable march: returns (psi_f, T_f, psi) after one RK step."""
    d = ops.dim
    out = March.apply(ops, float(dt), int(order), psi is not None, *u, *psi_f, *T_f, *([psi] if psi is not None else []))
    return list(out[:d]), list(out[d:2 * d]), (out[2 * d] if psi is not None else None)


def pullback_u(ops, src, psi_f, T_f, alpha=1.0, kern=2):
    """Differentiable pullback (kern: sampling spline, 1 linear / 2 quadratic / 3 cubic)."""
    return list(PullbackU.apply(ops, float(alpha), int(kern), *src, *psi_f, *T_f))


def pullback_s(ops, src, psi, alpha=1.0, kern=2):
    return PullbackS.apply(ops, float(alpha), int(kern), src, psi)


def sl_advect(ops, vel, dt, order, vec_src=(), s_src=(), kern=2):
    """Differentiable sl_advect: returns ([advected MAC fields], [scalars])."""
    d = ops.dim
    out = SLAdvect.apply(ops, float(dt), int(order), int(kern), len(vec_src), len(s_src), *vel,
                         *[c for m in vec_src for c in m], *s_src)
    nv = len(vec_src)
    return [list(out[m * d:(m + 1) * d]) for m in range(nv)], list(out[nv * d:])


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
