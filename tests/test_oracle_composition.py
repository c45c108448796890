"""Pins of the oracle's whole-step composition (OracleLFM.step and
backtrack_step), independent of the GPU path and of the reading of the
reference that both share.

Forward: in the unit box, u = (sin pi x cos pi y, -cos pi x sin pi y) is a
steady solution of the incompressible Euler equations with no-penetration
walls (stream function sin pi x sin pi y / pi, vorticity 2 pi^2 times it, so
vorticity is a function of the stream function); extruded along z it is one
in 3D.  OracleLFM.step must keep it, and a passive scalar it carries must
follow the exact characteristics (integrated here to 1e-10 with scipy), to
discretisation error that falls under refinement at a fixed physical time:
 - at second order where a run ends on a reinitialisation (the fields are
   then the BFECC pullback of the interval's start through its flow map);
 - at first order where it ends inside an interval with a fixed physical
   interval length: each of the interval's semi-Lagrangian steps applies the
   quadratic B-spline's O(dx^2) smoothing, and there are O(1/dx) of them.

Adjoint: the reference's adjoint is the continuous adjoint on flow maps, not
the derivative of the discrete forward map, so it differs from central finite
differences of the oracle's own loss by a discretisation gap; that gap must
shrink under refinement (first order is measured), in 2D and in 3D.

Each pin is also run against mistakes planted in the oracle (a dropped
Jacobian in the impulse pullback, a flipped time step in the passive's
semi-Lagrangian advection or in the forward map, a dropped adjoint source,
the convective source accumulated at the local step's dt): every planted
mistake fails the pin it is aimed at.  The reference: forward step
2D/vortex/lfm_midpoint_simulator.py:517-661 and 3D/lfm_midpoint_simulator.py:648-807;
adjoint sweep 2D/vortex/lfm_midpoint_diff_simulator.py:260-503 and
3D/lfm_midpoint_diff_simulator.py:421-594.
"""
import contextlib

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import lfm_oracle as O


# ------------------------------------------------------------- planting ---
@contextlib.contextmanager
def planted(name, wrap):
    """Replace oracle function `name` by wrap(original) for the block."""
    old = getattr(O, name)
    setattr(O, name, wrap(old))
    try:
        yield
    finally:
        setattr(O, name, old)


def _pullback_without_jacobian(old):
    def advect_u(u0, fmap, dx, kern=2):
        out = []
        for a in range(len(u0)):
            eye = np.zeros_like(fmap.T_f[a])
            eye[..., a] = 1.0
            out.append(O.pullback(u0, fmap.psi_f[a], eye, dx, kern))
        return out
    return advect_u


def _passive_advected_backwards(old):
    return lambda vel, s, dx, dt, order, kern=2: old(vel, s, dx, -dt, order, kern)


@contextlib.contextmanager
def forward_map_flipped():
    """phi (the forward map) marched with +dt instead of -dt."""
    old_march, old_init = O.FlowMap.march, O.OracleLFM.__init__
    current = []

    def init(self, *a, **k):
        old_init(self, *a, **k)
        current.append(self)

    def march(self, u, dt, order):
        return old_march(self, u, -dt if current and self is current[-1].phi_map else dt, order)

    O.FlowMap.march, O.OracleLFM.__init__ = march, init
    try:
        yield
    finally:
        O.FlowMap.march, O.OracleLFM.__init__ = old_march, old_init


# ------------------------------------------------------- steady flow pin ---
def _tg_faces(n, dx):
    out = []
    for a in range(len(n)):
        x = O.face_coords(n, a, dx)
        if a == 0:
            out.append(np.sin(np.pi * x[..., 0]) * np.cos(np.pi * x[..., 1]))
        elif a == 1:
            out.append(-np.cos(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 1]))
        else:
            out.append(np.zeros(x.shape[:-1]))
    return out


def _tg_velocity(z, dim):
    x = z.reshape(dim, -1)
    v = [np.sin(np.pi * x[0]) * np.cos(np.pi * x[1]), -np.cos(np.pi * x[0]) * np.sin(np.pi * x[1])]
    if dim == 3:
        v.append(np.zeros_like(x[2]))
    return np.concatenate(v)


def _blob(p, centre, sigma):
    return np.exp(-sum((p[..., k] - centre[k]) ** 2 for k in range(p.shape[-1])) / (2 * sigma ** 2))


def steady_flow_errors(nn, dim, T, R):
    """rel-L2 error of u and of the passive after time T (dt = dx/2) of the
    steady Taylor-Green flow on an nn^dim grid, reinitialising every R."""
    n = (nn,) * dim
    dx = 1.0 / nn
    dt = 0.5 * dx
    steps = int(round(T / dt))
    centre, sigma = (0.5, 0.3, 0.5)[:dim], 0.08 if dim == 2 else 0.1
    u0 = _tg_faces(n, dx)
    xc = O.cell_coords(n, dx)
    orc = O.OracleLFM(n, dx, dt, R, 2 if dim == 2 else 4)
    orc.init(u0, [_blob(xc, centre, sigma)])
    for _ in range(steps):
        orc.step()
    # exact characteristics: s(x, t) = s0(X(-t; x))
    z0 = np.concatenate([xc[..., k].ravel() for k in range(dim)])
    sol = solve_ivp(lambda t, z: -_tg_velocity(z, dim), (0.0, steps * dt), z0, rtol=1e-10, atol=1e-12,
                    method="DOP853")
    X = np.moveaxis(sol.y[:, -1].reshape((dim,) + n), 0, -1)
    s_exact = _blob(X, centre, sigma)
    e_u = np.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(orc.u, u0)) / sum(np.sum(b ** 2) for b in u0))
    e_s = np.linalg.norm(orc.s[0] - s_exact) / np.linalg.norm(s_exact)
    return e_u, e_s


def orders(errs, sizes):
    return [float(np.log(errs[i] / errs[i + 1]) / np.log(sizes[i + 1] / sizes[i])) for i in range(len(errs) - 1)]


def _orders_of(dim, sizes, T, R_of):
    e = [steady_flow_errors(nn, dim, T, R_of(nn)) for nn in sizes]
    return e, orders([x[0] for x in e], sizes), orders([x[1] for x in e], sizes)


# 2D at reinitialisation points: 16 / 32 / 64 steps at 32^2 / 64^2 / 128^2, R = 8
REINIT_2D = dict(dim=2, sizes=(32, 64, 128), T=0.25, R_of=lambda nn: 8)
# 2D inside an interval, interval 3/64 of unit time: 20 / 40 / 80 steps, R = 3 / 6 / 12
MID_2D = dict(dim=2, sizes=(32, 64, 128), T=0.3125, R_of=lambda nn: 3 * nn // 32)


def test_steady_flow_2d_second_order_at_reinitialisation():
    e, ou, os_ = _orders_of(**REINIT_2D)
    assert min(ou) > 1.8 and min(os_) > 1.8, (e, ou, os_)
    assert e[-1][0] < 3e-4 and e[-1][1] < 2e-3, e


def test_steady_flow_2d_first_order_inside_interval():
    e, ou, os_ = _orders_of(**MID_2D)
    assert min(ou) > 0.9 and min(os_) > 0.9, (e, ou, os_)
    assert e[-1][0] < 3e-3 and e[-1][1] < 2e-2, e


def test_steady_flow_3d():
    # 6 / 12 steps at 16^3 / 32^3: on a reinitialisation with R = 3, inside
    # an interval with R = 4 / 8
    e, ou, os_ = _orders_of(dim=3, sizes=(16, 32), T=0.1875, R_of=lambda nn: 3)
    assert ou[0] > 1.8 and os_[0] > 1.8, (e, ou, os_)
    e, ou, os_ = _orders_of(dim=3, sizes=(16, 32), T=0.1875, R_of=lambda nn: nn // 4)
    assert ou[0] > 0.9 and os_[0] > 0.9, (e, ou, os_)


def _caught(setup, ctx):
    with ctx():
        e, ou, os_ = _orders_of(**dict(setup, sizes=setup["sizes"][:2]))
    e0, ou0, os0 = _orders_of(**dict(setup, sizes=setup["sizes"][:2]))
    # an order well below the correct one, or an error several times larger
    return (ou[0] < ou0[0] - 0.5 or os_[0] < os0[0] - 0.5 or e[-1][0] > 4 * e0[-1][0]
            or e[-1][1] > 4 * e0[-1][1]), (e, ou, os_, e0)


@pytest.mark.parametrize("mistake,setup", [
    ("pullback_without_jacobian", "reinit"), ("forward_map_flipped", "reinit"),
    ("passive_advected_backwards", "mid")])
def test_steady_flow_pin_catches_planted_mistakes(mistake, setup):
    ctx = {"pullback_without_jacobian": lambda: planted("advect_u", _pullback_without_jacobian),
           "passive_advected_backwards": lambda: planted("sl_advect_s", _passive_advected_backwards),
           "forward_map_flipped": forward_map_flipped}[mistake]
    ok, info = _caught(REINIT_2D if setup == "reinit" else MID_2D, ctx)
    assert ok, (mistake, info)


# ---------------------------------------------------- adjoint gap pin ------
# Smooth problems whose physical size does not change under refinement: the
# reinitialisation interval is 1/16 of unit time (R = 1/16 / dt steps) and the
# run is two intervals; the loss is the reference's sum (s_T - s*)^2.
THETA_2D = np.array([1.0, -0.8, 0.9, -1.1, 0.35, 0.62, 0.45, 0.68, 0.38, 0.40, 0.66, 0.62, 2.4, 2.0, 2.2, 2.6])
DIRS_2D = [np.array(t) for t in ([1.0, 0.45, 0.55, 3.0], [-1.0, 0.6, 0.4, 2.5], [0.7, 0.35, 0.35, 3.2],
                                 [-0.6, 0.55, 0.6, 2.8])]
TUBES_3D = np.array([[1.0, 0.45, 0.5, 0.5, 0, 0, 1, 0.2], [-0.9, 0.6, 0.45, 0.55, 0.6, 0.8, 0, 0.18],
                     [0.8, 0.5, 0.62, 0.4, 1, 0, 0, 0.22]])
DIRS_3D = [np.array([[1.0, 0.5, 0.45, 0.5, 0, 1, 0, 0.22]]), np.array([[-1.0, 0.55, 0.55, 0.45, 0.8, 0, 0.6, 0.2]])]


def _problem(dim, n, dx):
    from dfm import scenes

    if dim == 2:
        u0 = scenes.split_central_vector(scenes.vortex_velocity_2d(THETA_2D, n, dx))
        s0 = scenes.disk_density(n, dx, radius=0.18, width=0.03)
        tgt = scenes.gaussian_blobs(n, dx, 7, 3, sigma=0.1)
        dirs = [scenes.split_central_vector(scenes.vortex_velocity_2d(t, n, dx)) for t in DIRS_2D]
    else:
        u0 = scenes.split_central_vector(scenes.vortex_tubes_3d(TUBES_3D, n, dx))
        s0 = scenes.disk_density(n, dx, radius=0.22, width=0.06)
        tgt = scenes.gaussian_blobs(n, dx, 9, 2, sigma=0.15)
        dirs = [scenes.split_central_vector(scenes.vortex_tubes_3d(t, n, dx)) for t in DIRS_3D]
    return u0, s0, tgt, dirs


def adjoint_fd_gap(dim, nn, h=1e-3):
    """rel-L2 over the directions d of <lam_u(0), d> (the oracle's adjoint
    sweep) against central differences (L(u0 + h d) - L(u0 - h d)) / 2h of
    the oracle's own forward loss."""
    n = (nn,) * dim
    dx = 1.0 / nn
    dt = 0.5 * dx
    R = int(round((1.0 / 16) / dt))
    steps = 2 * R
    u0, s0, tgt, dirs = _problem(dim, n, dx)

    def run(u):
        orc = O.OracleLFM(n, dx, dt, R, 2 if dim == 2 else 4)
        orc.init(u, [s0])
        for _ in range(steps):
            orc.step()
        return float(np.sum((orc.s[0] - tgt) ** 2)), orc

    _, orc = run(u0)
    orc.init_gradient([tgt])
    for _ in range(steps):
        orc.backtrack_step()
    g = np.array([sum(np.sum(l * c) for l, c in zip(orc.lam_u, d)) for d in dirs])
    fd = np.array([(run([a + h * b for a, b in zip(u0, d)])[0] - run([a - h * b for a, b in zip(u0, d)])[0]) / (2 * h)
                   for d in dirs])
    return float(np.linalg.norm(g - fd) / np.linalg.norm(fd))


def test_adjoint_gap_shrinks_2d():
    gaps = [adjoint_fd_gap(2, nn) for nn in (32, 64, 128)]
    # measured: 4.8e-2, 2.4e-2, 1.1e-2 (first order)
    assert gaps[0] / gaps[1] > 1.6 and gaps[1] / gaps[2] > 1.6, gaps
    assert gaps[-1] < 0.02, gaps


def test_adjoint_gap_shrinks_3d():
    gaps = [adjoint_fd_gap(3, nn) for nn in (16, 24)]
    # measured: 0.19, 0.087 (and 0.050 at 32^3): second order
    assert orders(gaps, (16, 24))[0] > 1.4, gaps
    assert gaps[-1] < 0.12, gaps


def _ab3_at_local_dt(old):
    return lambda f, init, fmap, H1, H2, ind, dx, dt: old(f, init, fmap, H1, H2, ind, dx, 0.5 * dt)


def _scalar_source_dropped(old):
    return lambda s, lam, dx: [np.zeros_like(c) for c in old(s, lam, dx)]


def _convective_source_doubled(old):
    return lambda u, w, dx, kern: [2 * c for c in old(u, w, dx, kern)]


@pytest.mark.parametrize("name,wrap", [("accumulate_ab3", _ab3_at_local_dt),
                                       ("adj_scalar_source", _scalar_source_dropped),
                                       ("adj_conv_source", _convective_source_doubled)])
def test_adjoint_gap_pin_catches_planted_mistakes(name, wrap):
    with planted(name, wrap):
        gaps = [adjoint_fd_gap(2, nn) for nn in (32, 64)]
    # the correct sweep: 4.8e-2 -> 2.4e-2
    assert gaps[0] / gaps[1] < 1.3 or gaps[1] > 0.07, gaps
