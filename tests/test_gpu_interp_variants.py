"""The linear (N_1) and cubic (N_3) sampling variants of the pullbacks, BFECC
and semi-Lagrangian advection (the _k entry points), against the FP64 oracle
on the same float32-rounded inputs: kernels at rel-L2 <= 1e-5, whole forward
and adjoint trajectories of the simulator with interp_kern 1 / 3 at the
trajectory bounds, and the transposes behind the autograd Functions by the
dot-product identity.  The oracle's N_1 interpolation is pinned against
scipy's RegularGridInterpolator and exact affine reproduction (CPU tests),
its N_3 against scipy's B-spline basis."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, fmap_to_dev, mac_to_dev, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
GRIDS = {2: ((24, 32), 1 / 32), 3: ((12, 16, 20), 1 / 20)}


def rand_mac(rng, n):
    return round32([rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))])


def smooth_mac(rng, n, dx):
    out = []
    for a in range(len(n)):
        X = O.face_coords(n, a, dx)
        v = np.zeros(X.shape[:-1])
        for _ in range(3):
            k = rng.uniform(1, 4, len(n)) * np.pi
            ph = rng.uniform(0, 2 * np.pi, len(n))
            v += np.prod(np.cos(k * X + ph), axis=-1) * rng.uniform(0.3, 0.8)
        out.append(v)
    return round32(out)


def random_map(rng, n, dx, disp=2.5):
    """Displaced points (some pushed past the walls, so the clamps act) and
    perturbed Jacobian columns."""
    om = O.FlowMap(n, dx, np.float64)
    om.psi = round32(om.psi + disp * dx * rng.uniform(-1, 1, om.psi.shape))
    for a in range(len(n)):
        om.psi_f[a] = round32(om.psi_f[a] + disp * dx * rng.uniform(-1, 1, om.psi_f[a].shape))
        om.T_f[a] = round32(om.T_f[a] + 0.2 * rng.standard_normal(om.T_f[a].shape))
    return om


@pytest.mark.parametrize("kern", [1, 3])
@pytest.mark.parametrize("dim", [2, 3])
def test_pullbacks_and_bfecc(dim, kern, device):
    from dfm.ops import MacField, Ops

    n, dx = GRIDS[dim]
    rng = np.random.default_rng(100 + 10 * dim + kern)
    ops = Ops(n, dx)
    u0, add = rand_mac(rng, n), rand_mac(rng, n)
    om1, om2 = random_map(rng, n, dx), random_map(rng, n, dx)
    fm1, fm2 = fmap_to_dev(om1, device), fmap_to_dev(om2, device)
    # impulse pullback with its epilogue
    out = mac_to_dev(add, device)
    ops.pullback_u(mac_to_dev(u0, device), fm1, out, alpha=0.7, beta=-1.3, add=mac_to_dev(add, device), kern=kern)
    ref = [0.7 * r - 1.3 * c for r, c in zip(O.advect_u(u0, om1, dx, kern), add)]
    within(f"pullback_u k{kern}", rel_l2_fields(mac_to_np(out), ref), TOL)
    # BFECC of a MAC field
    err, out = MacField.empty(n, device), MacField.empty(n, device)
    ops.bfecc_u(mac_to_dev(u0, device), fm1, fm2, err, out, kern=kern)
    within(f"bfecc_u k{kern}", rel_l2_fields(mac_to_np(out), O.bfecc(u0, om1, om2, dx, kern=kern)), TOL)
    # scalar pullback and BFECC
    s0 = round32(rng.standard_normal(n))
    s_d = torch.tensor(s0, dtype=torch.float32, device=device)
    so = torch.empty_like(s_d)
    ops.pullback_s(s_d, fm1.psi, so, kern=kern)
    within(f"pullback_s k{kern}", rel_l2(so.cpu().numpy(), O.advect_scalar(s0, om1.psi, dx, kern)), TOL)
    es = torch.empty_like(s_d)
    ops.bfecc_s(s_d, fm1.psi, fm2.psi, es, so, kern=kern)
    within(f"bfecc_s k{kern}", rel_l2(so.cpu().numpy(), O.bfecc_scalar(s0, om1.psi, om2.psi, dx, kern=kern)), TOL)
    # the variant differs from the quadratic result (the kernel is really switched)
    q = torch.empty_like(s_d)
    ops.pullback_s(s_d, fm1.psi, q)
    assert rel_l2(so.cpu().numpy(), q.cpu().numpy()) > 1e-2


@pytest.mark.parametrize("kern", [1, 3])
@pytest.mark.parametrize("dim,order", [(2, 2), (3, 4)])
def test_sl_advect(dim, order, kern, device):
    from dfm.ops import MacField, Ops

    n, dx = GRIDS[dim]
    rng = np.random.default_rng(200 + 10 * dim + order + kern)
    vel = smooth_mac(rng, n, dx)
    src = rand_mac(rng, n)
    ss = [round32(rng.standard_normal(n)) for _ in range(2)]
    dt = 0.8 * dx
    dst = MacField.empty(n, device)
    s_src = [torch.tensor(s, dtype=torch.float32, device=device) for s in ss]
    s_dst = [torch.empty_like(t) for t in s_src]
    Ops(n, dx).sl_advect(mac_to_dev(vel, device), dt, order, [mac_to_dev(src, device)], [dst], s_src, s_dst, kern=kern)
    within(f"sl_advect u k{kern}", rel_l2_fields(mac_to_np(dst), O.sl_advect_u(vel, src, dx, dt, order, kern)), TOL)
    for s, d in zip(ss, s_dst):
        within(f"sl_advect s k{kern}", rel_l2(d.cpu().numpy(), O.sl_advect_s(vel, s, dx, dt, order, kern)), TOL)


@pytest.mark.parametrize("kern", [1, 3])
@pytest.mark.parametrize("dim,n,R,steps", [(2, (32, 32), 4, 8), (3, (16, 16, 16), 3, 6)])
def test_trajectory_matches_oracle(dim, n, R, steps, kern, device):
    """Forward and adjoint of LFM_Diff_Simulator(interp_kern=kern) against
    OracleLFM(interp_kern=kern)."""
    from dfm import LFM_Diff_Simulator, scenes

    sc = scenes.scene_2d(n, seed=4, vortices=4) if dim == 2 else scenes.scene_3d(n, seed=4, tubes=6)
    dx = sc["dx"]
    dt = 0.5 * dx
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=200, interp_kern=kern)
    u0, s0, tgt = round32(sc["u0"]), round32(sc["density"]), round32(sc["target"])
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, dt, R, sim.order, interp_kern=kern)
    orc.init(u0, [s0])
    for _ in range(steps):
        sim.step_midpoint()
        orc.step()
    within(f"u k{kern}", rel_l2_fields(mac_to_np(sim.u), orc.u), FWD_TOL)
    within(f"s k{kern}", rel_l2(sim.passive[0].cpu().numpy(), orc.s[0]), FWD_TOL)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    while not sim.backtrack_step_midpoint()[1]:
        orc.backtrack_step()
    orc.backtrack_step()
    within(f"lam_u k{kern}", rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u), ADJ_TOL)
    within(f"lam_s k{kern}", rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0]), ADJ_TOL)


@pytest.mark.parametrize("kern", [1, 3])
@pytest.mark.parametrize("dim", [2, 3])
def test_variant_transposes(dim, kern, device):
    """<A x, y> = <x, A^T y> for the pullbacks and the advection with the
    variant kernel."""
    from dfm.ops import MacField, Ops

    n, dx = GRIDS[dim]
    rng = np.random.default_rng(300 + 10 * dim + kern)
    ops = Ops(n, dx)
    fm = fmap_to_dev(random_map(rng, n, dx), device)
    rnd = lambda shp: torch.tensor(rng.standard_normal(shp).astype(np.float32), device=device)  # noqa: E731

    def dot(a, b):
        return sum(float(torch.sum(x.double() * y.double())) for x, y in zip(a, b))

    x = MacField([rnd(O.face_shape(n, a)) for a in range(dim)])
    y = MacField([rnd(O.face_shape(n, a)) for a in range(dim)])
    out = MacField.empty(n, device)
    ops.pullback_u(x, fm, out, kern=kern)
    gx = MacField.empty(n, device, zero=True)
    ops.pullback_u_vjp(y, fm, gx, kern=kern)
    lhs, rhs = dot(out.c, y.c), dot(x.c, gx.c)
    within(f"pullback_u^T k{kern}", abs(lhs - rhs) / abs(lhs), 1e-5)
    s, ys = rnd(n), rnd(n)
    so = torch.empty_like(s)
    ops.pullback_s(s, fm.psi, so, kern=kern)
    gs = torch.zeros_like(s)
    ops.pullback_s_vjp(ys, fm.psi, gs, kern=kern)
    lhs, rhs = dot([so], [ys]), dot([s], [gs])
    within(f"pullback_s^T k{kern}", abs(lhs - rhs) / abs(lhs), 1e-5)
    vel = mac_to_dev(smooth_mac(rng, n, dx), device)
    vd, sd = MacField.empty(n, device), torch.empty_like(s)
    ops.sl_advect(vel, dx, 2, [x], [vd], [s], [sd], kern=kern)
    gv, gsl = MacField.empty(n, device, zero=True), torch.zeros_like(s)
    ops.sl_advect_vjp(vel, dx, 2, [y], [gv], [ys], [gsl], kern=kern)
    lhs, rhs = dot(vd.c + [sd], y.c + [ys]), dot(x.c + [s], gv.c + [gsl])
    within(f"sl_advect^T k{kern}", abs(lhs - rhs) / abs(lhs), 1e-5)
