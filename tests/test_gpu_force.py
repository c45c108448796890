"""Gaussian control forces on the GPU against the oracle: the force field
(FP32 cut-off test bit-for-bit: the same faces are inside), its parameter
adjoint (rel. 1e-5; deterministic), and the simulators with per-step control
forces (forward fields rel-L2 <= 1e-5, adjoint fields <= 1e-5 and per-step
force gradients <= 1e-3)."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, mac_to_dev, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu

# the reference's inverse-radius choices 1 / (k dx)^2 (2D/shape/optimize_shape_G_R.py:21-36)
RADII = (1, 2, 3, 4, 5, 6, 8, 9, 10)


def make_forces(rng, count, dim, dx, lo=0.2, hi=0.8, radii=RADII):
    c = rng.uniform(lo, hi, (count, dim)).astype(np.float32)
    r = np.array([1.0 / (rng.choice(radii) * dx) ** 2 for _ in range(count)], np.float32)
    ang = rng.uniform(0, 2 * np.pi, count)
    mag = rng.uniform(5e-3, 2e-2, count) * 2.0
    s = rng.standard_normal((count, dim)).astype(np.float32)
    s *= (mag / np.linalg.norm(s, axis=1))[:, None].astype(np.float32)
    if dim == 2:
        s[:, 0], s[:, 1] = mag * np.cos(ang), mag * np.sin(ang)
    return c, r, s


def to_dev(*arrs, device):
    return [torch.tensor(np.ascontiguousarray(a, np.float32), device=device) for a in arrs]


@pytest.mark.parametrize("n,count", [((96, 80), 300), ((24, 20, 28), 60)])
def test_force_field_matches_oracle(n, count, device):
    from dfm.ops import MacField, Ops

    dim = len(n)
    dx = 1.0 / n[1]
    rng = np.random.default_rng(21 + dim)
    c, r, s = make_forces(rng, count, dim, dx)
    # edge cases: a centre outside the domain, r = 0 (every face, weight 1),
    # a NaN centre (no face), and rows beyond the count (ignored)
    c[3] = [-0.1] + [0.5] * (dim - 1)
    r[4] = 0.0
    c[5, 0] = np.nan
    used = count - 7
    ops = Ops(n, dx)
    f = MacField.empty(n, device)
    for t in f.c:
        t.fill_(7.0)
    ops.gaussian_force(*to_dev(c, r, s, device=device), used, f)
    ref = O.gaussian_force(n, dx, c, r, s, used)
    got = mac_to_np(f)
    for a in range(dim):
        scale = np.abs(ref[a]).max()
        assert np.abs(got[a] - ref[a]).max() <= 2e-6 * scale, a
    # positive strengths: a face is non-zero exactly when some force admits it
    sp = np.abs(s) + 1e-3
    r2 = r.copy()
    r2[4] = 1.0 / (3 * dx) ** 2
    ops.gaussian_force(*to_dev(c, r2, sp, device=device), used, f)
    ref = O.gaussian_force(n, dx, c, r2, sp, used)
    for a, g in enumerate(mac_to_np(f)):
        assert np.array_equal(g > 0, ref[a] > 0), a
    # no forces: every face 0
    ops.gaussian_force(*to_dev(c, r, s, device=device), 0, f)
    assert all(float(t.abs().max()) == 0.0 for t in f.c)


@pytest.mark.parametrize("n,count", [((96, 80), 300), ((24, 20, 28), 60)])
def test_force_adjoint_matches_oracle_and_is_deterministic(n, count, device):
    from dfm.ops import Ops

    dim = len(n)
    dx = 1.0 / n[1]
    rng = np.random.default_rng(31 + dim)
    c, r, s = make_forces(rng, count, dim, dx)
    lam = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    ops = Ops(n, dx)
    dc, dr, ds = to_dev(c, r, s, device=device)
    outs = []
    for _ in range(2):
        g = [torch.full((count, dim), 3.0, device=device) for _ in range(4)]
        ops.gaussian_force_adjoint(dc, dr, ds, count, mac_to_dev(lam, device), *g)
        outs.append([t.cpu().numpy() for t in g])
    ref = O.gaussian_force_adjoint(n, dx, c, r, s, count, lam)
    for name, got, want in zip(("adj_center", "adj_strength", "pen_center", "pen_strength"), outs[0], ref):
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max(), name
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


def _force_schedule(steps, count, dim, dx, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(steps):
        c, r, s = make_forces(rng, count, dim, dx, radii=(2, 3, 4))
        s *= 40.0  # strong enough to move the flow at these short runs
        out.append({"c": c, "r": r, "s": s, "n": count})
    return out


@pytest.mark.parametrize("dim,n,R,steps,checkpoint", [(2, (32, 32), 4, 8, False), (2, (32, 32), 4, 8, True),
                                                      (3, (16, 16, 16), 3, 6, False)])
def test_simulator_with_control_forces_matches_oracle(dim, n, R, steps, checkpoint, device):
    from dfm import LFM_Diff_Simulator, scenes

    sc = scenes.scene_2d(n, seed=5, vortices=4) if dim == 2 else scenes.scene_3d(n, seed=5, tubes=6)
    dx = sc["dx"]
    dt = 0.5 * dx
    count = 12
    sched = _force_schedule(steps, count, dim, dx, seed=9)
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=200, control_num=16,
                             checkpoint=checkpoint)
    u0, s0, tgt = round32(sc["u0"]), round32(sc["density"]), round32(sc["target"])
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, dt, R, sim.order, control_forces=True)
    orc.init(u0, [s0])
    for k in range(steps):
        sim.step_midpoint(control_force_para=sched[k])
        orc.step(force=sched[k])
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), FWD_TOL)
    within("s", rel_l2(sim.passive[0].cpu().numpy(), orc.s[0]), FWD_TOL)
    # the forces changed the flow
    free = O.OracleLFM(n, dx, dt, R, sim.order)
    free.init(u0, [s0])
    for k in range(steps):
        free.step()
    assert rel_l2_fields(free.u, orc.u) > 1e-3
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    assert abs(sim.loss - orc.loss) <= 1e-5 * abs(orc.loss)
    done = False
    while not done:
        k = sim.step_num[None] - 1
        _, done = sim.backtrack_step_midpoint(control_force_para=sched[k])
        orc.backtrack_step(force=sched[k])
        names = ("adj_gaussion_force_center", "adj_gaussion_force_strength", "penalty_gaussion_force_center",
                 "penalty_gaussion_force_strength")
        for nm, want in zip(names, orc.force_grad):
            got = getattr(sim, nm).to_numpy()
            assert got.shape == (16, dim)
            assert np.all(got[count:] == 0)
            within("force_grad_" + nm, rel_l2(got[:count], want), 1e-3)
    within("lam_u", rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u), ADJ_TOL)
    within("lam_s", rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0]), ADJ_TOL)
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), ADJ_TOL)
    within("s", rel_l2(sim.passive[0].cpu().numpy(), orc.s[0]), ADJ_TOL)


def test_control_force_api(device):
    from dfm import LFM_Diff_Simulator, LFM_Simulator, scenes

    sc = scenes.scene_2d((32, 32), seed=1, vortices=3)
    dx = sc["dx"]
    plain = LFM_Simulator(32, 32, dx, 0.5 * dx, 2, None, total_steps=2)
    plain.init(theta=sc["theta"], init_passive_np=sc["density"])
    para = _force_schedule(1, 4, 2, dx, seed=2)[0]
    with pytest.raises(ValueError):
        plain.step_midpoint(control_force_para=para)
    sim = LFM_Diff_Simulator(32, 32, dx, 0.5 * dx, 2, None, total_steps=2, control_num=4)
    sim.init(theta=sc["theta"], init_passive_np=sc["density"])
    with pytest.raises(ValueError):
        sim.step_midpoint(control_force_para=dict(para, n=5))
    # apply_gausion_force fills the given fields like the reference's method
    sim.apply_gausion_force(sim.control_force_x, sim.control_force_y, para["c"], para["r"], para["s"], para["n"])
    ref = O.gaussian_force((32, 32), dx, para["c"], para["r"], para["s"], para["n"])
    within("sim.control_force_x.to_numpy", rel_l2(sim.control_force_x.to_numpy(), ref[0]), 1e-5)
    within("sim.control_force_y.to_numpy", rel_l2(sim.control_force_y.to_numpy(), ref[1]), 1e-5)
