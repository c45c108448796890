"""Differentiation on the GPU: the vortex initializer and its vector-Jacobian
product kernel, per-interval checkpoint/recompute of the adjoint sweep, the
torch.autograd wrappers, and the whole-loss gradient against central finite
differences (h = 1e-3, 4 seeded directions) on the 2D vortex control problem.

The gradient being checked is the reference's continuous adjoint, which is
not the exact derivative of the discrete forward map: on this problem the
FP64 oracle's adjoint gradient has whole-vector cosine ~0.999 and rel-L2
~5e-2 against central differences (measured with the oracle on the same
setup), so the finite-difference bounds below are set at that gap, while the
GPU-vs-oracle gradient bound is the tight one (rel-L2 <= 1e-5)."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def host_faces(theta, n, dx):
    from dfm import scenes

    return scenes.split_central_vector(scenes.vortex_velocity_2d(theta, n, dx))


def test_vortex_init_matches_host_restatement(device):
    from dfm import scenes
    from dfm.autograd import vortex_velocity_2d

    n, dx = (64, 48), 1.0 / 48
    theta = scenes.vortex_params_2d(11, 8).astype(np.float32)
    ux, uy = vortex_velocity_2d(torch.tensor(theta, device=device), n, dx)
    ref = host_faces(theta.astype(np.float64), n, dx)
    assert ux.shape == (65, 48) and uy.shape == (64, 49)
    within("[ux.cpu_.numpy", rel_l2_fields([ux.cpu().numpy(), uy.cpu().numpy()], ref), 1e-6)


def test_vortex_single_closed_form(device):
    """One vortex: the cell-centred field is the swirl w g(r) 2.5 (-p_y, p_x),
    g(r) = -0.01 (1 - exp(-r^2/a^2)) / r^2; at faces the split averages two
    cells.  Checked at an interior x-face against the closed form."""
    from dfm.autograd import vortex_velocity_2d

    n, dx = (32, 32), 1.0 / 32
    w, cx, cy, rr = 1.3, 0.47, 0.52, 0.8
    th = torch.tensor([w, cx, cy, rr], dtype=torch.float32, device=device)
    ux, _ = vortex_velocity_2d(th, n, dx)
    i, j = 20, 9
    a = 0.05 * np.float64(np.float32(rr))
    w, cx, cy = (np.float64(np.float32(v)) for v in (w, cx, cy))

    def cell_ux(ci, cj):
        px, py = (ci + 0.5) * dx - cx, (cj + 0.5) * dx - cy
        r = np.hypot(px, py) + 1e-6
        return w * -0.01 * (1 - np.exp(-r * r / (a * a))) / (r * r) * 2.5 * -py

    expect = 0.5 * (cell_ux(i, j) + cell_ux(i - 1, j))
    assert abs(float(ux[i, j]) - expect) <= 1e-6 * abs(expect) + 1e-9


def test_vortex_vjp_matches_finite_differences(device):
    """<lam, du/dtheta_i> from the VJP kernel against central differences of
    the FP64 host initializer (the reference forms the same inner product by
    finite differences, 2D/vortex/optimize_vortex.py:193-226)."""
    from dfm import scenes
    from dfm.autograd import vortex_velocity_2d

    n, dx = (48, 40), 1.0 / 40
    theta = scenes.vortex_params_2d(5, 6).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(9)
    lam = [rng.standard_normal((49, 40)), rng.standard_normal((48, 41))]
    th = torch.tensor(theta, dtype=torch.float32, device=device, requires_grad=True)
    ux, uy = vortex_velocity_2d(th, n, dx)
    lt = [torch.tensor(l, dtype=torch.float32, device=device) for l in lam]
    ((ux * lt[0]).sum() + (uy * lt[1]).sum()).backward()
    g = th.grad.double().cpu().numpy()
    lam32 = [l.astype(np.float32).astype(np.float64) for l in lam]
    eps = 1e-6
    fd = np.zeros_like(theta)
    for i in range(theta.size):
        e = np.zeros_like(theta)
        e[i] = eps
        up, um = host_faces(theta + e, n, dx), host_faces(theta - e, n, dx)
        fd[i] = sum(np.sum(l * (a - b)) for l, a, b in zip(lam32, up, um)) / (2 * eps)
    assert rel_l2(g, fd) < 1e-4, (g, fd)


def make_2d_problem(n, vortices, seed, R, steps, device, tol=1e-7):
    from dfm import LFM_Diff_Simulator, scenes

    dx = 1.0 / n[1]
    theta = scenes.vortex_params_2d(seed, vortices).astype(np.float32)
    s0 = round32(scenes.disk_density(n, dx))
    tgt = round32(scenes.gaussian_blobs(n, dx, seed + 1000, 3))
    sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=tol, max_iters=200)
    return sim, theta, s0, tgt, dx


def test_checkpoint_recompute_is_bitwise_and_bounded(device):
    """Per-interval checkpointing (replay of an interval's forward steps when
    the adjoint sweep enters it) gives the stored-everything gradient bitwise,
    holding at most one interval of midpoint velocities and one forward map."""
    from dfm import LFM_Diff_Simulator

    n, R, steps = (32, 32), 4, 12
    _, theta, s0, tgt, dx = make_2d_problem(n, 4, 3, R, steps, device)
    out = {}
    for mode in ("record", "checkpoint"):
        sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=1e-7,
                                 checkpoint=(mode == "checkpoint"))
        sim.init(theta=theta, init_passive_np=s0)
        for _ in range(steps):
            sim.step_midpoint()
        if mode == "checkpoint":
            assert len(sim.mid_u) == 0 and len(sim.flow_maps) == 0 and len(sim.u_start) == steps // R
        sim.init_gradient(target=tgt)
        done = False
        while not done:
            _, done = sim.backtrack_step_midpoint()
            if mode == "checkpoint":
                assert len(sim.mid_u) <= R and len(sim.flow_maps) <= 1
        out[mode] = (mac_to_np(sim.adj_u), sim.adj_passive[0].cpu().numpy(), sim.loss)
    for a, b in zip(out["record"][0], out["checkpoint"][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out["record"][1], out["checkpoint"][1])
    assert out["record"][2] == out["checkpoint"][2]


def test_rollout_gradient_equals_simulator_adjoint(device):
    """autograd through LFMRollout (seeded with dL/ds_T of the density loss)
    reproduces the explicit init_gradient / backtrack sweep."""
    from dfm import LFM_Diff_Simulator
    from dfm.autograd import density_loss

    n, R, steps = (32, 32), 4, 8
    sim, theta, s0, tgt, dx = make_2d_problem(n, 4, 4, R, steps, device)
    sim.init(theta=theta, init_passive_np=s0)
    u0 = [c.clone() for c in sim.u.c]
    for _ in range(steps):
        sim.step_midpoint()
    sim.init_gradient(target=tgt)
    for _ in range(steps):
        sim.backtrack_step_midpoint()
    ref_u, ref_s = mac_to_np(sim.adj_u), sim.adj_passive[0].cpu().numpy()

    sim2 = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=1e-7)
    u0g = [c.clone().requires_grad_(True) for c in u0]
    s0g = torch.tensor(s0, dtype=torch.float32, device=device, requires_grad=True)
    loss = density_loss(sim2, steps, s0g, u0g, torch.tensor(tgt, dtype=torch.float32, device=device))
    loss.backward()
    assert abs(float(loss.detach()) - sim.loss) <= 1e-6 * sim.loss
    within("[g.grad.cpu_.numpy_for_g_in_u0g]", rel_l2_fields([g.grad.cpu().numpy() for g in u0g], ref_u), 1e-6)
    within("s0g.grad.cpu_.numpy", rel_l2(s0g.grad.cpu().numpy(), ref_s), 1e-6)


def test_u0_gradient_3d_matches_oracle(device):
    """The 3D form (gradient w.r.t. the initial velocity field) through
    autograd at 16^3 against the FP64 oracle's adjoint."""
    from dfm import LFM_Diff_Simulator, scenes
    from dfm.autograd import density_loss

    n, R, steps = (16, 16, 16), 3, 6
    sc = scenes.scene_3d(n, seed=2, tubes=5)
    dx = sc["dx"]
    u0 = round32(sc["u0"])
    s0, tgt = round32(sc["density"]), round32(sc["target"])
    sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=1e-7)
    u0g = [torch.tensor(c, dtype=torch.float32, device=device, requires_grad=True) for c in u0]
    s0t = torch.tensor(s0, dtype=torch.float32, device=device)
    loss = density_loss(sim, steps, s0t, u0g, torch.tensor(tgt, dtype=torch.float32, device=device))
    loss.backward()
    orc = O.OracleLFM(n, dx, 0.5 * dx, R, 4)
    orc.init(u0, [s0])
    for _ in range(steps):
        orc.step()
    orc.init_gradient([tgt])
    for _ in range(steps):
        orc.backtrack_step()
    assert abs(float(loss) - orc.loss) <= 1e-5 * orc.loss
    within("[g.grad.cpu_.numpy_for_g_in_u0g]", rel_l2_fields([g.grad.cpu().numpy() for g in u0g], orc.lam_u), ADJ_TOL)


def test_theta_gradient_vs_central_differences_and_oracle(device):
    """The 2D vortex control (theta, 8 vortices x 4) at 64^2: the autograd
    gradient (adjoint sweep + VJP kernel) against (a) the FP64 oracle's
    adjoint pulled back through the host initializer (rel-L2 <= 1e-3) and
    (b) central differences of the GPU loss, h = 1e-3, along 4 seeded unit
    directions and component-wise (bounded by the continuous-adjoint gap)."""
    from dfm.autograd import density_loss, vortex_velocity_2d

    n, R, steps = (64, 64), 4, 8
    sim, theta, s0, tgt, dx = make_2d_problem(n, 8, 0, R, steps, device)
    s0t = torch.tensor(s0, dtype=torch.float32, device=device)
    tgtt = torch.tensor(tgt, dtype=torch.float32, device=device)

    def loss_of(th):
        return density_loss(sim, steps, s0t, list(vortex_velocity_2d(th, n, dx)), tgtt)

    th = torch.tensor(theta, device=device, requires_grad=True)
    L0 = loss_of(th)
    L0.backward()
    g = th.grad.double().cpu().numpy()
    gn = np.linalg.norm(g)

    # (a) oracle: FP64 forward/adjoint from the same (FP32-rounded) initial field
    th64 = theta.astype(np.float64)
    orc = O.OracleLFM(n, dx, 0.5 * dx, R, 2)
    orc.init(round32(host_faces(th64, n, dx)), [s0])
    for _ in range(steps):
        orc.step()
    orc.init_gradient([tgt])
    for _ in range(steps):
        orc.backtrack_step()
    eps = 1e-6
    g_orc = np.zeros_like(th64)
    for i in range(th64.size):
        e = np.zeros_like(th64)
        e[i] = eps
        up, um = host_faces(th64 + e, n, dx), host_faces(th64 - e, n, dx)
        g_orc[i] = sum(np.sum(l * (a - b)) for l, a, b in zip(orc.lam_u, up, um)) / (2 * eps)
    within("g", rel_l2(g, g_orc), ADJ_TOL)

    # (b) central differences of the GPU loss
    h = 1e-3
    with torch.no_grad():
        def L(t):
            return float(loss_of(torch.tensor(t.astype(np.float32), device=device)))

        rng = np.random.default_rng(1)
        for k in range(4):
            d = rng.standard_normal(theta.size)
            d /= np.linalg.norm(d)
            fd = (L(th64 + h * d) - L(th64 - h * d)) / (2 * h)
            print(f"direction {k}: central FD {fd:+.5e}  adjoint {g @ d:+.5e}")
            assert abs(fd - g @ d) <= 0.05 * gn
        gfd = np.zeros_like(th64)
        for i in range(th64.size):
            e = np.zeros_like(th64)
            e[i] = h
            gfd[i] = (L(th64 + e) - L(th64 - e)) / (2 * h)
    cos = float(g @ gfd / (gn * np.linalg.norm(gfd)))
    print(f"component-wise: cosine {cos:.5f}, rel-L2 {rel_l2(g, gfd):.3e}")
    assert cos > 0.99
    within("g", rel_l2(g, gfd), 0.15)
