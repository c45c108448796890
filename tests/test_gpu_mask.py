"""Obstacles on the GPU: the mask_* field kernels, the obstacle-aware MGPCG
projection (MGPCG_2_solid) against the oracle's exact masked projection, and
a whole 2D forward+adjoint trajectory with a cylinder obstacle against the
FP64 oracle."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, mac_to_dev, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def ball_mask(n, centre, radius):
    X = O.cell_coords(n, 1.0 / n[1])
    return (np.sqrt(np.sum((X - np.asarray(centre)) ** 2, axis=-1)) < radius).astype(np.float32)


@pytest.mark.parametrize("n", [(40, 36), (18, 16, 14)])
def test_mask_kernels_match_oracle(n, device):
    from dfm.ops import Ops

    rng = np.random.default_rng(21)
    dim = len(n)
    mask = ball_mask(n, [0.5] * dim, 0.3)
    mask[0, 0] = 0.5                       # neither solid nor fluid
    mask[-1, -1] = 1.0                     # solid corner cell
    ops = Ops(n, 1.0 / n[1])
    m = torch.tensor(mask, device=device)
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    bvel = round32([rng.standard_normal(n) for _ in range(dim)])
    d = mac_to_dev(u, device)
    ops.mask_velocity(m, d, torch.tensor(np.stack(bvel), dtype=torch.float32, device=device))
    ref = [c.copy() for c in u]
    O.mask_velocity(ref, mask, bvel)
    for a, b in zip(mac_to_np(d), ref):
        assert np.array_equal(a, b)
    d = mac_to_dev(u, device)
    ops.mask_velocity(m, d, None)
    ref = [c.copy() for c in u]
    O.mask_velocity(ref, mask, None)
    for a, b in zip(mac_to_np(d), ref):
        assert np.array_equal(a, b)
    s = rng.standard_normal(n).astype(np.float32)
    for adjoint in (False, True):
        t = torch.tensor(s, device=device)
        ops.mask_scalar(m, t, adjoint=adjoint)
        r = s.copy()
        O.mask_scalar(r, mask, adjoint=adjoint)
        assert np.abs(t.cpu().numpy() - r).max() <= 1e-6 * np.abs(r).max()


@pytest.mark.parametrize("n,levels", [((64, 48), 3), ((32, 32, 24), 3)])
def test_masked_projection_matches_exact(n, levels, device):
    from dfm.ops import MGPCG

    dim = len(n)
    rng = np.random.default_rng(8 + dim)
    mask = ball_mask(n, [0.45] + [0.5] * (dim - 1), 0.2)
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    solver = MGPCG(n, 1.0 / n[1], levels=levels, tol=1e-7, max_iters=300)
    solver.set_mask(torch.tensor(mask, device=device))
    d = mac_to_dev(u, device)
    solver.Poisson(d)
    it, res = solver.status()
    ref = [c.copy() for c in u]
    O.project_masked_exact(ref, mask > 0)
    out = mac_to_np(d)
    err = rel_l2_fields(out, ref)
    print(f"n={n} masked: iters={it} res={res:.2e} rel-L2 vs exact={err:.2e}")
    assert err < 5e-6
    for a in range(dim):
        assert np.all(out[a][O._face_solid(mask > 0, a)] == 0)
    div = O.divergence(out)
    assert np.abs(div[mask == 0]).max() < 1e-4 * max(np.abs(c).max() for c in u)
    # removing the mask restores the plain projection
    solver.set_mask(None)
    d = mac_to_dev(u, device)
    solver.Poisson(d)
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    within("mac_to_np_d", rel_l2_fields(mac_to_np(d), ref), 2e-6)


def test_simulator_with_obstacle_matches_oracle(device):
    """2D, 32^2, R = 4, 8 steps, a cylinder obstacle: forward fields and the
    adjoint against the FP64 oracle with the same mask call sites."""
    from dfm import LFM_Diff_Simulator, scenes

    n, R, steps = (32, 32), 4, 8
    sc = scenes.scene_2d(n, seed=6, vortices=4)
    dx = sc["dx"]
    dt = 0.5 * dx
    mask = ball_mask(n, [0.3, 0.6], 0.12)
    u0, s0, tgt = round32(sc["u0"]), round32(sc["density"]), round32(sc["target"])
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=300, mg_levels=3)
    sim.set_boundary_mask(mask)
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, dt, R, 2)
    orc.set_mask(mask)
    orc.init(u0, [s0])
    for _ in range(steps):
        sim.step_midpoint()
        orc.step()
    eu = rel_l2_fields(mac_to_np(sim.u), orc.u)
    es = rel_l2(sim.passive[0].cpu().numpy(), orc.s[0])
    print(f"forward rel-L2 u={eu:.2e} s={es:.2e}")
    within("u", eu, FWD_TOL)
    within("s", es, FWD_TOL)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    for _ in range(steps):
        sim.backtrack_step_midpoint()
        orc.backtrack_step()
    ea = rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u)
    eas = rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0])
    print(f"adjoint rel-L2 lam_u={ea:.2e} lam_s={eas:.2e}")
    within("lam_u", ea, ADJ_TOL)
    within("lam_s", eas, ADJ_TOL)
    # the obstacle is respected: no flow through solid-cell faces
    for a in range(2):
        assert np.all(mac_to_np(sim.u)[a][O._face_solid(mask >= 1, a)] == 0)
