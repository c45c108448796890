"""GPU path against the committed golden fixtures (FP64 oracle outputs on
seeded FP32 inputs, tests/golden/): forward fields rel-L2 <= 1e-5, adjoint
fields rel-L2 <= 1e-5, projection <= 2e-6, vortex initializer <= 1e-6 and its
VJP <= 1e-4."""
import os

import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, mac_to_dev, mac_to_np, rel_l2, rel_l2_fields, within

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["lfm2d_32", "lfm2d_48x40", "lfm3d_16"])
def test_simulator_matches_golden(name, device):
    from dfm import LFM_Diff_Simulator

    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    dim, n = int(g["dim"]), tuple(int(v) for v in g["n"])
    steps = int(g["steps"])
    sim = LFM_Diff_Simulator(*n, float(g["dx"]), float(g["dt"]), int(g["R"]), None, total_steps=steps,
                             order=int(g["order"]), tol=1e-7, max_iters=200)
    sim.init(u0=[g["u0_%d" % a] for a in range(dim)], init_passive_np=g["s0"])
    for _ in range(steps):
        sim.step_midpoint()
    within("u", rel_l2_fields(mac_to_np(sim.u), [g["u_T_%d" % a] for a in range(dim)]), FWD_TOL)
    within("s", rel_l2(sim.passive[0].cpu().numpy(), g["s_T"]), FWD_TOL)
    sim.init_gradient(target=g["target"])
    assert abs(sim.loss - float(g["loss"])) <= 1e-5 * float(g["loss"])
    for _ in range(steps):
        sim.backtrack_step_midpoint()
    within("lam_u", rel_l2_fields(mac_to_np(sim.adj_u), [g["lam_u0_%d" % a] for a in range(dim)]), ADJ_TOL)
    within("lam_s", rel_l2(sim.adj_passive[0].cpu().numpy(), g["lam_s0"]), ADJ_TOL)


def test_projection_matches_golden(device):
    from dfm.ops import MGPCG

    g = np.load(os.path.join(GOLDEN, "projection3d_24x20x16.npz"))
    n = tuple(int(v) for v in g["n"])
    s = MGPCG(n, 1.0 / n[1], levels=3, tol=1e-7, max_iters=100)
    d = mac_to_dev([g["u_%d" % a] for a in range(3)], device)
    s.Poisson(d)
    within("mac_to_np_d", rel_l2_fields(mac_to_np(d), [g["Pu_%d" % a] for a in range(3)]), 2e-6)


def test_vortex_initializer_matches_golden(device):
    from dfm.autograd import vortex_velocity_2d

    g = np.load(os.path.join(GOLDEN, "vortex2d_40x36.npz"))
    n, dx = tuple(int(v) for v in g["n"]), float(g["dx"])
    th = torch.tensor(g["theta"], device=device, requires_grad=True)
    ux, uy = vortex_velocity_2d(th, n, dx)
    within("[ux.detach_.cpu_.numpy", rel_l2_fields([ux.detach().cpu().numpy(), uy.detach().cpu().numpy()], [g["u_0"], g["u_1"]]), 1e-6)
    lam = [torch.tensor(g["lam_%d" % a], device=device) for a in range(2)]
    ((ux * lam[0]).sum() + (uy * lam[1]).sum()).backward()
    within("th.grad.cpu_.numpy", rel_l2(th.grad.cpu().numpy(), g["vjp"]), 1e-4)
