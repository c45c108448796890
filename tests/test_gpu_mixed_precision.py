"""The mixed-precision form (BF16 storage of MG levels >= 2, FP32 V-cycle
arithmetic, FP64 outer PCG with iterative refinement) on the GPU:
 - the projection with BF16 global levels converges to the exact (DCT)
   projection as the FP32 form does (it only changes the preconditioner),
   at sizes where level 2 is a global LDS-tiled level and where it is the
   tail's first level;
 - a whole forward + adjoint trajectory in this mode against the FP64
   oracle, at the same bounds as the FP32 trajectories;
 - at 256^3 (the mixed-precision configuration's grid): BF16 reported in use
   on level 2, projected divergence at the FP32 floor, a forward + adjoint
   interval that stays finite and within 1e-5 of the FP32 form's fields."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, mac_to_dev, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,levels,tail_level,bf16_level", [
    ((64, 64, 32), 4, 3, 2),      # level 2 (16 x 16 x 8) tiled in BF16, level 3 in the tail
    ((128, 128, 128), 5, 0, 2),   # levels 2.. (32^3 down) in the BF16 tail
    ((1024, 1024), 0, 0, 2),      # 2D: level 2 (256^2) tiled in BF16, the tail from 128^2
    ((32, 32, 32), 0, 0, 3),      # 3 levels, all in one FP32 tail: no BF16 level
])
def test_bf16_levels_project_like_fp32(n, levels, tail_level, bf16_level, device):
    from dfm import _lib as L
    from dfm.ops import MGPCG

    dx = 1.0 / n[1]
    rng = np.random.default_rng(3 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))])
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    out = {}
    for bf16 in (False, True):
        s = MGPCG(n, dx, levels=levels, tol=1e-7, max_iters=100, bf16_coarse=bf16, tail_level=tail_level,
                  bottom_exact=True)
        d = mac_to_dev(u, device)
        s.Poisson(d)
        # FP32: no level stored in BF16 (the level count L)
        assert s.get_option(L.MG_BF16_LEVEL) == (bf16_level if bf16 else s.get_option(L.MG_LEVELS))
        it, res = s.status()
        out[bf16] = (mac_to_np(d), it)
        print(f"n={n} bf16={bf16}: iters={it} res={res:.2e} bf16_level={s.get_option(L.MG_BF16_LEVEL)} "
              f"tail={s.get_option(L.MG_TAIL)}", flush=True)
        within("projection_vs_exact", rel_l2_fields(out[bf16][0], ref), 2e-6)
    assert out[True][1] <= out[False][1] + 2


def test_bf16_trajectory_matches_oracle(device):
    """3D forward + adjoint with BF16 global level 2 (64 x 64 x 32, 4 levels,
    tail from level 3) against the FP64 oracle (exact projection)."""
    from dfm import LFM_Diff_Simulator, _lib as L, scenes

    n, R, steps = (64, 64, 32), 2, 4
    sc = scenes.scene_3d(n, seed=5, tubes=5)
    dx = sc["dx"]
    dt = 0.5 * dx
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=200, mg_levels=4,
                             mg_bf16_coarse=True, mg_tail_level=3, mg_bottom_exact=True)
    u0, s0, tgt = round32(sc["u0"]), round32(sc["density"]), round32(sc["target"])
    sim.init(u0=u0, init_passive_np=s0)
    assert sim.solver.get_option(L.MG_BF16_LEVEL) == 2
    orc = O.OracleLFM(n, dx, dt, R, sim.order)
    orc.init(u0, [s0])
    for _ in range(steps):
        sim.step_midpoint()
        orc.step()
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), FWD_TOL)
    within("s", rel_l2(sim.passive[0].cpu().numpy(), orc.s[0]), FWD_TOL)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    assert abs(sim.loss - orc.loss) <= 1e-5 * orc.loss
    while not sim.backtrack_step_midpoint()[1]:
        orc.backtrack_step()
    orc.backtrack_step()
    within("lam_u", rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u), ADJ_TOL)
    within("lam_s", rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0]), ADJ_TOL)
    assert sim.noconv_count == 0


def _interval_256(bf16, sc):
    from dfm import LFM_Diff_Simulator

    n = (256, 256, 256)
    dx = sc["dx"]
    sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, 5, None, total_steps=5, tol=1e-7, max_iters=200,
                             mg_bf16_coarse=bf16, mg_bottom_exact=True)
    sim.init(u0=sc["u0"], init_passive_np=sc["density"])
    for _ in range(5):
        sim.step_midpoint()
    u_T = [c.clone() for c in sim.u.c]
    s_T = sim.passive[0].clone()
    sim.init_gradient(target=sc["target"])
    for _ in range(5):
        sim.backtrack_step_midpoint()
    torch.cuda.synchronize()
    return sim, u_T, s_T


def test_bf16_full_size_256(device):
    from dfm import _lib as L, scenes

    n = (256, 256, 256)
    sc = scenes.scene_3d(n, seed=2, tubes=4)
    sc = {k: (round32(v) if k in ("u0", "density", "target") else v) for k, v in sc.items()}
    fp32, u32, s32 = _interval_256(False, sc)
    bf16, u16, s16 = _interval_256(True, sc)
    assert bf16.solver.get_option(L.MG_BF16_LEVEL) == 2 and bf16.solver.get_option(L.MG_BF16_COARSE) == 1
    assert fp32.solver.get_option(L.MG_BF16_LEVEL) == fp32.solver.get_option(L.MG_LEVELS) == 6
    assert bf16.noconv_count == 0 and fp32.noconv_count == 0
    lam = torch.cat([c.flatten() for c in bf16.adj_u.c] + [bf16.adj_passive[0].flatten()])
    assert bool(torch.isfinite(lam).all())
    div = sum(torch.diff(u16[a].double(), dim=a) for a in range(3))
    assert float(div.abs().max()) < 1e-4 * max(float(c.abs().max()) for c in u16)
    t = lambda xs: [x.cpu().numpy() for x in xs]  # noqa: E731
    within("u_T_bf16_vs_fp32", rel_l2_fields(t(u16), t(u32)), FWD_TOL)
    within("s_T_bf16_vs_fp32", rel_l2(s16.cpu().numpy(), s32.cpu().numpy()), FWD_TOL)
    within("lam_u_bf16_vs_fp32", rel_l2_fields(t(bf16.adj_u.c), t(fp32.adj_u.c)), ADJ_TOL)
    within("lam_s_bf16_vs_fp32", rel_l2(bf16.adj_passive[0].cpu().numpy(), fp32.adj_passive[0].cpu().numpy()), ADJ_TOL)
