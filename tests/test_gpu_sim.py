"""Whole-trajectory parity of the drop-in simulators against the FP64 oracle
(forward fields rel-L2 <= 1e-5 after several reinitialisation intervals,
adjoint fields / gradient rel-L2 <= 1e-5), plus size-independent properties
at the benchmark grid sizes."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu



def run_pair(dim, n, R, steps, seed):
    from dfm import LFM_Diff_Simulator, scenes

    if dim == 2:
        sc = scenes.scene_2d(n, seed=seed, vortices=4)
    else:
        sc = scenes.scene_3d(n, seed=seed, tubes=6)
    dx = sc["dx"]
    dt = 0.5 * dx
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=200)
    u0 = round32(sc["u0"])
    s0 = round32(sc["density"])
    tgt = round32(sc["target"])
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, dt, R, sim.order)
    orc.init(u0, [s0])
    return sim, orc, tgt


@pytest.mark.parametrize("dim,n,R,steps", [(2, (32, 32), 4, 8), (3, (16, 16, 16), 3, 6)])
def test_forward_and_adjoint_match_oracle(dim, n, R, steps, device):
    sim, orc, tgt = run_pair(dim, n, R, steps, seed=3)
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), FWD_TOL)
    for k in range(steps):
        _, final = sim.step_midpoint()
        orc.step()
        assert final == (k == steps - 1)
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), FWD_TOL)
    within("s", rel_l2(sim.passive[0].cpu().numpy(), orc.s[0]), FWD_TOL)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    assert abs(sim.loss - orc.loss) <= 1e-5 * abs(orc.loss)
    done = False
    while not done:
        _, done = sim.backtrack_step_midpoint()
        orc.backtrack_step()
    assert sim.step_num[None] == 0
    within("lam_u", rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u), ADJ_TOL)
    within("lam_s", rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0]), ADJ_TOL)
    # the reverse-advected primal returns to the initial state (backward_u)
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), ADJ_TOL)


@pytest.mark.parametrize("dim,n,R,steps,clamp,short", [(2, (32, 32), 4, 8, True, True), (2, (32, 32), 4, 8, True, False),
                                                     (2, (32, 32), 4, 8, False, True), (3, (16, 16, 16), 3, 6, True, True)])
def test_bfecc_variants_match_oracle(dim, n, R, steps, clamp, short, device):
    """BFECC_clamp and use_short_BFECC (the reference's plume stage-2
    configuration uses both) through whole forward and adjoint trajectories."""
    from dfm import LFM_Diff_Simulator, scenes

    sc = scenes.scene_2d(n, seed=3, vortices=4) if dim == 2 else scenes.scene_3d(n, seed=3, tubes=6)
    dx = sc["dx"]
    dt = 0.5 * dx
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=200, BFECC_clamp=clamp,
                             use_short_BFECC=short)
    u0, s0, tgt = round32(sc["u0"]), round32(sc["density"]), round32(sc["target"])
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, dt, R, sim.order, bfecc_clamp=clamp, short_bfecc=short)
    orc.init(u0, [s0])
    for _ in range(steps):
        sim.step_midpoint()
        orc.step()
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), FWD_TOL)
    within("s", rel_l2(sim.passive[0].cpu().numpy(), orc.s[0]), FWD_TOL)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    while not sim.backtrack_step_midpoint()[1]:
        orc.backtrack_step()
    orc.backtrack_step()
    within("lam_u", rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u), ADJ_TOL)
    within("lam_s", rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0]), ADJ_TOL)
    within("u", rel_l2_fields(mac_to_np(sim.u), orc.u), ADJ_TOL)


def test_reference_api_surface(device):
    from dfm import LFM_Diff_Simulator, scenes

    sc = scenes.scene_2d((32, 32), seed=1, vortices=3)
    sim = LFM_Diff_Simulator(32, 32, sc["dx"], 0.5 * sc["dx"], 2, "unused_log_dir", total_steps=2)
    sim.init(theta=sc["theta"], init_passive_np=sc["density"])
    assert sim.u_x.shape == (33, 32) and sim.u_y.shape == (32, 33)
    assert sim.step_num[None] == -1
    out, final = sim.forward_step_midpoint(True, True)
    assert (out, final) == (False, False)
    out, final = sim.forward_step_midpoint(True, True)
    assert final
    w = sim.calculate_w()
    assert w.shape == (32, 32) and np.isfinite(w).all()
    sim.init_gradient(target=sc["target"])
    assert sim.adj_u_x.to_numpy().shape == (33, 32)
    flags = [sim.backtrack_step_midpoint()[1] for _ in range(2)]
    assert flags == [False, True]
    assert np.isfinite(sim.adj_u_x.to_numpy()).all()


@pytest.mark.parametrize("n", [(256, 256), (128, 128, 128)])
def test_full_size_projection_and_uniform_march(n, device):
    """At the benchmark sizes: projected divergence small, projection
    idempotent, and a uniform flow marches the flow map exactly."""
    from dfm import MGPCG, FlowMap, MacField, Ops

    dim = len(n)
    dx = 1.0 / n[1]
    rng = np.random.default_rng(11)
    u = MacField([torch.tensor(rng.standard_normal(O.face_shape(n, a)), dtype=torch.float32, device=device)
                  for a in range(dim)])
    solver = MGPCG(n, dx, tol=1e-6)
    solver.Poisson(u)
    it, res = solver.status()
    assert res < 1e-6 and it < 60
    div = sum(torch.diff(u.c[a], dim=a) for a in range(dim))
    assert float(div.abs().max()) < 1e-4
    c = [0.3, -0.2, 0.1][:dim]
    U = MacField([torch.full(u.c[a].shape, c[a], device=device) for a in range(dim)])
    fm = FlowMap(n, device)
    ops = Ops(n, dx)
    ops.reset_identity(fm)
    psi0 = fm.psi.clone()
    ops.march(U, fm, 0.5 * dx, 4)
    shift = torch.tensor(c, device=device).view(dim, *([1] * dim)) * 0.5 * dx
    assert float((fm.psi - (psi0 - shift)).abs().max()) < 1e-6
    for a in range(dim):
        assert float((fm.T_f[a][a] - 1).abs().max()) < 1e-6


@pytest.mark.parametrize("dim,checkpoint", [(2, False), (3, False), (2, True)])
def test_overlapped_adjoint_projections_are_bitwise_sequential(dim, checkpoint, device):
    """The adjoint step's two independent projections, queued on two MGPCG
    handles and streams before either waits, and the flow-map marches (and
    source accumulation) run on a side stream beside the projections, give
    bitwise the results of everything in sequence on one stream."""
    from dfm import LFM_Diff_Simulator, scenes

    n, R, steps = ((32, 32), 4, 8) if dim == 2 else ((16, 16, 16), 3, 6)
    sc = scenes.scene_2d(n, seed=9, vortices=4) if dim == 2 else scenes.scene_3d(n, seed=9, tubes=5)
    dx = sc["dx"]
    out = []
    for overlap in (False, True):
        sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=1e-7,
                                 overlap_projections=overlap, overlap_march=overlap, checkpoint=checkpoint)
        sim.init(u0=round32(sc["u0"]), init_passive_np=round32(sc["density"]))
        for _ in range(steps):
            sim.step_midpoint()
        sim.init_gradient(target=round32(sc["target"]))
        for _ in range(steps):
            sim.backtrack_step_midpoint()
        out.append((mac_to_np(sim.adj_u), mac_to_np(sim.u), sim.adj_passive[0].cpu().numpy()))
    for a, b in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("dim,n,R,steps", [(2, (32, 32), 4, 8), (3, (16, 16, 16), 3, 6)])
def test_reference_file_layout_round_trip(dim, n, R, steps, tmp_path, device):
    """export_cache writes the run in the reference's RW_Cache layout, and an
    adjoint sweep resumed from those files (import_cache in a fresh simulator)
    is bitwise the sweep of the simulator that ran the forward pass."""
    from dfm import LFM_Diff_Simulator, scenes

    sc = scenes.scene_2d(n, seed=3, vortices=4) if dim == 2 else scenes.scene_3d(n, seed=3, tubes=6)
    dx = sc["dx"]
    kw = dict(total_steps=steps, tol=1e-7, max_iters=200)
    a = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, str(tmp_path), **kw)
    a.init(u0=round32(sc["u0"]), init_passive_np=round32(sc["density"]))
    for _ in range(steps):
        a.step_midpoint()
    a.export_cache()
    xy = "phi_xy" if dim == 2 else "phi_xyz"
    s = "s1" if dim == 2 else "s"
    big = tuple(m + 1 for m in n)
    files = {
        "midpoint/u_0_%d.npy" % (steps - 1): (100,) + big + (dim,),
        "flowmap_%s/%s_0_%d.npy" % (xy, xy, steps // R - 1): (100,) + big + (dim * dim,),
        "flowmap_F/F_0_%d.npy" % (steps // R - 1): (100,) + big + (dim * dim,),
        "flowmap_phi/phi_0_%d.npy" % (steps // R - 1): (100,) + n + (dim,),
        "passive/%s_0_%d.npy" % (s, steps - R): (100,) + n + (1,),
        "final/u_x_%d.npy" % steps: (n[0] + 1,) + n[1:],
        "final/%s_%d.npy" % (s, steps): n,
    }
    for f, shape in files.items():
        assert np.load(tmp_path / f).shape == shape, f
    assert np.load(tmp_path / "final/final_ind_0.npy").tolist() == [steps]
    b = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, str(tmp_path), **kw)
    b.import_cache()
    assert b.step_num[None] == steps - 1
    tgt = round32(sc["target"])
    for sim in (a, b):
        sim.init_gradient(target=tgt)
        while not sim.backtrack_step_midpoint()[1]:
            pass
    assert a.loss == b.loss
    for x, y in zip(a.adj_u.c + [a.adj_passive[0]], b.adj_u.c + [b.adj_passive[0]]):
        assert torch.equal(x, y)


def test_interval_targets_match_oracle(tmp_path, device):
    """Intermediate targets at interval starts (the reference's q7 terms):
    loss and adjoint against the oracle, with targets given directly and read
    from a target run's passive series in the reference's layout."""
    from dfm import LFM_Diff_Simulator, scenes

    n, R, steps = (32, 32), 4, 8
    sc = scenes.scene_2d(n, seed=4, vortices=4)
    dx = sc["dx"]
    rng = np.random.default_rng(6)
    targets = {st: [round32(np.clip(sc["density"] + 0.2 * rng.standard_normal(n), 0, 1))] for st in (0, 4)}
    sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=1e-7, max_iters=200)
    u0, s0, tgt = round32(sc["u0"]), round32(sc["density"]), round32(sc["target"])
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, 0.5 * dx, R, sim.order)
    orc.init(u0, [s0])
    orc.interval_targets = targets
    for _ in range(steps):
        sim.step_midpoint()
        orc.step()
    sim.set_interval_targets(targets)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    final_loss = orc.loss
    while not sim.backtrack_step_midpoint()[1]:
        orc.backtrack_step()
    orc.backtrack_step()
    assert orc.loss > 1.5 * final_loss  # the interval terms are on
    assert abs(sim.loss - orc.loss) <= 1e-5 * orc.loss
    within("lam_s", rel_l2(sim.adj_passive[0].cpu().numpy(), orc.lam_s[0]), ADJ_TOL)
    within("lam_u", rel_l2_fields(mac_to_np(sim.adj_u), orc.lam_u), ADJ_TOL)
    # the same targets as a target run's passive series on disk
    from dfm import rw_cache as W

    (tmp_path / "passive").mkdir()
    cache = W.RWCache()
    for st in range(5):
        t = targets[st][0] if st in targets else np.zeros(n)
        cache.write(W.as_channels(t, 2), "s1", str(tmp_path / "passive"), st)
    cache.force_write("s1", str(tmp_path / "passive"))
    b = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, str(tmp_path), total_steps=steps, tol=1e-7, max_iters=200)
    b.init(u0=u0, init_passive_np=s0)
    for _ in range(steps):
        b.step_midpoint()
    b.set_interval_targets(b.interval_targets_from_dir())
    b.init_gradient(target=tgt)
    while not b.backtrack_step_midpoint()[1]:
        pass
    assert b.loss == sim.loss
    assert torch.equal(b.adj_passive[0], sim.adj_passive[0])


@pytest.mark.parametrize("checkpoint", [False, True])
def test_reserved_storage_reused_bitwise(checkpoint, device):
    """reserve() allocates the per-step storage up front and init() returns
    it to the simulator's pool: a second run of the same problem, and a
    longer one within the reservation, allocate no device memory and give
    bitwise the first run's fields and gradient."""
    from dfm import LFM_Diff_Simulator, scenes

    n, R = (24, 20, 16), 3
    sc = scenes.scene_3d(n, seed=4, tubes=4)
    dx = sc["dx"]
    sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=2 * R, tol=1e-7, max_iters=200,
                             checkpoint=checkpoint)
    sim.reserve(4 * R)

    def run(steps):
        sim.total_steps = steps
        sim.init(u0=sc["u0"], init_passive_np=sc["density"])
        for _ in range(steps):
            sim.step_midpoint()
        sim.init_gradient(target=sc["target"])
        for _ in range(steps):
            sim.backtrack_step_midpoint()
        torch.cuda.synchronize()
        # results kept on the host, so device memory holds only the simulator's
        return [c.cpu() for c in sim.adj_u.c] + [sim.adj_passive[0].cpu()], sim.loss

    first, l1 = run(2 * R)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    second, l2 = run(2 * R)
    longer, _ = run(4 * R)
    torch.cuda.synchronize()
    assert torch.cuda.memory_allocated() == before
    assert l1 == l2
    for a, b in zip(first, second):
        assert torch.equal(a, b)
    assert all(bool(torch.isfinite(t).all()) for t in longer)


def test_nonfinite_fields_raise(device):
    """NaN / Inf reach the caller as DfmError(DFM_E_NONFINITE): counted
    exactly by dfm_count_nonfinite; in the final forward state at
    init_gradient; in the velocity at the next projection; through a NaN
    target in the adjoint."""
    from dfm import LFM_Diff_Simulator, _lib as L, scenes
    from dfm.ops import Ops

    n, R = (20, 16, 12), 2
    ops = Ops(n, 1.0 / 16)
    a = torch.zeros(100, device=device)
    b = torch.zeros(7, device=device)
    a[[3, 50, 99]] = float("nan")
    b[0] = float("inf")
    b[6] = float("-inf")
    cnt = torch.zeros(1, dtype=torch.int32, device=device)
    ops.count_nonfinite([a, b], cnt)
    assert int(cnt.item()) == 5
    ops.count_nonfinite([], cnt)
    assert int(cnt.item()) == 0

    sc = scenes.scene_3d(n, seed=1, tubes=3)

    def sim_after(steps):
        sim = LFM_Diff_Simulator(*n, sc["dx"], 0.5 * sc["dx"], R, None, total_steps=steps, tol=1e-7, max_iters=200)
        sim.init(u0=sc["u0"], init_passive_np=sc["density"])
        for _ in range(steps):
            sim.step_midpoint()
        return sim

    sim = sim_after(R)
    sim.passive[0][1, 2, 3] = float("nan")
    with pytest.raises(L.DfmError) as e:
        sim.init_gradient(target=sc["target"])
    assert e.value.code == L.DFM_E_NONFINITE

    sim = sim_after(1)
    sim.u.c[0][4, 4, 4] = float("inf")
    with pytest.raises(L.DfmError) as e:
        sim.step_midpoint()
    assert e.value.code == L.DFM_E_NONFINITE

    sim = sim_after(R)
    tgt = np.array(sc["target"], dtype=np.float64)
    tgt[2, 2, 2] = np.nan
    with pytest.raises(L.DfmError) as e:
        sim.init_gradient(target=tgt)
        for _ in range(R):
            sim.backtrack_step_midpoint()
    assert e.value.code == L.DFM_E_NONFINITE
