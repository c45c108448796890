"""GPU kernel parity: each libdfm.so entry point (through the C-ABI) against
the FP64 oracle on the same float32-rounded inputs.

Tolerance: rel-L2 <= 1e-5 for FP32 kernels (the forward-field bar), on
non-square grids so an axis mix-up cannot cancel."""
import numpy as np
import pytest
import torch

from gpu_util import fmap_to_dev, fmap_to_np, mac_to_dev, mac_to_np, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
GRIDS = {2: ((24, 32), 1 / 32), 3: ((12, 16, 20), 1 / 20)}


def rand_mac(rng, n, scale=1.0):
    return round32([scale * rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))])


def smooth_mac(rng, n, dx):
    """Smooth-ish divergence-free-ish velocity of magnitude ~1."""
    dim = len(n)
    out = []
    for a in range(dim):
        X = O.face_coords(n, a, dx)
        v = np.zeros(X.shape[:-1])
        for _ in range(3):
            k = rng.uniform(1, 4, dim) * np.pi
            ph = rng.uniform(0, 2 * np.pi, dim)
            v += np.prod(np.cos(k * X + ph), axis=-1) * rng.uniform(0.3, 0.8)
        out.append(v)
    return round32(out)


def random_map(rng, n, dx, disp=2.0):
    om = O.FlowMap(n, dx, np.float64)
    om.psi = round32(om.psi + disp * dx * rng.uniform(-1, 1, om.psi.shape))
    for a in range(len(n)):
        om.psi_f[a] = round32(om.psi_f[a] + disp * dx * rng.uniform(-1, 1, om.psi_f[a].shape))
        om.T_f[a] = round32(om.T_f[a] + 0.2 * rng.standard_normal(om.T_f[a].shape))
    return om


def ops_for(n, dx):
    from dfm.ops import Ops

    return Ops(n, dx)


@pytest.mark.parametrize("dim", [2, 3])
def test_reset_identity(dim, device):
    n, dx = GRIDS[dim]
    from dfm.ops import FlowMap

    fm = FlowMap(n, device)
    ops_for(n, dx).reset_identity(fm)
    om = O.FlowMap(n, dx, np.float64)
    psi, psi_f, T_f = fmap_to_np(fm)
    within("psi", rel_l2(psi, om.psi), 1e-7)
    for a in range(dim):
        within("psi_f[a]", rel_l2(psi_f[a], om.psi_f[a]), 1e-7)
        assert np.array_equal(T_f[a], om.T_f[a])


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 4), (3, 2), (3, 4)])
def test_march(dim, order, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(10 + dim * 3 + order)
    u = smooth_mac(rng, n, dx)
    om = random_map(rng, n, dx)
    fm = fmap_to_dev(om, device)
    dt = 0.5 * dx
    ops_for(n, dx).march(mac_to_dev(u, device), fm, dt, order)
    O.march(u, om.psi, None, dx, dt, order)
    for a in range(dim):
        O.march(u, om.psi_f[a], om.T_f[a], dx, dt, order)
    psi, psi_f, T_f = fmap_to_np(fm)
    within("psi", rel_l2(psi, om.psi), TOL)
    for a in range(dim):
        within("psi_f[a]", rel_l2(psi_f[a], om.psi_f[a]), TOL)
        within("T_f[a]", rel_l2(T_f[a], om.T_f[a]), TOL)


@pytest.mark.parametrize("dim,order", [(2, 2), (3, 4)])
def test_march_capped_is_march(dim, order, device):
    """dfm_march_capped (at most k march workgroups per CU, through an LDS
    request the kernel does not use) gives dfm_march's maps bit for bit, for
    caps that fit in the default per-workgroup LDS (2, 3, 16) and one that
    needs the opt-in (1); caps outside 0..16 are refused."""
    from dfm import _lib as L

    n, dx = GRIDS[dim]
    rng = np.random.default_rng(70 + dim)
    u = smooth_mac(rng, n, dx)
    om = random_map(rng, n, dx)
    dt = 0.5 * dx
    ops = ops_for(n, dx)
    ref = fmap_to_dev(om, device)
    ops.march(mac_to_dev(u, device), ref, dt, order)
    want = fmap_to_np(ref)
    for cap in (1, 2, 3, 16):
        fm = fmap_to_dev(om, device)
        ops.march(mac_to_dev(u, device), fm, dt, order, blocks_per_cu=cap)
        got = fmap_to_np(fm)
        assert np.array_equal(got[0], want[0])
        for a in range(dim):
            assert np.array_equal(got[1][a], want[1][a]) and np.array_equal(got[2][a], want[2][a])
    for bad in (-1, 17):
        with pytest.raises(L.DfmError):
            ops.march(mac_to_dev(u, device), fmap_to_dev(om, device), dt, order, blocks_per_cu=bad)


@pytest.mark.parametrize("dim", [2, 3])
def test_pullback_u_and_epilogue(dim, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(20 + dim)
    u0 = rand_mac(rng, n)
    add = rand_mac(rng, n)
    om = random_map(rng, n, dx, disp=3.0)
    fm = fmap_to_dev(om, device)
    out = mac_to_dev(add, device)
    ops_for(n, dx).pullback_u(mac_to_dev(u0, device), fm, out, alpha=0.7, beta=-1.3, add=mac_to_dev(add, device))
    ref = [0.7 * r - 1.3 * c for r, c in zip(O.advect_u(u0, om, dx), add)]
    within("mac_to_np_out", rel_l2_fields(mac_to_np(out), ref), TOL)


@pytest.mark.parametrize("dim", [2, 3])
def test_pullback_s_and_bfecc_s(dim, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(30 + dim)
    s0 = round32(rng.standard_normal(n))
    om1 = random_map(rng, n, dx)
    om2 = random_map(rng, n, dx)
    fm1, fm2 = fmap_to_dev(om1, device), fmap_to_dev(om2, device)
    ops = ops_for(n, dx)
    s_d = torch.tensor(s0, dtype=torch.float32, device=device)
    out = torch.empty_like(s_d)
    ops.pullback_s(s_d, fm1.psi, out)
    within("out.cpu_.numpy", rel_l2(out.cpu().numpy(), O.advect_scalar(s0, om1.psi, dx)), TOL)
    err = torch.empty_like(s_d)
    ops.bfecc_s(s_d, fm1.psi, fm2.psi, err, out)
    within("out.cpu_.numpy", rel_l2(out.cpu().numpy(), O.bfecc_scalar(s0, om1.psi, om2.psi, dx)), TOL)


@pytest.mark.parametrize("dim", [2, 3])
def test_bfecc_u(dim, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(40 + dim)
    u0 = rand_mac(rng, n)
    om1 = random_map(rng, n, dx)
    om2 = random_map(rng, n, dx)
    from dfm.ops import MacField

    out = MacField.empty(n, device)
    err = MacField.empty(n, device)
    ops_for(n, dx).bfecc_u(mac_to_dev(u0, device), fmap_to_dev(om1, device), fmap_to_dev(om2, device), err, out)
    within("mac_to_np_out", rel_l2_fields(mac_to_np(out), O.bfecc(u0, om1, om2, dx)), TOL)


@pytest.mark.parametrize("dim", [2, 3])
def test_bfecc_clamped(dim, device):
    """BFECC_clamp = True for MAC and cell fields against the oracle, and the
    clamp's own property: every value lies within the first pullback's
    neighbour range; the clamp is active on a visible fraction of faces."""
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(45 + dim)
    u0 = rand_mac(rng, n)
    s0 = round32(rng.standard_normal(n))
    om1 = random_map(rng, n, dx)
    om2 = random_map(rng, n, dx)
    from dfm.ops import MacField

    out, err, tmp = MacField.empty(n, device), MacField.empty(n, device), MacField.empty(n, device)
    ops = ops_for(n, dx)
    f1, f2 = fmap_to_dev(om1, device), fmap_to_dev(om2, device)
    ops.bfecc_u_clamp(mac_to_dev(u0, device), f1, f2, err, tmp, out)
    ref = O.bfecc(u0, om1, om2, dx, clamp=True)
    within("mac_to_np_out", rel_l2_fields(mac_to_np(out), ref), TOL)
    plain = O.bfecc(u0, om1, om2, dx)
    changed = sum(np.count_nonzero(np.abs(a - b) > 1e-9) for a, b in zip(ref, plain))
    assert changed > 0.01 * sum(c.size for c in ref)
    so, se, st = (torch.empty(n, device=device) for _ in range(3))
    ops.bfecc_s_clamp(torch.tensor(s0, dtype=torch.float32, device=device), f1.psi, f2.psi, se, st, so)
    within("so.cpu_.numpy", rel_l2(so.cpu().numpy(), O.bfecc_scalar(s0, om1.psi, om2.psi, dx, clamp=True)), TOL)


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 4), (3, 2), (3, 4)])
def test_sl_advect_multi_source(dim, order, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(50 + dim * 5 + order)
    vel = smooth_mac(rng, n, dx)
    srcs = [rand_mac(rng, n), rand_mac(rng, n)]
    ss = [round32(rng.standard_normal(n)) for _ in range(3)]
    dt = -0.5 * dx if order == 2 else 0.5 * dx
    from dfm.ops import MacField

    dsts = [MacField.empty(n, device) for _ in srcs]
    s_src = [torch.tensor(s, dtype=torch.float32, device=device) for s in ss]
    s_dst = [torch.empty_like(t) for t in s_src]
    ops_for(n, dx).sl_advect(mac_to_dev(vel, device), dt, order, [mac_to_dev(s, device) for s in srcs], dsts,
                             s_src, s_dst)
    for src, dst in zip(srcs, dsts):
        within("mac_to_np_dst", rel_l2_fields(mac_to_np(dst), O.sl_advect_u(vel, src, dx, dt, order)), TOL)
    for s, d in zip(ss, s_dst):
        within("d.cpu_.numpy", rel_l2(d.cpu().numpy(), O.sl_advect_s(vel, s, dx, dt, order)), TOL)


@pytest.mark.parametrize("dim", [2, 3])
def test_accumulate_ab3_and_second_source(dim, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(60 + dim)
    om = random_map(rng, n, dx)
    fm = fmap_to_dev(om, device)
    ops = ops_for(n, dx)
    init = rand_mac(rng, n)
    h1 = [np.zeros_like(c) for c in init]
    h2 = [np.zeros_like(c) for c in init]
    d_init, d_h1, d_h2 = mac_to_dev(init, device), mac_to_dev(h1, device), mac_to_dev(h2, device)
    for ind in range(4):
        f = rand_mac(rng, n)
        f2 = rand_mac(rng, n)
        ops.accumulate(mac_to_dev(f, device), fm, d_init, d_h1, d_h2, ind, -0.02, f2=mac_to_dev(f2, device),
                       dt2=-0.01)
        O.accumulate_ab3(f, init, om, h1, h2, ind, dx, -0.02)
        O.accumulate(f2, init, om, dx, -0.01)
        within("mac_to_np_d_init", rel_l2_fields(mac_to_np(d_init), init), TOL)
        within("mac_to_np_d_h1", rel_l2_fields(mac_to_np(d_h1), h1), TOL)
        within("mac_to_np_d_h2", rel_l2_fields(mac_to_np(d_h2), h2), TOL)


@pytest.mark.parametrize("dim,kern", [(2, 3), (2, 4), (3, 3), (3, 4)])
def test_adj_conv_source(dim, kern, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(70 + dim * 5 + kern)
    u = rand_mac(rng, n)
    w = rand_mac(rng, n)
    from dfm.ops import MacField

    r = MacField.empty(n, device)
    ops_for(n, dx).adj_conv_source(mac_to_dev(u, device), mac_to_dev(w, device), kern, r)
    within("mac_to_np_r", rel_l2_fields(mac_to_np(r), O.adj_conv_source(u, w, dx, kern)), TOL)


@pytest.mark.parametrize("dim", [2, 3])
def test_adj_scalar_source_summed(dim, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(80 + dim)
    ss = [round32(rng.standard_normal(n)) for _ in range(2)]
    ls = [round32(rng.standard_normal(n)) for _ in range(2)]
    from dfm.ops import MacField

    r = MacField.empty(n, device)
    ops_for(n, dx).adj_scalar_source([torch.tensor(s, device=device, dtype=torch.float32) for s in ss],
                                     [torch.tensor(v, device=device, dtype=torch.float32) for v in ls], r)
    parts = [O.adj_scalar_source(s, v, dx) for s, v in zip(ss, ls)]
    ref = [parts[0][a] + parts[1][a] for a in range(dim)]
    within("mac_to_np_r", rel_l2_fields(mac_to_np(r), ref), TOL)


@pytest.mark.parametrize("dim", [2, 3])
def test_axpy(dim, device):
    n, dx = GRIDS[dim]
    rng = np.random.default_rng(90 + dim)
    x, y, z = rand_mac(rng, n), rand_mac(rng, n), rand_mac(rng, n)
    dx_, dy_, dz_ = mac_to_dev(x, device), mac_to_dev(y, device), mac_to_dev(z, device)
    ops_for(n, dx).axpy(dx_, 0.25, dy_, dx_, beta=-0.5, z=dz_)
    within("mac_to_np_dx", rel_l2_fields(mac_to_np(dx_), [a + 0.25 * b - 0.5 * c for a, b, c in zip(x, y, z)]), 1e-6)


# FP32 V-cycle/residual with FP64 dots, FP64 pressure and FP64 iterative
# refinement of the true residual: the projected field sits within ~1e-6 of
# the exact projection at every size (without refinement 256^2 stalled at 1e-5).
@pytest.mark.parametrize("n,levels,bound", [((64, 48), 0, 2e-6), ((32, 32, 16), 0, 2e-6), ((256, 256), 6, 2e-6),
                                            ((64, 64, 64), 4, 2e-6)])
def test_mgpcg_projection_matches_exact(n, levels, bound, device):
    from dfm.ops import MGPCG

    dim = len(n)
    dx = 1.0 / n[1]
    rng = np.random.default_rng(100 + len(n) + n[0])
    u = smooth_mac(rng, n, dx)
    u = [c + 0.05 * rng.standard_normal(c.shape) for c in u]
    u = round32(u)
    solver = MGPCG(n, dx, levels=levels, tol=1e-7, max_iters=200)
    d = mac_to_dev(u, device)
    solver.Poisson(d)
    iters, res = solver.status()
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    out = mac_to_np(d)
    err = rel_l2_fields(out, ref)
    print(f"n={n} iters={iters} res={res:.2e} rel-L2 vs exact={err:.2e}")
    assert iters < 60, iters
    assert err < bound
    div = O.divergence(out)
    assert np.abs(div).max() < 1e-4 * max(np.abs(c).max() for c in u)
    # idempotent: a second projection changes nothing beyond FP32 noise, and
    # its right-hand side (FP32 rounding noise) still converges in relative
    # terms (no absolute guard in alpha / beta to stall on)
    solver.Poisson(d)
    within("mac_to_np_d", rel_l2_fields(mac_to_np(d), out), 5e-6)
    assert solver.status()[0] < 60 and solver.noconv_count == 0


@pytest.mark.parametrize("n", [(32, 32), (16, 16, 16), (64, 64, 32)])
def test_mgpcg_scale_invariant(n, device):
    """With a relative tolerance the solve is scale-free: a field scaled by
    1e-6 (a right-hand side far below the reference's absolute 1e-12 guard)
    takes the same iterations and gives the scaled projection."""
    from dfm.ops import MGPCG

    dx = 1.0 / n[1]
    rng = np.random.default_rng(77 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))])
    res = []
    for scale in (1.0, 1e-6):
        solver = MGPCG(n, dx, tol=1e-7, max_iters=200)
        d = mac_to_dev([c * scale for c in u], device)
        solver.Poisson(d)
        res.append((solver.status()[0], [c / scale for c in mac_to_np(d)]))
    assert abs(res[0][0] - res[1][0]) <= 1, (res[0][0], res[1][0])
    within("scaled_projection", rel_l2_fields(res[1][1], res[0][1]), 1e-5)


@pytest.mark.parametrize("n,levels", [((128, 128, 128), 5), ((64, 64, 64), 4), ((256, 256), 6), ((16, 16, 16), 2)])
def test_mgpcg_tail_and_bf16_modes(n, levels, device):
    """The one-workgroup LDS tail (FP32) reproduces the per-level launch
    sequence; BF16 storage of levels >= 2 still converges to the exact
    projection (it only changes the preconditioner)."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    dx = 1.0 / n[1]
    rng = np.random.default_rng(5 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))])
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    out = {}
    for name, kw in [("notail", dict(tail=False)), ("tail", dict()), ("bf16", dict(bf16_coarse=True)),
                     ("exact", dict(bottom_exact=True)), ("bf16exact", dict(bf16_coarse=True, bottom_exact=True))]:
        s = MGPCG(n, dx, levels=levels, tol=1e-7, max_iters=100, **kw)
        d = mac_to_dev(u, device)
        s.Poisson(d)
        it, res = s.status()
        out[name] = (mac_to_np(d), it, s.get_option(L.MG_TAIL), s.get_option(L.MG_BF16_COARSE))
        print(f"n={n} {name}: iters={it} res={res:.2e} tail_level={out[name][2]} bf16={out[name][3]} "
              f"err={rel_l2_fields(out[name][0], ref):.2e}")
        within("out[name][0]", rel_l2_fields(out[name][0], ref), 2e-6)
    assert out["notail"][2] == levels
    assert abs(out["tail"][1] - out["notail"][1]) <= 1
    within("out[_tail_][0]", rel_l2_fields(out["tail"][0], out["notail"][0]), 1e-6)


@pytest.mark.parametrize("check_every", [1, 2])
def test_mgpcg_graph_replays_match_eager(check_every, device):
    """The captured PCG-iteration graph, replayed across several solves with
    eager work and host syncs in between, gives the eager launches' results
    on every solve."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    n = (64, 64, 32)
    dx = 1 / 64
    g = MGPCG(n, dx, tol=1e-6, check_every=check_every)
    e = MGPCG(n, dx, tol=1e-6, graph=False, check_every=check_every)
    rng = np.random.default_rng(31)
    for rep in range(4):
        u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(3)])
        dg, de = mac_to_dev(u, device), mac_to_dev(u, device)
        g.Poisson(dg)
        torch.cuda.synchronize()
        e.Poisson(de)
        assert g.get_option(L.MG_GRAPH) == 2
        assert g.status()[0] == e.status()[0]
        assert rel_l2_fields(mac_to_np(dg), mac_to_np(de)) < 1e-6, rep


@pytest.mark.parametrize("n,levels", [((40, 36, 20), 3), ((72, 56), 3), ((128, 128, 128), 5)])
def test_mgpcg_repeat_solves_bitwise_identical(n, levels, device):
    """The grid reductions (per-block partials, summed in block order by the
    last block to arrive) are deterministic: the same input solved repeatedly
    by one handle (graph replays included) gives bitwise-identical fields and
    iteration counts.  Sizes with odd coarse rows exercise the index
    decomposition's odd-pair edge cases."""
    from dfm.ops import MGPCG

    dx = 1.0 / n[1]
    rng = np.random.default_rng(17 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))])
    s = MGPCG(n, dx, levels=levels, tol=1e-7, max_iters=100)
    outs, its = [], []
    for rep in range(5):
        d = mac_to_dev(u, device)
        s.Poisson(d)
        outs.append(mac_to_np(d))
        its.append(s.status()[0])
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)
    assert len(set(its)) == 1, its
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    within("outs[0]", rel_l2_fields(outs[0], ref), 2e-6)


@pytest.mark.parametrize("n,levels,tail,masked", [((72, 56), 3, 0, False), ((256, 192), 5, 1, False),
                                                  ((40, 36, 20), 3, 0, False), ((128, 128, 128), 5, 1, False),
                                                  ((96, 80), 4, 0, True), ((32, 32, 24), 3, 0, True),
                                                  ((64, 48, 32), 3, 0, False), ((128, 96), 4, 0, False)])
def test_mgpcg_fused_sweeps_bitwise(n, levels, tail, masked, device):
    """The LDS-tiled pre/post smoothing of the coarse levels (option 2) and
    one launch per red-black sweep (the second colour recomputes its
    neighbours' first-colour values, option 1) give bitwise the fields and
    iteration counts of two half-sweep launches (option 0), with and without
    the LDS tail and with an obstacle mask."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    dim = len(n)
    rng = np.random.default_rng(5 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    mask = None
    if masked:
        c = np.stack(np.meshgrid(*[(np.arange(m) + 0.5) / n[1] for m in n], indexing="ij"), -1)
        mask = (np.linalg.norm(c - 0.4, axis=-1) < 0.2).astype(np.float32)
    outs = []
    for fused in (2, 1, 0):
        s = MGPCG(n, 1.0 / n[1], levels=levels, tol=1e-7, max_iters=200)
        s.set_option(L.MG_TAIL, tail)
        s.set_option(L.MG_FUSED_SWEEPS, fused)
        assert s.get_option(L.MG_FUSED_SWEEPS) == fused
        if masked:
            s.set_mask(torch.tensor(mask, device=device))
        res = []
        for rep in range(2):
            d = mac_to_dev(u, device)
            s.Poisson(d)
            res.append((mac_to_np(d), s.status()[0]))
        outs.append(res)
    for other in outs[1:]:
        for (fa, ia), (fb, ib) in zip(outs[0], other):
            assert ia == ib
            for a, b in zip(fa, fb):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [(100, 60), (48, 36, 20)])
def test_mgpcg_auto_levels_on_grids_that_halve_few_times(n, device):
    """Automatic levels stop where the grid no longer halves exactly (100 =
    4 x 25 gives 3 levels), and the projection still matches the exact one."""
    from dfm.ops import MGPCG

    dim = len(n)
    rng = np.random.default_rng(3)
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    s = MGPCG(n, 1.0 / n[1], tol=1e-7, max_iters=300)
    d = mac_to_dev(u, device)
    s.Poisson(d)
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    within("mac_to_np_d", rel_l2_fields(mac_to_np(d), ref), 2e-6)


def _pressure_of(solver, n, device):
    """The handle's FP64 pressure (dfm_mgpcg_pressure) copied into a tensor."""
    import ctypes

    from dfm import _lib as L

    ptr = ctypes.c_void_p()
    L.check(L.lib().dfm_mgpcg_pressure(solver.h, ctypes.byref(ptr)), "mgpcg_pressure")
    out = torch.empty(int(np.prod(n)), dtype=torch.float64, device=device)
    torch.cuda.synchronize()
    # the HIP runtime libdfm.so is linked against (the one torch uses)
    memcpy = L.lib().hipMemcpy
    memcpy.restype = ctypes.c_int
    memcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert memcpy(out.data_ptr(), ptr, out.numel() * 8, 3) == 0  # device to device
    return out.cpu().numpy().reshape(n)


@pytest.mark.parametrize("n,levels", [((64, 48, 32), 3), ((256, 192), 5), ((40, 36, 20), 3), ((192, 160), 4),
                                      ((128, 128, 128), 5)])
def test_mgpcg_split_layout_matches_natural(n, levels, device):
    """The colour-split finest level (red cells in one half of each level-0
    array, black in the other; the default where eligible) gives the
    natural layout's projection up to rounding (its FP64 reductions sum in
    another order), within one iteration, the exact projection, the same
    pressure through dfm_mgpcg_pressure (natural order) and the same
    solution through the b -> x entry."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    dim = len(n)
    rng = np.random.default_rng(23 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    b = rng.standard_normal(n)
    b = round32(b - b.mean())
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    xe = O.neumann_solve_dct(b.astype(np.float64))
    got = {}
    for split in (True, False):
        s = MGPCG(n, 1.0 / n[1], levels=levels, tol=1e-7, max_iters=200, split_layout=split)
        assert s.get_option(L.MG_SPLIT_LAYOUT) == (1 if split else 0)
        d = mac_to_dev(u, device)
        s.Poisson(d)
        it = s.status()[0]
        p = _pressure_of(s, n, device)
        x = torch.zeros(n, dtype=torch.float32, device=device)
        s.solve(torch.tensor(b, dtype=torch.float32, device=device), x)
        xs = x.cpu().numpy().astype(np.float64)
        got[split] = (mac_to_np(d), it, p - p.mean(), xs - xs.mean())
        within("got[split][0]", rel_l2_fields(got[split][0], ref), 2e-6)
        within("got[split][3]", rel_l2(got[split][3], xe), 1e-5)
    (fs, its, ps, xs), (fn, itn, pn, xn) = got[True], got[False]
    assert abs(its - itn) <= 1
    within("fs", rel_l2_fields(fs, fn), 1e-6)
    within("ps", rel_l2(ps, pn), 1e-5)
    within("xs", rel_l2(xs, xn), 1e-5)


@pytest.mark.parametrize("n,levels,bf16", [((64, 48, 32), 3, False), ((40, 36, 24), 3, False),
                                           ((128, 128, 128), 5, True), ((40, 36, 20), 3, False)])
def test_mgpcg_row_kernels_match_pair_kernels(n, levels, bf16, device):
    """The finest colour-split level's row kernels (4 cells of a row and
    colour per thread) give the per-pair kernels' projection up to rounding
    (the smoothed values are the same; the z.r sums add in another order),
    within one iteration, and the exact projection.  (40, 36, 20) has an
    innermost extent that is not a multiple of 8: the option is on but the
    pair kernels run, so the two solves are bitwise equal."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    rng = np.random.default_rng(31 + n[0] + n[2])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(3)])
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    got = {}
    for rows in (1, 0):
        s = MGPCG(n, 1.0 / n[1], levels=levels, tol=1e-7, max_iters=200, bf16_coarse=bf16)
        assert s.get_option(L.MG_ROW_KERNELS) == 1
        s.set_option(L.MG_ROW_KERNELS, rows)
        assert s.get_option(L.MG_ROW_KERNELS) == rows
        d = mac_to_dev(u, device)
        s.Poisson(d)
        got[rows] = ([c.clone() for c in d.c], s.status()[0])
        within("row-kernel projection vs exact", rel_l2_fields(mac_to_np(d), ref), 2e-5 if bf16 else 2e-6)
    (fr, itr), (fp, itp) = got[1], got[0]
    assert abs(itr - itp) <= 1
    if n[2] % 8:
        assert itr == itp and all(torch.equal(a, b) for a, b in zip(fr, fp))
    else:
        within("row vs pair kernels", rel_l2_fields([c.cpu().numpy() for c in fr], [c.cpu().numpy() for c in fp]),
               1e-6)


@pytest.mark.parametrize("n,levels,bf16,exact", [((64, 48, 32), 3, False, False), ((64, 48, 32), 3, False, True),
                                                 ((48, 40, 40), 3, True, False), ((128, 128, 128), 5, True, False),
                                                 ((128, 128, 128), 5, False, True), ((40, 36, 20), 3, False, False)])
def test_mgpcg_natural_row_kernels_bitwise(n, levels, bf16, exact, device):
    """On natural-order levels (every level here: the finest is not
    colour-split) the row half-sweeps, which update a row's 4 cells of one
    colour from two 16-byte loads per neighbour row, are the per-pair
    half-sweeps operation for operation: the solves with them on every
    eligible level (option 2), on the wide levels only (1: here only the
    128^3 level) and with none (0) give the same bits, iteration counts and
    pressure.  (48, 40, 40) has a level (24, 20, 20) whose innermost extent
    is not a multiple of 8, so row and pair sweeps alternate between levels;
    (40, 36, 20) runs no row sweep.  The BF16 cases read a BF16 coarse
    correction in the fused prolongation; the exact cases end the V-cycle
    with the DCT bottom solve in the one-workgroup LDS tail.  Option values
    outside 0..2 are refused."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    rng = np.random.default_rng(7 + n[0] + n[2])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(3)])
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    got = {}
    for rows in (2, 1, 0):
        s = MGPCG(n, 1.0 / n[1], levels=levels, tol=1e-7, max_iters=200, bf16_coarse=bf16, split_layout=False,
                  bottom_exact=exact)
        assert s.get_option(L.MG_SPLIT_LAYOUT) == 0
        assert s.get_option(L.MG_BOTTOM_EXACT) == int(exact)
        s.set_option(L.MG_ROW_KERNELS, rows)
        assert s.get_option(L.MG_ROW_KERNELS) == rows
        for bad in (-1, 3):
            with pytest.raises(L.DfmError):
                s.set_option(L.MG_ROW_KERNELS, bad)
        assert s.get_option(L.MG_ROW_KERNELS) == rows
        d = mac_to_dev(u, device)
        s.Poisson(d)
        got[rows] = ([c.clone() for c in d.c], s.status()[0], _pressure_of(s, n, device))
        within("natural row-kernel projection vs exact", rel_l2_fields(mac_to_np(d), ref), 2e-5 if bf16 else 2e-6)
    fp, itp, pp = got[0]
    for rows in (2, 1):
        fr, itr, pr = got[rows]
        assert itr == itp
        assert all(torch.equal(a, b) for a, b in zip(fr, fp))
        assert np.array_equal(pr, pp)


def test_mgpcg_first_solve_on_side_stream(device):
    """Creating a handle only allocates.  Its reduction counters, control
    block and x are zeroed on the stream of the first call that launches
    work, so the pressure accessor refuses a handle that has not solved yet,
    and a first solve on a non-default stream projects like one on the
    default stream."""
    import ctypes

    from dfm import _lib as L
    from dfm.ops import MGPCG

    n = (96, 64)
    rng = np.random.default_rng(5)
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(2)])
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    s = MGPCG(n, 1.0 / n[1], tol=1e-7, max_iters=200)
    ptr = ctypes.c_void_p()
    assert L.lib().dfm_mgpcg_pressure(s.h, ctypes.byref(ptr)) == -1
    side = torch.cuda.Stream()
    d = mac_to_dev(u, device)
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        s.Poisson(d)
    side.synchronize()
    within("mac_to_np_d", rel_l2_fields(mac_to_np(d), ref), 2e-6)
    p = _pressure_of(s, n, device)
    assert np.isfinite(p).all() and np.abs(p).max() > 0


def test_mgpcg_split_layout_eligibility(device):
    """The split layout needs an even innermost extent, no obstacle mask and
    a finest level outside the LDS tail (grids whose levels all fit in one
    workgroup's LDS are solved there whole)."""
    from dfm import _lib as L
    from dfm.ops import MGPCG

    assert MGPCG((48, 40, 32), 1 / 40, levels=2).get_option(L.MG_SPLIT_LAYOUT) == 1
    assert MGPCG((48, 40, 31), 1 / 40, levels=1).get_option(L.MG_SPLIT_LAYOUT) == 0
    assert MGPCG((256, 200), 1 / 200, levels=2).get_option(L.MG_SPLIT_LAYOUT) == 1
    assert MGPCG((256, 199), 1 / 199, levels=1).get_option(L.MG_SPLIT_LAYOUT) == 0
    assert MGPCG((16, 16, 16), 1 / 16, levels=1).get_option(L.MG_SPLIT_LAYOUT) == 0  # level 0 in the tail
    assert MGPCG((48, 40, 32), 1 / 40, levels=2, split_layout=False).get_option(L.MG_SPLIT_LAYOUT) == 0
    s = MGPCG((256, 256), 1 / 256, levels=3)
    assert s.get_option(L.MG_SPLIT_LAYOUT) == 1
    s.set_mask(torch.zeros((256, 256), device=device))
    assert s.get_option(L.MG_SPLIT_LAYOUT) == 0
    s.set_mask(None)
    assert s.get_option(L.MG_SPLIT_LAYOUT) == 1


def test_mgpcg_solve_and_status(device):
    from dfm.ops import MGPCG

    n = (32, 32, 32)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(n)
    b -= b.mean()
    solver = MGPCG(n, 1 / 32, levels=3, tol=1e-6)
    x = torch.zeros(n, dtype=torch.float32, device=device)
    solver.solve(torch.tensor(b, dtype=torch.float32, device=device), x)
    it, res = solver.status()
    assert res < 1e-6 and 0 < it < 40
    xe = O.neumann_solve_dct(round32(b))
    xg = x.cpu().numpy().astype(np.float64)
    within("xg_xg.mean", rel_l2(xg - xg.mean(), xe), 1e-5)


@pytest.mark.parametrize("n", [(48, 40), (20, 16, 24)])
def test_central_vector_split_and_max_speed(n, device):
    """get_central_vector / split_central_vector / calc_max_speed against
    numpy restatements: the central average, the clamped split (a constant
    central field splits to the same constant; split of a field's central
    average differs from it only by the two-cell smoothing), and the maximum
    speed with its 1e-3 floor."""
    from dfm.ops import MacField, Ops

    dim = len(n)
    rng = np.random.default_rng(40 + dim)
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    ops = Ops(n, 1.0 / n[1])
    du = mac_to_dev(u, device)
    c = torch.empty((dim, *n), device=device)
    ops.central_vector(du, c)
    cen = []
    for a in range(dim):
        hi = [slice(None)] * dim
        lo = [slice(None)] * dim
        hi[a], lo[a] = slice(1, None), slice(0, -1)
        cen.append(0.5 * (u[a][tuple(hi)] + u[a][tuple(lo)]))
    assert np.array_equal(c.cpu().numpy(), np.stack(cen).astype(np.float32))
    # split: faces average the two adjacent cells, clamped at the walls
    out = MacField.empty(n, device)
    ops.split_central_vector(c, out)
    for a in range(dim):
        pad = np.concatenate([np.take(cen[a], [0], axis=a), cen[a], np.take(cen[a], [-1], axis=a)], axis=a)
        hi = [slice(None)] * dim
        lo = [slice(None)] * dim
        hi[a], lo[a] = slice(1, None), slice(0, -1)
        want = 0.5 * (pad[tuple(hi)] + pad[tuple(lo)])
        assert np.abs(out.c[a].cpu().numpy() - want).max() <= 1e-6 * np.abs(want).max()
    const = torch.full((dim, *n), 0.75, device=device)
    ops.split_central_vector(const, out)
    assert all(bool(torch.all(t == 0.75)) for t in out.c)
    # maximum speed
    m = torch.empty(1, device=device)
    ops.max_speed(du, m)
    want = max(1e-3, float(np.sqrt(sum(x.astype(np.float64) ** 2 for x in cen)).max()))
    assert abs(float(m.item()) - want) <= 1e-6 * want
    ops.max_speed(MacField([torch.zeros_like(t) for t in du.c]), m)
    assert float(m.item()) == np.float32(1e-3)


# The 3D adjoint sources run on an LDS-tiled fixed-stencil kernel (8x8x32
# output tiles); the per-point gather kernel stays behind an option.  Both
# form the same weights and contract z, y, x in the same order, so they agree
# bit for bit; grids with partial tiles on every axis, both splines.
TILE_GRIDS = [((12, 16, 20), 1 / 20), ((9, 17, 33), 1 / 33), ((40, 8, 70), 1 / 64)]


@pytest.mark.parametrize("gi", range(len(TILE_GRIDS)))
@pytest.mark.parametrize("kern", [3, 4])
def test_tiled_adj_conv_bitwise_gather_and_oracle(gi, kern, device):
    from dfm.ops import FM_TILED_SOURCES, MacField, set_flowmap_option

    n, dx = TILE_GRIDS[gi]
    rng = np.random.default_rng(500 + 7 * gi + kern)
    u, w = rand_mac(rng, n), rand_mac(rng, n)
    ud, wd = mac_to_dev(u, device), mac_to_dev(w, device)
    ops = ops_for(n, dx)
    r_t, r_g = MacField.empty(n, device), MacField.empty(n, device)
    prev = set_flowmap_option(FM_TILED_SOURCES, 1)
    try:
        ops.adj_conv_source(ud, wd, kern, r_t)
        set_flowmap_option(FM_TILED_SOURCES, 0)
        ops.adj_conv_source(ud, wd, kern, r_g)
    finally:
        set_flowmap_option(FM_TILED_SOURCES, prev)
    for a in range(3):
        assert torch.equal(r_t.c[a], r_g.c[a]), f"component {a}"
    within("tiled_adj_conv", rel_l2_fields(mac_to_np(r_t), O.adj_conv_source(u, w, dx, kern)), TOL)


@pytest.mark.parametrize("gi", range(len(TILE_GRIDS)))
@pytest.mark.parametrize("nscal", [0, 1, 3])
def test_tiled_adj_scalar_bitwise_gather_and_oracle(gi, nscal, device):
    from dfm.ops import FM_TILED_SOURCES, MacField, set_flowmap_option

    n, dx = TILE_GRIDS[gi]
    rng = np.random.default_rng(600 + 7 * gi + nscal)
    ss = [round32(rng.standard_normal(n)) for _ in range(nscal)]
    ls = [round32(rng.standard_normal(n)) for _ in range(nscal)]
    sd = [torch.tensor(s, device=device, dtype=torch.float32) for s in ss]
    ld = [torch.tensor(v, device=device, dtype=torch.float32) for v in ls]
    ops = ops_for(n, dx)
    r_t, r_g = MacField.empty(n, device), MacField.empty(n, device)
    prev = set_flowmap_option(FM_TILED_SOURCES, 1)
    try:
        ops.adj_scalar_source(sd, ld, r_t)
        set_flowmap_option(FM_TILED_SOURCES, 0)
        ops.adj_scalar_source(sd, ld, r_g)
    finally:
        set_flowmap_option(FM_TILED_SOURCES, prev)
    for a in range(3):
        assert torch.equal(r_t.c[a], r_g.c[a]), f"component {a}"
    if nscal == 0:
        assert all(float(c.abs().max()) == 0.0 for c in r_t.c)
        return
    parts = [O.adj_scalar_source(s, v, dx) for s, v in zip(ss, ls)]
    ref = [sum(p[a] for p in parts) for a in range(3)]
    within("tiled_adj_scalar", rel_l2_fields(mac_to_np(r_t), ref), TOL)


# At S >= 64 the FP32 clamp bound S - 2 - 1e-6 rounds to S - 2, so a cubic
# sample at the upper wall has its (zero-weight) fourth tap clamped back to
# S - 1; the tiled kernel must read that tap's value, not an unloaded LDS
# slot.  A first call on NaN fields leaves NaN in LDS, so a stale read would
# show as a non-finite result.
@pytest.mark.parametrize("n", [(70, 66, 64), (64, 96, 130)])
def test_tiled_sources_upper_wall_clamp(n, device):
    from dfm.ops import FM_TILED_SOURCES, MacField, set_flowmap_option

    dx = 1.0 / n[2]
    rng = np.random.default_rng(700 + n[0])
    u, w = rand_mac(rng, n), rand_mac(rng, n)
    ops = ops_for(n, dx)
    nan = MacField([torch.full(c.shape, float("nan"), device=device) for c in mac_to_dev(u, device).c])
    r_t, r_g, junk = MacField.empty(n, device), MacField.empty(n, device), MacField.empty(n, device)
    ud, wd = mac_to_dev(u, device), mac_to_dev(w, device)
    s = [torch.tensor(round32(rng.standard_normal(n)), device=device, dtype=torch.float32)]
    s_nan = [torch.full(n, float("nan"), device=device)]
    prev = set_flowmap_option(FM_TILED_SOURCES, 1)
    try:
        ops.adj_conv_source(nan, nan, 3, junk)
        ops.adj_conv_source(ud, wd, 3, r_t)
        ops.adj_scalar_source(s_nan, s_nan, junk)
        st = MacField.empty(n, device)
        ops.adj_scalar_source(s, s, st)
        set_flowmap_option(FM_TILED_SOURCES, 0)
        ops.adj_conv_source(ud, wd, 3, r_g)
        sg = MacField.empty(n, device)
        ops.adj_scalar_source(s, s, sg)
    finally:
        set_flowmap_option(FM_TILED_SOURCES, prev)
    for a in range(3):
        assert bool(torch.isfinite(r_t.c[a]).all()) and bool(torch.isfinite(st.c[a]).all())
        assert torch.equal(r_t.c[a], r_g.c[a]), f"adj_conv component {a}"
        assert torch.equal(st.c[a], sg.c[a]), f"adj_scalar component {a}"


@pytest.mark.parametrize("n,split", [((48, 40, 32), True), ((48, 40, 32), False), ((96, 64), True)])
def test_mgpcg_zero_rhs_after_solve(n, split, device):
    """x starts at zero without being written (the first x/r update does not
    read it).  A projection whose right-hand side is 0 runs no iteration:
    after a solve that left x non-zero, it must still return the field
    unchanged, a zero pressure and a zero b -> x solution."""
    from dfm.ops import MGPCG

    dim = len(n)
    rng = np.random.default_rng(11 + n[0])
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)])
    s = MGPCG(n, 1.0 / n[1], tol=1e-7, max_iters=200, split_layout=split)
    s.Poisson(mac_to_dev(u, device))
    assert s.status()[0] > 0 and float(np.abs(_pressure_of(s, n, device)).max()) > 0
    z = [np.zeros(O.face_shape(n, a)) for a in range(dim)]
    d = mac_to_dev(z, device)
    s.Poisson(d)
    assert s.status()[0] == 0
    assert all(float(c.abs().max()) == 0.0 for c in d.c)
    assert float(np.abs(_pressure_of(s, n, device)).max()) == 0.0
    x = torch.full(n, 7.0, dtype=torch.float32, device=device)
    s.solve(torch.zeros(n, dtype=torch.float32, device=device), x)
    assert float(x.abs().max()) == 0.0
    # and a real solve after it starts from zero again
    d = mac_to_dev(u, device)
    s.Poisson(d)
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    within("after_zero", rel_l2_fields(mac_to_np(d), ref), 2e-6)
