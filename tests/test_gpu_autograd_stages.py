"""Per-stage differentiation on the GPU: the vector-Jacobian product kernels
behind dfm.autograd's March / PullbackU / PullbackS / SLAdvect / Project, and
the whole-loss gradient of a 3D rollout against central differences.

* Transposes of the linear stages (map or velocity fixed) satisfy the
  dot-product identity <A x, y> = <x, A^T y> with A the forward kernel itself
  (rel <= 1e-5; FP32 scatter sums).
* Derivatives w.r.t. points (march, pullback map, advecting velocity) are
  checked against central differences of the FP64 oracle's forward stage
  (h = 1e-6), evaluated at the GPU's FP32 inputs: a dropped term, a wrong sign
  or a transposed Jacobian in the reverse sweep changes the directional
  derivative by O(1) and fails the 1e-4 bound.
* A chain march -> pullback -> projection (+ scalar pullback) composed from
  the autograd Functions is the derivative of the discrete map: its gradient
  matches FP64 differences of the oracle's chain to 1e-4.
* LFMRollout in 3D is the reference's continuous adjoint, which differs from
  the discrete derivative by the adjoint equations' discretisation gap.  On
  this setup (32^3, 5 tubes, R = 3, 6 steps) the FP64 oracle's own adjoint
  misses FP64 central differences by at most 5.8e-3 |g| |d| along smooth
  directions (1.9e-2 at 16^3: the gap shrinks with refinement), so the
  FP32 GPU check along the same kind of directions is bounded at 1e-2."""
import numpy as np
import pytest
import torch

from gpu_util import rel_l2, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu

DIMS = {2: (24, 20), 3: (16, 16, 12)}


def smooth_field(shape, rng, modes=3):
    """Sum of a few low cosine modes (unit-cube coordinates)."""
    X = np.meshgrid(*[np.linspace(0.0, 1.0, s) for s in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(modes):
        k = rng.integers(1, 3, size=len(shape))
        ph = rng.uniform(0.0, 2.0 * np.pi, size=len(shape))
        f += rng.standard_normal() * np.prod([np.cos(np.pi * k[i] * X[i] + ph[i]) for i in range(len(shape))], axis=0)
    return f


def smooth_mac(n, rng, amp, drift=0.0):
    return [round32(amp * smooth_field(O.face_shape(n, a), rng) + drift * (a + 1)) for a in range(len(n))]


def dev(arrs, device):
    return [torch.tensor(np.asarray(a, np.float32), device=device) for a in arrs]


def dot(xs, ys):
    return sum(float(np.sum(np.asarray(x, np.float64) * np.asarray(y, np.float64))) for x, y in zip(xs, ys))


def tnp(ts):
    return [t.detach().cpu().numpy().astype(np.float64) for t in ts]


def displaced_map(ops, n, device, rng, dx):
    """A flow map marched off the identity (psi displaced, T not the identity)."""
    from dfm.ops import FlowMap, MacField

    fm = FlowMap(n, device)
    ops.reset_identity(fm)
    w = MacField(dev(smooth_mac(n, rng, 1.0, 0.2), device))
    for _ in range(3):
        ops.march(w, fm, 0.7 * dx, 4)
    return fm


def aos(t):
    """(dim, *shape) device tensor -> (*shape, dim) FP64 numpy (the oracle's layout)."""
    return np.ascontiguousarray(np.moveaxis(t.detach().cpu().numpy().astype(np.float64), 0, -1))


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_stage_transposes(device, dim):
    from dfm.ops import MacField, Ops

    n = DIMS[dim]
    dx = 1.0 / n[1]
    rng = np.random.default_rng(dim)
    ops = Ops(n, dx)
    fm = displaced_map(ops, n, device, rng, dx)
    rnd = lambda shp: torch.tensor(rng.standard_normal(shp).astype(np.float32), device=device)  # noqa: E731
    x = MacField([rnd(O.face_shape(n, a)) for a in range(dim)])
    y = MacField([rnd(O.face_shape(n, a)) for a in range(dim)])

    # pullback of a MAC field: out_a = alpha T_a . x(psi_a)
    out = MacField.empty(n, device)
    ops.pullback_u(x, fm, out, alpha=0.8)
    gx = MacField.empty(n, device, zero=True)
    ops.pullback_u_vjp(y, fm, gx, alpha=0.8)
    lhs, rhs = dot(tnp(out.c), tnp(y.c)), dot(tnp(x.c), tnp(gx.c))
    within("pullback_u <Ax,y> vs <x,A'y>", abs(lhs - rhs) / abs(lhs), 1e-5)

    # pullback of a cell scalar
    s, ys = rnd(n), rnd(n)
    so = torch.empty_like(s)
    ops.pullback_s(s, fm.psi, so, alpha=1.3)
    gs = torch.zeros_like(s)
    ops.pullback_s_vjp(ys, fm.psi, gs, alpha=1.3)
    lhs, rhs = dot(tnp([so]), tnp([ys])), dot(tnp([s]), tnp([gs]))
    within("pullback_s <Ax,y> vs <x,A'y>", abs(lhs - rhs) / abs(lhs), 1e-5)

    # semi-Lagrangian advection of two MAC fields and two scalars
    vel = MacField(dev(smooth_mac(n, rng, 1.0, 0.3), device))
    s2, ys2 = rnd(n), rnd(n)
    x2 = MacField([rnd(O.face_shape(n, a)) for a in range(dim)])
    y2 = MacField([rnd(O.face_shape(n, a)) for a in range(dim)])
    for order in (2, 4):
        vd = [MacField.empty(n, device), MacField.empty(n, device)]
        sd = [torch.empty_like(s), torch.empty_like(s)]
        ops.sl_advect(vel, 1.5 * dx, order, [x, x2], vd, [s, s2], sd)
        gv = [MacField.empty(n, device, zero=True), MacField.empty(n, device, zero=True)]
        gsl = [torch.zeros_like(s), torch.zeros_like(s)]
        ops.sl_advect_vjp(vel, 1.5 * dx, order, [y, y2], gv, [ys, ys2], gsl)
        lhs = dot(tnp(vd[0].c + vd[1].c + sd), tnp(y.c + y2.c + [ys, ys2]))
        rhs = dot(tnp(x.c + x2.c + [s, s2]), tnp(gv[0].c + gv[1].c + gsl))
        within(f"sl_advect order {order} <Ax,y> vs <x,A'y>", abs(lhs - rhs) / abs(lhs), 1e-5)


@pytest.mark.parametrize("dim", [2, 3])
def test_projection_is_symmetric(device, dim):
    """<P x, y> = <x, P y>: the projection's VJP is the projection."""
    from dfm.autograd import project
    from dfm.ops import MGPCG

    n = DIMS[dim]
    dx = 1.0 / n[1]
    rng = np.random.default_rng(10 + dim)
    solver = MGPCG(n, dx, tol=1e-7)
    x = [t.requires_grad_(True) for t in dev([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)], device)]
    y = dev([rng.standard_normal(O.face_shape(n, a)) for a in range(dim)], device)
    px = project(solver, x)
    loss = sum(torch.sum(a.double() * b.double()) for a, b in zip(px, y))
    loss.backward()
    lhs = float(loss.detach())
    rhs = dot(tnp(x), tnp([t.grad for t in x]))
    nrm = np.sqrt(dot(tnp(x), tnp(x)) * dot(tnp(y), tnp(y)))
    within("projection symmetry", abs(lhs - rhs) / nrm, 1e-5)
    # and the projection is the oracle's exact one
    ref = [a.copy() for a in tnp(x)]
    O.project_exact(ref)
    within("projection vs exact", rel_l2(np.concatenate([a.ravel() for a in tnp(px)]),
                                         np.concatenate([a.ravel() for a in ref])), 1e-4)


def oracle_march_loss(u, maps, w, dx, dt, order):
    """sum <w, marched map> with the FP64 oracle's march; maps / w: lists of
    (psi, T or None) in the oracle's (..., d) layout."""
    tot = 0.0
    for (psi, T), (wp, wT) in zip(maps, w):
        p = psi.copy()
        t = T.copy() if T is not None else None
        O.march(u, p, t, dx, dt, order)
        tot += float(np.sum(p * wp)) + (float(np.sum(t * wT)) if t is not None else 0.0)
    return tot


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 4), (3, 2), (3, 4)])
def test_march_vjp_vs_fp64_differences(device, dim, order):
    """dfm_march_vjp against central differences of the oracle's FP64 march,
    separately along velocity, point and Jacobian-column directions."""
    from dfm.ops import FlowMap, MacField, Ops

    n = DIMS[dim]
    dx = 1.0 / n[1]
    dt = 1.3 * dx
    rng = np.random.default_rng(20 + 2 * dim + order)
    ops = Ops(n, dx)
    fm0 = displaced_map(ops, n, device, rng, dx)
    u64 = smooth_mac(n, rng, 1.0, 0.25)
    u = MacField(dev(u64, device))
    sets = [(fm0.psi_f[a], fm0.T_f[a]) for a in range(dim)] + [(fm0.psi, None)]
    w = [(rng.standard_normal(p.shape), rng.standard_normal(p.shape) if T is not None else None) for p, T in sets]
    # GPU reverse sweep: gradients of sum <w, marched map>
    g_psi_f = [torch.tensor(w[a][0].astype(np.float32), device=device) for a in range(dim)]
    g_T_f = [torch.tensor(w[a][1].astype(np.float32), device=device) for a in range(dim)]
    g_psi = torch.tensor(w[dim][0].astype(np.float32), device=device)
    gfm = FlowMap.wrap(g_psi, g_psi_f, g_T_f)
    gu = MacField.empty(n, device, zero=True)
    ops.march_vjp(u, fm0, dt, order, gfm, gu)
    # oracle in FP64 at the same inputs
    maps = [(aos(p), aos(T) if T is not None else None) for p, T in sets]
    w_aos = [(np.ascontiguousarray(np.moveaxis(wp, 0, -1)), None if wT is None else np.ascontiguousarray(
        np.moveaxis(wT, 0, -1))) for wp, wT in w]
    g_maps = [(aos(gfm.psi_f[a]), aos(gfm.T_f[a])) for a in range(dim)] + [(aos(gfm.psi), None)]
    h = 1e-6

    def fd(du=None, dmaps=None):
        def shifted(sg):
            uu = [a + sg * h * b for a, b in zip(u64, du)] if du is not None else u64
            mm = maps if dmaps is None else [
                (p + sg * h * dp, None if T is None else T + sg * h * dT) for (p, T), (dp, dT) in zip(maps, dmaps)]
            return oracle_march_loss(uu, mm, w_aos, dx, dt, order)
        return (shifted(1.0) - shifted(-1.0)) / (2 * h)

    # velocity direction
    du = [smooth_field(c.shape, rng) for c in u64]
    ref, got = fd(du=du), dot(tnp(gu.c), du)
    print(f"march d/du: fd {ref:+.8e} vjp {got:+.8e}")
    within("march vjp velocity", abs(got - ref) / abs(ref), 1e-4)
    # point direction (every set), then Jacobian-column direction
    dp = [(0.3 * dx * rng.standard_normal(p.shape), None if T is None else np.zeros_like(T)) for p, T in maps]
    ref = fd(dmaps=dp)
    got = sum(dot([gp], [d]) for (gp, _), (d, _) in zip(g_maps, dp))
    print(f"march d/dpsi: fd {ref:+.8e} vjp {got:+.8e}")
    within("march vjp points", abs(got - ref) / abs(ref), 1e-4)
    dT = [(np.zeros_like(p), None if T is None else rng.standard_normal(T.shape)) for p, T in maps]
    ref = fd(dmaps=dT)
    got = sum(dot([gT], [d]) for (_, gT), (_, d) in zip(g_maps, dT) if d is not None)
    print(f"march d/dT: fd {ref:+.8e} vjp {got:+.8e}")
    within("march vjp Jacobian columns", abs(got - ref) / abs(ref), 1e-4)


@pytest.mark.parametrize("dim", [2, 3])
def test_stage_chain_gradient_vs_fp64_differences(device, dim):
    """Loss = <wu, P(T . u0(psi))> + <ws, s0(psi)> + <wa, SL(vel; a0)> with the
    map marched from the identity through vel: the autograd gradient
    (March, PullbackU, PullbackS, Project, SLAdvect) w.r.t. vel, u0, s0 and a0
    against FP64 central differences of the same chain in the oracle."""
    from dfm import autograd as A
    from dfm.ops import MGPCG, FlowMap, Ops

    n = DIMS[dim]
    dx = 1.0 / n[1]
    dt = 1.1 * dx
    rng = np.random.default_rng(40 + dim)
    ops = Ops(n, dx)
    solver = MGPCG(n, dx, tol=1e-7)
    vel64 = smooth_mac(n, rng, 1.0, 0.2)
    u064 = smooth_mac(n, rng, 1.0)
    s064 = round32(smooth_field(n, rng))
    a064 = round32(smooth_field(n, rng))
    wu = [rng.standard_normal(O.face_shape(n, a)) for a in range(dim)]
    ws, wa = rng.standard_normal(n), rng.standard_normal(n)

    vel = [t.requires_grad_(True) for t in dev(vel64, device)]
    u0 = [t.requires_grad_(True) for t in dev(u064, device)]
    s0 = torch.tensor(s064.astype(np.float32), device=device, requires_grad=True)
    a0 = torch.tensor(a064.astype(np.float32), device=device, requires_grad=True)
    idm = FlowMap(n, device)
    ops.reset_identity(idm)
    psi_f, T_f, psi = A.march(ops, vel, idm.psi_f, idm.T_f, dt, 4, psi=idm.psi)
    u1 = A.project(solver, A.pullback_u(ops, u0, psi_f, T_f))
    s1 = A.pullback_s(ops, s0, psi)
    _, (a1,) = A.sl_advect(ops, vel, dt, 2, s_src=[a0])
    loss = sum(torch.sum(x.double() * torch.tensor(w, device=device)) for x, w in zip(u1, wu)) + \
        torch.sum(s1.double() * torch.tensor(ws, device=device)) + torch.sum(a1.double() * torch.tensor(wa, device=device))
    loss.backward()

    def oracle_loss(v, u, s, a):
        fm = O.FlowMap(n, dx, np.float64)
        fm.march(v, dt, 4)
        uu = O.advect_u(u, fm, dx)
        O.project_exact(uu)
        return dot(uu, wu) + dot([O.advect_scalar(s, fm.psi, dx)], [ws]) + \
            dot([O.sl_advect_s(v, a, dx, dt, 2)], [wa])

    within("chain loss vs oracle", abs(float(loss) - oracle_loss(vel64, u064, s064, a064)) / abs(float(loss)), 1e-5)
    h = 1e-6
    dirs = {
        "vel": ([smooth_field(c.shape, rng) for c in vel64], None, None, None),
        "u0": (None, [smooth_field(c.shape, rng) for c in u064], None, None),
        "s0": (None, None, smooth_field(n, rng), None),
        "a0": (None, None, None, smooth_field(n, rng)),
    }
    grads = {"vel": tnp([t.grad for t in vel]), "u0": tnp([t.grad for t in u0]), "s0": tnp([s0.grad]),
             "a0": tnp([a0.grad])}
    for name, (dv, du, ds, da) in dirs.items():
        def at(sg):
            return oracle_loss([a + sg * h * b for a, b in zip(vel64, dv)] if dv is not None else vel64,
                               [a + sg * h * b for a, b in zip(u064, du)] if du is not None else u064,
                               s064 + sg * h * ds if ds is not None else s064,
                               a064 + sg * h * da if da is not None else a064)
        ref = (at(1.0) - at(-1.0)) / (2 * h)
        d = dv or du or [x for x in (ds, da) if x is not None]
        got = dot(grads[name], d)
        print(f"chain d/d{name}: fd {ref:+.8e} autograd {got:+.8e}")
        within(f"chain gradient {name}", abs(got - ref) / abs(ref), 1e-4)


def test_rollout_3d_vs_central_differences(device):
    """Whole-loss central differences of the 3D rollout (32^3, h = 1e-3 of
    |u0| along 4 smooth directions) against the autograd gradient (the
    continuous adjoint), bounded by the measured adjoint gap."""
    from dfm import LFM_Diff_Simulator, scenes
    from dfm.autograd import density_loss

    n, R, steps = (32, 32, 32), 3, 6
    sc = scenes.scene_3d(n, seed=2, tubes=5)
    dx = sc["dx"]
    u064 = round32(sc["u0"])
    s0 = torch.tensor(np.asarray(sc["density"], np.float32), device=device)
    tgt = torch.tensor(np.asarray(sc["target"], np.float32), device=device)
    sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, total_steps=steps, tol=1e-7)
    u0 = [t.requires_grad_(True) for t in dev(u064, device)]
    loss = density_loss(sim, steps, s0, u0, tgt)
    loss.backward()
    g = tnp([t.grad for t in u0])
    gn = np.sqrt(dot(g, g))
    un = np.sqrt(dot(u064, u064))
    rng = np.random.default_rng(1)
    worst = 0.0
    with torch.no_grad():
        def L(u):
            return float(density_loss(sim, steps, s0, dev(u, device), tgt))

        for k in range(4):
            d = [smooth_field(c.shape, rng) for c in u064]
            dn = np.sqrt(dot(d, d))
            d = [x * (un / dn) for x in d]  # |d| = |u0|
            h = 1e-3
            fd = (L([a + h * b for a, b in zip(u064, d)]) - L([a - h * b for a, b in zip(u064, d)])) / (2 * h)
            ad = dot(g, d)
            err = abs(fd - ad) / (gn * un)
            worst = max(worst, err)
            print(f"direction {k}: central FD {fd:+.6e}  adjoint {ad:+.6e}  |err|/(|g||d|) {err:.3e}")
    within("3D rollout adjoint vs central differences", worst, 1e-2)
