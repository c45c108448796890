"""FP64 (or FP32-mode) CPU restatement of the leapfrog flow-map step and its
adjoint -- TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Point kernels (B-spline interpolation, RK march, pullback, semi-Lagrangian
back-trace, MGPCG) live in ``dfm_oracle.cpp`` (built to ``liboracle.so`` by
``oracle/Makefile``); this module holds the field-level compositions and the
step orchestration, each citing the reference lines it restates.

What pins each function (beyond the shared reading of the reference):

=====================  ==========================================================
function               pinned by (tests/test_oracle_pins.py)
=====================  ==========================================================
N_k, dN_k              partition of unity, sum dN = 0, node weights, knot
                       continuity, dN = finite difference of N
interp / mac_eval      exact reproduction of constant / linear fields and their
                       gradients away from the clamp band; clamp behaviour by
                       brute force; MAC offsets with u = (x, 2y, 3z)
march                  uniform flow (exact), linear flow vs matrix exponential
                       (observed order 2 / 4), rotation det(T) ~ 1
pullback / bfecc       identity maps reproduce the B-spline smoothing in closed
                       form; linear fields pulled back exactly
sl_advect              uniform flow == exact shift of a linear field
project_exact          DCT-II eigen-decomposition of the Neumann Laplacian;
                       checked against scipy.sparse spsolve and divergence-free
                       output, P idempotent and symmetric
project_mgpcg          converges to project_exact
project_masked_exact   empty mask == project_exact; divergence-free on fluid
                       cells, zero on walls and solid-cell faces, idempotent,
                       sparse operator symmetric with zero row sums
mask_velocity /        zero mask is a no-op; solid-cell faces take the boundary
mask_scalar            velocity; a constant scalar stays constant; a solid cell
                       with no fluid neighbour gets 0; adjoint form zeroes
adj sources            closed forms on polynomial fields
gaussian_force         a force far from the walls integrates to
                       s (pi/r)(1 - e^-5) / dx^d; one force on a node gives s
                       there; forces add; the FP32 cut-off admits exactly the
                       nodes with fp32 (m*m) r < 5 (brute force per node)
gaussian_force_adjoint adj_strength / adj_center = central differences of
                       <f, lam> in s / c; penalty_* = central differences of
                       sum over forces of sum f_I^2 (f_I: force I alone)
OracleLFM (whole step) tests/test_oracle_composition.py: the steady
                       Taylor-Green flow and a passive it carries (exact
                       characteristics) converge at second order at
                       reinitialisations and first order inside an interval
                       (2D and 3D); the adjoint's gap to central differences of
                       the oracle's own loss shrinks under refinement (2D first
                       order, 3D second); planted mistakes (dropped Jacobian,
                       flipped time steps, dropped or mis-weighted adjoint
                       sources) fail these pins.  Bit-level agreement with the
                       reference's Taichi trajectories stays unpinned (Taichi
                       cannot run here).
=====================  ==========================================================
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so with gcc (no GPU toolchain involved)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        for sfx in ("d", "f"):
            getattr(_LIB, f"oracle_mgpcg_{sfx}").restype = ctypes.c_int
    return _LIB


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (n <= 0: unchanged);
    returns the count in use."""
    f = lib().oracle_set_threads
    f.restype = ctypes.c_int
    return int(f(ctypes.c_int(int(n))))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "f"


def _real(dtype, x):
    return ctypes.c_double(x) if np.dtype(dtype) == np.float64 else ctypes.c_float(x)


def _shape3(shape):
    s = list(shape) + [1] * (3 - len(shape))
    return np.ascontiguousarray(s, dtype=np.int32)


# ---------------------------------------------------------------- geometry --
def face_shape(n, a):
    return tuple(n[b] + (1 if b == a else 0) for b in range(len(n)))


def coords(shape, offs, dx, dtype=np.float64):
    """Positions (idx + offs) * dx of every node, shape `shape + (dim,)`
    (center/horizontal/vertical_coords_func, 2D/vortex/taichi_utils.py:137-150;
    x/y/z_coords_func, 3D/taichi_utils.py:95-128)."""
    axes = [(np.arange(s, dtype=np.float64) + o) * dx for s, o in zip(shape, offs)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.ascontiguousarray(np.stack(g, axis=-1).astype(dtype))


def cell_coords(n, dx, dtype=np.float64):
    return coords(n, [0.5] * len(n), dx, dtype)


def face_coords(n, a, dx, dtype=np.float64):
    return coords(face_shape(n, a), [0.0 if b == a else 0.5 for b in range(len(n))], dx, dtype)


def mac_bl(dim, a):
    return [0.0 if b == a else 0.5 for b in range(dim)]


# ------------------------------------------------------------ point kernels --
def interp(kern, f, bl, pts, dx, grad=False):
    """B-spline interp of f (kern 2/3/4) at pts; returns value[, gradient]."""
    dt = f.dtype
    dim = f.ndim
    f = np.ascontiguousarray(f)
    pts = np.ascontiguousarray(pts, dtype=dt)
    npts = pts.size // dim
    val = np.empty(pts.shape[:-1], dtype=dt)
    g = np.empty(pts.shape, dtype=dt) if grad else None
    blv = np.ascontiguousarray(list(bl) + [0.5] * (3 - dim), dtype=dt)
    getattr(lib(), f"oracle_interp_{_sfx(dt)}")(
        ctypes.c_int(kern), ctypes.c_int(dim), _p(f), _p(_shape3(f.shape)), _p(blv), _p(pts),
        ctypes.c_int64(npts), _real(dt, dx), _p(val), _p(g) if grad else None)
    return (val, g) if grad else val


def _ucomps(u):
    dt = u[0].dtype
    comps = [np.ascontiguousarray(c, dtype=dt) for c in u]
    ptrs = [_p(c) for c in comps] + [None] * (3 - len(comps))
    return comps, ptrs


def _cells(u):
    n = list(u[0].shape)
    n[0] -= 1
    return n


def mac_eval(u, pts, dx, kern=2, grad=False):
    """Sample the MAC velocity u (list of face fields) at pts:
    value (..., d) and optionally J (..., d, d) with J[a][b] = d u_a / d x_b."""
    dim = len(u)
    vals, grads = [], []
    for a in range(dim):
        r = interp(kern, u[a], mac_bl(dim, a), pts, dx, grad=grad)
        if grad:
            vals.append(r[0])
            grads.append(r[1])
        else:
            vals.append(r)
    v = np.stack(vals, axis=-1)
    if grad:
        return v, np.stack(grads, axis=-2)
    return v


def march(u, psi, T, dx, dt, order):
    """In-place RK march of points psi (..., d) and Jacobian columns T."""
    dtp = u[0].dtype
    dim = len(u)
    comps, ptrs = _ucomps(u)
    assert psi.flags.c_contiguous and psi.dtype == dtp
    if T is not None:
        assert T.flags.c_contiguous and T.dtype == dtp
    getattr(lib(), f"oracle_march_{_sfx(dtp)}")(
        ctypes.c_int(dim), ctypes.c_int(order), _p(_shape3(_cells(u))), _real(dtp, dx), _real(dtp, dt),
        *ptrs, _p(psi), _p(T) if T is not None else None, ctypes.c_int64(psi.size // dim))


def pullback(u0, psi, T, dx, kern=2):
    """T . u0(psi) per point (u0 sampled with the B-spline of order kern)."""
    dtp = u0[0].dtype
    dim = len(u0)
    comps, ptrs = _ucomps(u0)
    psi = np.ascontiguousarray(psi, dtype=dtp)
    T = np.ascontiguousarray(T, dtype=dtp)
    out = np.empty(psi.shape[:-1], dtype=dtp)
    getattr(lib(), f"oracle_pullback_{_sfx(dtp)}")(
        ctypes.c_int(kern), ctypes.c_int(dim), _p(_shape3(_cells(u0))), _real(dtp, dx), *ptrs, _p(psi), _p(T),
        ctypes.c_int64(psi.size // dim), _p(out))
    return out


def sl_points(vel, src, src_bl, x, dx, dt, order, kern=2):
    """Back-trace x through vel over dt (RK order), sample src there with the
    B-spline of order kern."""
    dtp = vel[0].dtype
    dim = len(vel)
    comps, ptrs = _ucomps(vel)
    src = np.ascontiguousarray(src, dtype=dtp)
    x = np.ascontiguousarray(x, dtype=dtp)
    out = np.empty(x.shape[:-1], dtype=dtp)
    blv = np.ascontiguousarray(list(src_bl) + [0.5] * (3 - dim), dtype=dtp)
    getattr(lib(), f"oracle_sl_{_sfx(dtp)}")(
        ctypes.c_int(kern), ctypes.c_int(dim), ctypes.c_int(order), _p(_shape3(_cells(vel))), _real(dtp, dx), _real(dtp, dt),
        *ptrs, _p(src), _p(_shape3(src.shape)), _p(blv), _p(x), ctypes.c_int64(x.size // dim), _p(out))
    return out


# ---------------------------------------------------------- field kernels ---
def sl_advect_u(vel, src, dx, dt, order, kern=2):
    """advect_u_grid (2D/vortex/flowmap.py:59-69,233-291; 3D/flowmap.py:202-282):
    component a of src sampled at the back-traced face-a positions."""
    n = _cells(vel)
    dim = len(vel)
    return [sl_points(vel, src[a], mac_bl(dim, a), face_coords(n, a, dx, vel[0].dtype), dx, dt, order, kern)
            for a in range(dim)]


def sl_advect_s(vel, s, dx, dt, order, kern=2):
    """advect_scalar_grid (2D/vortex/flowmap.py:71-81,293-329)."""
    n = _cells(vel)
    return sl_points(vel, s, [0.5] * len(vel), cell_coords(n, dx, vel[0].dtype), dx, dt, order, kern)


class FlowMap:
    """psi (cells) plus psi_f / T_f per face axis: one Jacobian column per
    face grid (the reference's psi, psi_x.., T_x.. fields)."""

    def __init__(self, n, dx, dtype):
        self.n, self.dx, self.dtype = tuple(n), dx, dtype
        self.dim = len(n)
        self.reset()

    def reset(self):
        """reset_to_identity (2D/vortex/flowmap.py:343-358, 3D/flowmap.py:336-356)."""
        n, dx, dt = self.n, self.dx, self.dtype
        self.psi = cell_coords(n, dx, dt)
        self.psi_f = [face_coords(n, a, dx, dt) for a in range(self.dim)]
        self.T_f = []
        for a in range(self.dim):
            T = np.zeros(face_shape(n, a) + (self.dim,), dtype=dt)
            T[..., a] = 1.0
            self.T_f.append(T)

    def march(self, u, dt, order):
        """RK_grid_only_psi on psi + RK_grid on each (psi_a, T_a)
        (march_phi_grid / march_psi_grid, 2D/vortex/lfm_midpoint_simulator.py:355-358,
        2D/vortex/lfm_midpoint_diff_simulator.py:215-218)."""
        march(u, self.psi, None, self.dx, dt, order)
        for a in range(self.dim):
            march(u, self.psi_f[a], self.T_f[a], self.dx, dt, order)

    def copy(self):
        m = FlowMap.__new__(FlowMap)
        m.n, m.dx, m.dtype, m.dim = self.n, self.dx, self.dtype, self.dim
        m.psi = self.psi.copy()
        m.psi_f = [p.copy() for p in self.psi_f]
        m.T_f = [t.copy() for t in self.T_f]
        return m


def advect_u(u0, fmap, dx, kern=2):
    """Impulse pullback u1_a = T_a . u0(psi_a) (2D/vortex/flowmap.py:83-97)."""
    return [pullback(u0, fmap.psi_f[a], fmap.T_f[a], dx, kern) for a in range(len(u0))]


def advect_scalar(s0, psi, dx, kern=2):
    """s1 = s0(psi), quadratic in the reference (2D/vortex/flowmap.py:99-107)."""
    return interp(kern, s0, [0.5] * s0.ndim, psi, dx)


def clamp_extrema(c, ref):
    """clamp_u (2D/vortex/flowmap.py:332-341, 3D/flowmap.py:323-334): c clamped
    to [min, max] of ref at the 2*dim axis neighbours, indices clamped to the
    array (sample's clamp-to-edge)."""
    lo = np.full(ref.shape, np.inf)
    hi = np.full(ref.shape, -np.inf)
    for a in range(ref.ndim):
        for d in (-1, 1):
            nb = np.take(ref, np.clip(np.arange(ref.shape[a]) + d, 0, ref.shape[a] - 1), axis=a)
            lo = np.minimum(lo, nb)
            hi = np.maximum(hi, nb)
    return np.minimum(np.maximum(c, lo), hi).astype(ref.dtype)


def bfecc(init, fwd, bwd, dx, clamp=False, kern=2):
    """BFECC (2D/vortex/flowmap.py:440-463):
    u = T.init(psi); err = (F.u(phi) - init)/2; u -= T.err(psi),
    where (T, psi) = `fwd` and (F, phi) = `bwd`; with BFECC_clamp (clamp)
    the result is clamp_u'd against the first pullback."""
    u = advect_u(init, fwd, dx, kern)
    e = advect_u(u, bwd, dx, kern)
    err = [(e[a] - init[a]) * 0.5 for a in range(len(init))]
    t = advect_u(err, fwd, dx, kern)
    out = [u[a] - t[a] for a in range(len(init))]
    return [clamp_extrema(o, r) for o, r in zip(out, u)] if clamp else out


def bfecc_scalar(init, psi, phi, dx, clamp=False, kern=2):
    """BFECC_scalar (2D/vortex/flowmap.py:465-481), clamp_u'd with BFECC_clamp."""
    s = advect_scalar(init, psi, dx, kern)
    e = advect_scalar(s, phi, dx, kern)
    err = (e - init) * 0.5
    out = s - advect_scalar(err, psi, dx, kern)
    return clamp_extrema(out, s) if clamp else out


def accumulate(f, init, fmap, dx, dt):
    """accumulate_init: init_a += (F_a . f(phi_a)) dt (2D/vortex/flowmap.py:508-520)."""
    g = advect_u(f, fmap, dx)
    for a in range(len(init)):
        init[a] += g[a] * dt


def accumulate_ab3(f, init, fmap, H1, H2, ind, dx, dt):
    """accumulate_init_new, AB3 (23,-16,5)/12 with the ind==1 quirk (q1)
    that re-uses the current value (2D/vortex/flowmap.py:523-553)."""
    g = advect_u(f, fmap, dx)
    for a in range(len(init)):
        if ind == 0:
            init[a] += g[a] * dt
        elif ind == 1:
            init[a] += (23 * g[a] - 16 * g[a] + 5 * H1[a]) / 12 * dt
        else:
            init[a] += (23 * g[a] - 16 * H1[a] + 5 * H2[a]) / 12 * dt
        H2[a] = H1[a].copy()
        H1[a] = g[a].copy()


# ------------------------------------------------------ control forces --
def _force_weights(n, a, dx, c, r, fp32_test):
    """(mask, e, p - c) of one force at the nodes of face grid a.  The test
    (m*m) r < 5 with m = |p - c| is evaluated in FP32, as the reference's FP32
    fields evaluate it, when fp32_test (else in FP64); e = exp(-|p - c|^2 r)
    in FP64."""
    if fp32_test:
        # FP32 node positions as FP32 fields hold them: float(idx + off) * float(dx)
        shp = face_shape(n, a)
        axes = [(np.arange(m) + (0.0 if b == a else 0.5)).astype(np.float32) * np.float32(dx)
                for b, m in enumerate(shp)]
        p = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
    else:
        p = face_coords(n, a, dx)
    if fp32_test:
        d32 = p - np.asarray(c, np.float32)
        n2 = np.zeros(d32.shape[:-1], np.float32)
        for b in range(len(n)):
            n2 = n2 + d32[..., b] * d32[..., b]
        m = np.sqrt(n2)
        inside = (m * m) * np.float32(r) < np.float32(5.0)
    d = p.astype(np.float64) - np.asarray(c, np.float64)
    q = np.sum(d * d, axis=-1) * float(r)
    if not fp32_test:
        inside = q < 5.0
    return inside, np.exp(-q), d


def gaussian_force(n, dx, center, inv_radius, strength, count, fp32_test=True):
    """apply_gausion_force_kernel (2D/vortex/lfm_midpoint_simulator.py:309-326;
    3D/lfm_midpoint_simulator.py:316-344): on every face of axis a,
    f_a = sum over the first `count` forces I with (m*m) r_I < 5 of
    exp(-m^2 r_I) s_I[a], m = |p - c_I|; zero elsewhere."""
    dim = len(n)
    f = [np.zeros(face_shape(n, a)) for a in range(dim)]
    for I in range(int(count)):
        for a in range(dim):
            inside, e, _ = _force_weights(n, a, dx, center[I], inv_radius[I], fp32_test)
            f[a] += np.where(inside, e * float(strength[I][a]), 0.0)
    return f


def gaussian_force_adjoint(n, dx, center, inv_radius, strength, count, lam, fp32_test=True):
    """calculate_adj_gausion_force_kernel (2D/shape/lfm_midpoint_diff_simulator.py:75-103;
    3D/lfm_midpoint_diff_simulator.py:70-110): per force I, over the faces
    inside its cut-off,
      adj_strength[a] = sum e lam_a,   adj_center = sum s[a] e 2 (p - c) r lam_a,
      pen_strength[a] = sum 2 (e s[a]) e,   pen_center = sum 2 (e s[a]) s[a] e 2 (p - c) r.
    Returns (adj_center, adj_strength, pen_center, pen_strength), (count, dim) each."""
    dim = len(n)
    cnt = int(count)
    ac, ast, pc, ps = (np.zeros((cnt, dim)) for _ in range(4))
    for I in range(cnt):
        r = float(inv_radius[I])
        for a in range(dim):
            inside, e, d = _force_weights(n, a, dx, center[I], r, fp32_test)
            e = np.where(inside, e, 0.0)
            sa = float(strength[I][a])
            la = np.asarray(lam[a], np.float64)
            ast[I, a] += np.sum(e * la)
            ps[I, a] += np.sum(2.0 * (e * sa) * e)
            for b in range(dim):
                g = e * 2.0 * d[..., b] * r
                ac[I, b] += np.sum(sa * g * la)
                pc[I, b] += np.sum(2.0 * (e * sa) * sa * g)
    return ac, ast, pc, ps


def adj_conv_source(u, w, dx, kern):
    """calculate_nabla_u_w: r_a = ((grad u)^T w)_a at face-a nodes, with the
    quartic (2D, kern=4) or cubic (3D, kern=3) kernel
    (2D/vortex/flowmap.py:361-377, 3D/flowmap.py:359-380)."""
    dim = len(u)
    n = _cells(u)
    out = []
    for a in range(dim):
        X = face_coords(n, a, dx, u[0].dtype)
        r = np.zeros(X.shape[:-1], dtype=u[0].dtype)
        for c in range(dim):
            _, g = interp(kern, u[c], mac_bl(dim, c), X, dx, grad=True)
            wv = interp(kern, w[c], mac_bl(dim, c), X, dx)
            r += g[..., a] * wv
        out.append(r)
    return out


def adj_scalar_source(s, lam, dx):
    """calculate_nabla_scalar_adjoint: r_a = d_a s(X_a) * lam(X_a), cubic
    (2D/vortex/flowmap.py:379-395, 3D/flowmap.py:382-403)."""
    dim = s.ndim
    n = list(s.shape)
    out = []
    for a in range(dim):
        X = face_coords(n, a, dx, s.dtype)
        _, g = interp(3, s, [0.5] * dim, X, dx, grad=True)
        lv = interp(3, lam, [0.5] * dim, X, dx)
        out.append(g[..., a] * lv)
    return out


# ---------------------------------------------------------------- Poisson ---
def apply_bc(u):
    """Zero the normal velocity on the box faces (apply_bc, 2D/vortex/mgpcg.py:294-326)."""
    for a, c in enumerate(u):
        sl0 = [slice(None)] * c.ndim
        sl1 = [slice(None)] * c.ndim
        sl0[a] = 0
        sl1[a] = -1
        c[tuple(sl0)] = 0
        c[tuple(sl1)] = 0


def divergence(u):
    """div without 1/dx (2D/vortex/mgpcg.py:328-336, 3D/mgpcg.py:401-411)."""
    d = None
    for a, c in enumerate(u):
        hi = [slice(None)] * c.ndim
        lo = [slice(None)] * c.ndim
        hi[a] = slice(1, None)
        lo[a] = slice(0, -1)
        t = c[tuple(hi)] - c[tuple(lo)]
        d = t if d is None else d + t
    return d


def subtract_grad_p(u, p):
    """u_a -= p[I] - p[I - e_a] on interior faces (2D/vortex/mgpcg.py:338-356);
    boundary faces are zeroed by the following apply_bc."""
    for a, c in enumerate(u):
        hi = [slice(None)] * c.ndim
        lo = [slice(None)] * c.ndim
        mid = [slice(None)] * c.ndim
        hi[a] = slice(1, None)
        lo[a] = slice(0, -1)
        mid[a] = slice(1, -1)
        c[tuple(mid)] -= (p[tuple(hi)] - p[tuple(lo)]).astype(c.dtype)


def neumann_solve_dct(b):
    """Exact mean-zero solution of A p = b, A the all-Neumann (#nbrs)I - adj
    Laplacian on cells: A is diagonalised by the DCT-II, eigenvalues
    sum_d (2 - 2 cos(pi k_d / n_d))."""
    from scipy.fft import dctn, idctn

    b = np.asarray(b, dtype=np.float64)
    bh = dctn(b, type=2, norm="ortho")
    lam = np.zeros(b.shape)
    for a, n in enumerate(b.shape):
        k = np.arange(n)
        l1 = 2.0 - 2.0 * np.cos(np.pi * k / n)
        shp = [1] * b.ndim
        shp[a] = n
        lam = lam + l1.reshape(shp)
    lam.flat[0] = np.inf
    return idctn(bh / lam, type=2, norm="ortho")


def project_exact(u):
    """Poisson projection (MGPCG_2.Poisson, 2D/vortex/mgpcg.py:363-368) with the
    pressure solve done exactly; u modified in place."""
    apply_bc(u)
    b = -divergence(u).astype(np.float64)
    b -= b.mean()
    p = neumann_solve_dct(b)
    subtract_grad_p(u, p)
    apply_bc(u)


# ------------------------------------------------------ obstacle masks -------
def _face_solid(solid, a):
    """Boolean per face of axis a: true where either adjacent cell is solid."""
    n = solid.shape
    out = np.zeros(face_shape(n, a), dtype=bool)
    hi = [slice(None)] * solid.ndim
    lo = [slice(None)] * solid.ndim
    hi[a] = slice(0, n[a])
    lo[a] = slice(1, n[a] + 1)
    out[tuple(hi)] |= solid
    out[tuple(lo)] |= solid
    return out


def apply_bc_masked(u, solid):
    """apply_bc of the obstacle solver (2D/vortex/mgpcg_solid.py:388-402): box
    walls and every face of a solid cell are zeroed."""
    apply_bc(u)
    for a, c in enumerate(u):
        c[_face_solid(solid, a)] = 0


def masked_laplacian(solid):
    """Sparse fluid-cell operator of the obstacle solver: row i = (#fluid
    neighbours) x_i - sum of fluid neighbours (num_fluid_neighbors2 /
    neighbor_sum2, 2D/vortex/mgpcg_solid.py:140-170); solid rows and columns
    are left out.  Returns (A, fluid flat indices)."""
    import scipy.sparse as sp

    n = solid.shape
    fluid = ~solid
    idx = -np.ones(n, dtype=np.int64)
    fl = np.flatnonzero(fluid.ravel())
    idx.ravel()[fl] = np.arange(fl.size)
    rows, cols, vals = [], [], []
    diag = np.zeros(fl.size)
    for a in range(solid.ndim):
        lo = [slice(None)] * solid.ndim
        hi = [slice(None)] * solid.ndim
        lo[a] = slice(0, n[a] - 1)
        hi[a] = slice(1, n[a])
        i0, i1 = idx[tuple(lo)], idx[tuple(hi)]
        both = (i0 >= 0) & (i1 >= 0)
        r, c = i0[both], i1[both]
        rows += [r, c]
        cols += [c, r]
        vals += [-np.ones(r.size), -np.ones(r.size)]
        np.add.at(diag, r, 1.0)
        np.add.at(diag, c, 1.0)
    rows.append(np.arange(fl.size))
    cols.append(np.arange(fl.size))
    vals.append(diag)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(fl.size, fl.size))
    return A, fl


def project_masked_exact(u, solid):
    """MGPCG_2_solid.Poisson (2D/vortex/mgpcg_solid.py:455-461) with the
    pressure solved exactly: apply_bc, divergence on fluid cells (0 in solid
    ones), the fluid-cell Poisson problem solved directly (its constant null
    space fixed by the mean-zero condition; the fluid region must be
    connected), p = 0 in solid cells, gradient subtraction, apply_bc."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    solid = np.asarray(solid, dtype=bool)
    apply_bc_masked(u, solid)
    b = -divergence(u).astype(np.float64)
    b[solid] = 0.0
    A, fl = masked_laplacian(solid)
    bf = b.ravel()[fl]
    bf = bf - bf.mean()
    # ground the first fluid cell (x_0 = 0): the remaining block of the
    # singular Neumann operator is SPD on a connected fluid region, and the
    # consistent right-hand side makes row 0 hold automatically; then recentre
    x = np.zeros(fl.size)
    x[1:] = spla.spsolve(sp.csc_matrix(A[1:, 1:]), bf[1:])
    x -= x.mean()
    p = np.zeros(solid.shape)
    p.ravel()[fl] = x
    subtract_grad_p(u, p)
    apply_bc_masked(u, solid)
    return p


def mask_velocity(u, mask, bvel=None):
    """mask_velocity_by_boundary / mask_adj_velocity_by_boundary
    (2D/vortex/init_conditions.py:548-556,571-578): every face of a cell with
    mask >= 1 takes that cell's boundary velocity (0 when bvel is None); a
    face between two such cells takes the high-side cell's value."""
    solid = np.asarray(mask) >= 1
    n = solid.shape
    for a, c in enumerate(u):
        val = np.zeros(n) if bvel is None else np.asarray(bvel[a], dtype=np.float64)
        lo = [slice(None)] * solid.ndim
        hi = [slice(None)] * solid.ndim
        lo[a] = slice(1, n[a] + 1)   # faces whose low-side cell is I - e_a
        hi[a] = slice(0, n[a])       # faces whose high-side cell is I
        v = c[tuple(lo)]
        v[solid] = val[solid]
        c[tuple(lo)] = v
        v = c[tuple(hi)]
        v[solid] = val[solid]
        c[tuple(hi)] = v


def mask_scalar(s, mask, adjoint=False):
    """mask_passive_by_boundary (solid cells take the mean of the fluid cells of
    their 3^d neighbourhood, accumulated in the reference's loop order) and
    mask_adj_passive_by_boundary (solid cells zeroed)
    (2D/vortex/init_conditions.py:558-569,580-585; 3D/init_conditions.py:547-578)."""
    mask = np.asarray(mask)
    out = s.copy()
    for I in zip(*np.nonzero(mask >= 1)):
        if adjoint:
            out[I] = 0.0
            continue
        acc, w = s.dtype.type(0), 0
        for off in np.ndindex(*(3,) * s.ndim):
            J = tuple(i + o - 1 for i, o in zip(I, off))
            if all(0 <= j < m for j, m in zip(J, s.shape)) and mask[J] <= 0:
                acc = acc + s[J]
                w += 1
        out[I] = acc / w if w > 0 else 0.0
    s[...] = out


def mgpcg_solve(b, levels, pre=2, bottom=50, tol=1e-12, relative=False, max_iters=400):
    """Restated MGPCG solve of A x = b; returns (x, iters, rel_residual, status)."""
    b = np.ascontiguousarray(b)
    dtp = b.dtype
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    res = ctypes.c_double(0)
    st = getattr(lib(), f"oracle_mgpcg_{_sfx(dtp)}")(
        ctypes.c_int(b.ndim), _p(_shape3(b.shape)), ctypes.c_int(levels), ctypes.c_int(pre),
        ctypes.c_int(bottom), _p(b), _p(x), ctypes.c_int(max_iters), ctypes.c_double(tol),
        ctypes.c_int(1 if relative else 0), ctypes.byref(it), ctypes.byref(res))
    return x, it.value, res.value, st


def default_levels(n, base_level=3):
    """n_mg_levels = log2(min N) - base_level + 1 (2D/vortex/mgpcg.py:28)."""
    import math

    return int(math.log2(min(n))) - base_level + 1


def project_mgpcg(u, levels=None, pre=2, bottom=50, tol=1e-6, relative=True, max_iters=400):
    """Poisson projection with the restated MGPCG pressure solve."""
    n = _cells(u)
    apply_bc(u)
    b = -divergence(u)
    x, it, res, st = mgpcg_solve(b, levels or default_levels(n), pre, bottom, tol, relative, max_iters)
    subtract_grad_p(u, x)
    apply_bc(u)
    return it, res


# -------------------------------------------------------------- simulator ---
class OracleLFM:
    """Forward LFM midpoint step and its adjoint sweep, restating
    LFM_Simulator.step_flow_map (2D/vortex/lfm_midpoint_simulator.py:472-661;
    3D/lfm_midpoint_simulator.py:648-807) and
    LFM_Diff_Simulator.init_gradient / backtrack_flow_map
    (2D/vortex/lfm_midpoint_diff_simulator.py:126-503;
    3D/lfm_midpoint_diff_simulator.py:336-594) for the configuration with
    viscosity 0, backward_u=True and a final-time density loss (the
    per-interval intermediate-target terms, q7, are off); BFECC_clamp and
    use_short_BFECC are constructor flags.

    Storage follows the reference's disk cache semantics, kept in RAM:
    midpoint velocity per step, passive at the start of every step, forward
    map (phi, F) per interval.
    """

    def __init__(self, n, dx, dt, reinit_every, order, adj_kernel=None, project="exact", dtype=np.float64,
                 mg_levels=None, mg_tol=1e-6, control_forces=False, bfecc_clamp=False, short_bfecc=False,
                 interp_kern=2):
        self.control_forces = control_forces  # add_control_force: also switches the j == 0 passive rebuild
        # B-spline order sampling the advected fields in the BFECC pullbacks and
        # the semi-Lagrangian advection (2 = the reference's; 1 = linear N_1,
        # 3 = cubic N_3 as optional variants); the map march, the back-trace
        # velocity and accumulate_init stay quadratic
        self.kern = interp_kern
        self.clamp = bfecc_clamp               # BFECC_clamp
        self.short = short_bfecc               # use_short_BFECC
        self.n = tuple(n)
        self.dim = len(n)
        self.dx, self.dt, self.R, self.order = dx, dt, reinit_every, order
        self.adj_kernel = adj_kernel or (4 if self.dim == 2 else 3)
        self.dtype = dtype
        self.project_mode = project
        self.mg_levels = mg_levels
        self.mg_tol = mg_tol
        self.phi_map = FlowMap(n, dx, dtype)
        self.psi_map = FlowMap(n, dx, dtype)

    mask = None
    bvel = None
    interval_targets = None  # {interval-start step: [target per passive]}
    on_stage = None          # optional callable(name): told as each projection of a step completes

    def _stage(self, name):
        if self.on_stage is not None:
            self.on_stage(name)

    def set_mask(self, mask, bvel=None):
        """Obstacle mask on cells (solid where > 0 for the projection, >= 1 for
        the field masks), as the 2D reference sets boundary_mask; the
        projection becomes MGPCG_2_solid's (solved exactly) and the mask_*
        kernels run where the reference calls them."""
        self.mask = None if mask is None else np.asarray(mask, dtype=np.float64)
        self.bvel = bvel

    def project(self, u):
        if self.mask is not None:
            project_masked_exact(u, self.mask > 0)
        elif self.project_mode == "exact":
            project_exact(u)
        else:
            project_mgpcg(u, levels=self.mg_levels, tol=self.mg_tol)
        self._stage("projection")

    def _mask_fwd(self):
        """mask_velocity_by_boundary + mask_passive_by_boundary after a forward
        projection (2D/vortex/lfm_midpoint_simulator.py:429-433,543-547,612-616,632-636)."""
        if self.mask is None:
            return
        mask_velocity(self.u, self.mask, self.bvel)
        for s in self.s:
            mask_scalar(s, self.mask)

    def _mask_adj(self):
        """The forward masks on u and s plus mask_adj_velocity_by_boundary and
        mask_adj_passive_by_boundary on the adjoint fields
        (2D/vortex/lfm_midpoint_diff_simulator.py:179-189,405-413,451-459)."""
        if self.mask is None:
            return
        mask_velocity(self.u, self.mask, self.bvel)
        mask_velocity(self.lam_u, self.mask, None)
        for s in self.s:
            mask_scalar(s, self.mask)
        for l in self.lam_s:
            mask_scalar(l, self.mask, adjoint=True)

    # forward ---------------------------------------------------------------
    def init(self, u0, passives):
        """LFM_Simulator.init (2D/vortex/lfm_midpoint_simulator.py:370-439) after
        the initial field is built: project, reset phi/F, init_u = u."""
        dt = self.dtype
        self.u = [np.array(c, dtype=dt) for c in u0]
        self.project(self.u)
        self.s = [np.array(s, dtype=dt) for s in passives]
        self.init_s = [s.copy() for s in self.s]
        self._mask_fwd()
        self.init_u = [c.copy() for c in self.u]
        self.phi_map.reset()
        self.step_num = -1
        self.mid = {}
        self.s_start = {}
        self.maps = {}

    def _force(self, force):
        return gaussian_force(self.n, self.dx, force["c"], force["r"], force["s"], force["n"])

    def step(self, force=None):
        """step_midpoint + step_flow_map (2D/vortex/lfm_midpoint_simulator.py:472-661).
        force: control_force_para {"c", "r", "s", "n"} of this step (the
        half-step force before the first projection, :527-537; the full-step
        force and its accumulate_init after advection, :597-602)."""
        self.step_num += 1
        j = self.step_num % self.R
        dx, dt, order = self.dx, self.dt, self.order
        self.s_start[self.step_num] = [s.copy() for s in self.s]
        ad_u = [c.copy() for c in self.u]
        self.u = sl_advect_u(ad_u, ad_u, dx, 0.5 * dt, order, self.kern)
        f = None
        if force is not None:
            f = self._force(force)
            for a in range(self.dim):
                self.u[a] = self.u[a] + 0.5 * dt * f[a]
        self.project(self.u)
        self._mask_fwd()
        self.mid[self.step_num] = [c.copy() for c in self.u]
        self.phi_map.march(self.u, -dt, order)
        if self.short:
            # one-step maps through the midpoint velocity and BFECC from the
            # step's start (use_short_BFECC, 2D/vortex/lfm_midpoint_simulator.py:560-583)
            psi_t, phi_t = FlowMap(self.n, dx, self.dtype), FlowMap(self.n, dx, self.dtype)
            psi_t.march(self.u, dt, order)
            phi_t.march(self.u, -dt, order)
            self.s = [bfecc_scalar(s, psi_t.psi, phi_t.psi, dx, self.clamp, self.kern) for s in self.s]
            self.u = bfecc(ad_u, psi_t, phi_t, dx, self.clamp, self.kern)
        else:
            self.s = [sl_advect_s(self.u, s, dx, dt, order, self.kern) for s in self.s]
            self.u = sl_advect_u(self.u, ad_u, dx, dt, order, self.kern)
        if f is not None:
            for a in range(self.dim):
                self.u[a] = self.u[a] + dt * f[a]
            accumulate(f, self.init_u, self.phi_map, dx, dt)
        self.project(self.u)
        self._mask_fwd()
        if j == self.R - 1:
            # backtrack_psi_grid: last step with the end-of-step u (q3)
            self.psi_map.reset()
            self.psi_map.march(self.u, dt, order)
            for st in reversed(range(self.step_num - j, self.step_num)):
                self.psi_map.march(self.mid[st], dt, order)
            self.u = bfecc(self.init_u, self.psi_map, self.phi_map, dx, self.clamp, self.kern)
            self.project(self.u)
            self._mask_fwd()
            self.s = [bfecc_scalar(si, self.psi_map.psi, self.phi_map.psi, dx, self.clamp, self.kern) for si in self.init_s]
            self.init_s = [s.copy() for s in self.s]
            self.maps[self.step_num // self.R] = self.phi_map.copy()
            self.phi_map.reset()
            self.init_u = [c.copy() for c in self.u]

    # adjoint ---------------------------------------------------------------
    def init_gradient(self, targets):
        """LFM_Diff_Simulator.init_gradient (2D/vortex/lfm_midpoint_diff_simulator.py:126-207):
        lam_s = 2 (s - s*), lam_u = 0, loss = sum (s - s*)^2."""
        self.final_ind = self.step_num + 1
        self.lam_u = [np.zeros_like(c) for c in self.u]
        self.lam_s = [2.0 * (s - t.astype(self.dtype)) for s, t in zip(self.s, targets)]
        self.loss = float(sum(np.sum((s.astype(np.float64) - t) ** 2) for s, t in zip(self.s, targets)))
        self._mask_adj()
        self.step_num = self.final_ind
        self.init_lam_u = [c.copy() for c in self.lam_u]
        self.init_lam_s = [c.copy() for c in self.lam_s]
        self.init_u = [c.copy() for c in self.u]
        self.init_s = [s.copy() for s in self.s]
        self.psi_map.reset()
        self.H1 = [np.zeros_like(c) for c in self.u]
        self.H2 = [np.zeros_like(c) for c in self.u]

    def backtrack_step(self, force=None):
        """backtrack_step_midpoint + backtrack_flow_map
        (2D/vortex/lfm_midpoint_diff_simulator.py:220-503).  force: the
        step's control_force_para; its parameter gradients (from lam_u as the
        step begins) land in self.force_grad, and u takes the force back
        (2D/shape/lfm_midpoint_diff_simulator.py:209-213,296-300)."""
        self.step_num -= 1
        j = self.step_num % self.R
        dx, dt, order = self.dx, self.dt, self.order
        mid = self.mid[self.step_num]
        f = None
        if force is not None:
            f = self._force(force)
            self.force_grad = gaussian_force_adjoint(self.n, dx, force["c"], force["r"], force["s"], force["n"],
                                                     self.lam_u)
        r_u = adj_conv_source(self.u, self.lam_u, dx, self.adj_kernel)
        r_s = [adj_scalar_source(s, l, dx) for s, l in zip(self.s, self.lam_s)]
        if self.short:
            # use_short_BFECC (2D/vortex/lfm_midpoint_diff_simulator.py:283-342,
            # 3D/lfm_midpoint_diff_simulator.py:440-496): one-step
            # maps through mid, BFECC of lam_u, lam_s, s and u, and the
            # convective source at -2 dt
            psi_t, phi_t = FlowMap(self.n, dx, self.dtype), FlowMap(self.n, dx, self.dtype)
            psi_t.march(mid, -dt, order)
            phi_t.march(mid, dt, order)
            lam = bfecc(self.lam_u, psi_t, phi_t, dx, self.clamp, self.kern)
            self.lam_u = [lam[a] + (-2 * dt) * r_u[a] for a in range(self.dim)]
            self.lam_s = [bfecc_scalar(l, psi_t.psi, phi_t.psi, dx, self.clamp, self.kern) for l in self.lam_s]
            for rs in r_s:
                for a in range(self.dim):
                    self.lam_u[a] = self.lam_u[a] + (-dt) * rs[a]
            self.s = [bfecc_scalar(s, psi_t.psi, phi_t.psi, dx, self.clamp, self.kern) for s in self.s]
            self.u = bfecc(self.u, psi_t, phi_t, dx, self.clamp, self.kern)
        else:
            self.lam_u = sl_advect_u(mid, self.lam_u, dx, -dt, order, self.kern)
            self.s = [sl_advect_s(mid, s, dx, -dt, order, self.kern) for s in self.s]
            self.u = sl_advect_u(mid, self.u, dx, -dt, order, self.kern)
            for a in range(self.dim):
                self.lam_u[a] = self.lam_u[a] + (-dt) * r_u[a]
            new_lam_s = []
            for l, rs in zip(self.lam_s, r_s):
                new_lam_s.append(sl_advect_s(mid, l, dx, -dt, order, self.kern))
                for a in range(self.dim):
                    self.lam_u[a] = self.lam_u[a] + (-dt) * rs[a]
            self.lam_s = new_lam_s
        self.psi_map.march(mid, dt, order)
        accumulate_ab3(r_u, self.init_lam_u, self.psi_map, self.H1, self.H2, self.R - 1 - j, dx, -2 * dt)
        for rs in r_s:
            accumulate(rs, self.init_lam_u, self.psi_map, dx, -dt)
        self.project(self.lam_u)
        if f is not None:
            for a in range(self.dim):
                self.u[a] = self.u[a] - dt * f[a]
            accumulate(f, self.init_u, self.psi_map, dx, -dt)
        self.project(self.u)
        self._mask_adj()
        if j == 0:
            fwd = self.maps[self.step_num // self.R]
            self.lam_u = bfecc(self.init_lam_u, fwd, self.psi_map, dx, self.clamp, self.kern)
            self.project(self.lam_u)
            self.init_lam_u = [c.copy() for c in self.lam_u]
            self.u = bfecc(self.init_u, fwd, self.psi_map, dx, self.clamp, self.kern)
            self.project(self.u)
            self.init_u = [c.copy() for c in self.u]
            self._mask_adj()
            if self.control_forces:
                # rebuilt through the interval's maps (2D/shape/lfm_midpoint_diff_simulator.py:339-343)
                self.s = [bfecc_scalar(si, fwd.psi, self.psi_map.psi, dx, self.clamp, self.kern) for si in self.init_s]
            else:
                self.s = [s.copy() for s in self.s_start[self.step_num]]
            self.init_s = [s.copy() for s in self.s]
            self.lam_s = [bfecc_scalar(l, fwd.psi, self.psi_map.psi, dx, self.clamp, self.kern) for l in self.init_lam_s]
            tg = self.interval_targets.get(self.step_num) if self.interval_targets else None
            if tg is not None:
                # intermediate targets (q7): lam_s += s - s*, loss += sum (s* - s)^2
                # (2D/vortex/lfm_midpoint_diff_simulator.py:478-497)
                for k, t in enumerate(tg):
                    d = self.s[k] - np.asarray(t, np.float64)
                    self.lam_s[k] = self.lam_s[k] + d
                    self.loss += float(np.sum(d ** 2))
            self.init_lam_s = [l.copy() for l in self.lam_s]
            self.psi_map.reset()
        return self.step_num == 0
