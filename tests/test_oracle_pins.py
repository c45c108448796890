"""Pins of the CPU oracle against mathematics and independent library
routines (CPU only).  Each test is chosen so that a plausible restatement
mistake -- a wrong tap window, sign, index offset, clamp or coefficient --
fails it; none re-types the oracle's own formula."""
import ctypes

import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import lfm_oracle as O

KNOTS = {2: [-1.5, -0.5, 0.5, 1.5], 3: [-2, -1, 0, 1, 2], 4: [-2.5, -1.5, -0.5, 0.5, 1.5, 2.5]}
SUPPORT = {2: 1.5, 3: 2.0, 4: 2.5}


def bspline(k, x, deriv=0):
    f = O.lib().oracle_bspline_d
    f.restype = ctypes.c_double
    return np.array([f(k, deriv, ctypes.c_double(v)) for v in np.atleast_1d(x)])


def scipy_basis(k, x, deriv=0):
    b = BSpline.basis_element(KNOTS[k], extrapolate=False)
    if deriv:
        b = b.derivative()
    v = b(np.atleast_1d(x))
    return np.nan_to_num(v)


# --------------------------------------------------------------- B-splines --
@pytest.mark.parametrize("k", [2, 3, 4])
def test_bspline_matches_scipy_basis_element(k):
    x = np.linspace(-3.2, 3.2, 1001)
    np.testing.assert_allclose(bspline(k, x), scipy_basis(k, x), atol=1e-14)
    # derivative away from the knots (where scipy's one-sided value is ambiguous)
    xs = x[np.min(np.abs(x[:, None] - np.array(KNOTS[k])[None]), axis=1) > 1e-6]
    np.testing.assert_allclose(bspline(k, xs, 1), scipy_basis(k, xs, 1), atol=1e-12)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_bspline_partition_of_unity_and_moments(k):
    f = np.linspace(0, 1, 37, endpoint=False)
    i = np.arange(-4, 5)
    W = bspline(k, (f[:, None] - i[None]).ravel()).reshape(len(f), len(i))
    dW = bspline(k, (f[:, None] - i[None]).ravel(), 1).reshape(len(f), len(i))
    np.testing.assert_allclose(W.sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(dW.sum(1), 0.0, atol=1e-13)
    # sum_i i N(x - i) = x ; sum_i i^2 N(x - i) = x^2 + (k+1)/12 (B-spline variance)
    np.testing.assert_allclose((W * i).sum(1), f, atol=1e-13)
    np.testing.assert_allclose((W * i ** 2).sum(1), f ** 2 + (k + 1) / 12.0, atol=1e-12)


def test_bspline_node_values():
    np.testing.assert_allclose(bspline(2, [0, 1, -1, 2]), [0.75, 0.125, 0.125, 0])
    np.testing.assert_allclose(bspline(3, [0, 1, -1, 2]), [2 / 3, 1 / 6, 1 / 6, 0])
    np.testing.assert_allclose(bspline(4, [0, 1, 2, 3]), [115 / 192, 19 / 96, 1 / 384, 0])


@pytest.mark.parametrize("k", [2, 3, 4])
def test_bspline_derivative_is_fd_and_continuous(k):
    x = np.linspace(-3, 3, 601) + 1e-3
    h = 1e-6
    fd = (bspline(k, x + h) - bspline(k, x - h)) / (2 * h)
    np.testing.assert_allclose(bspline(k, x, 1), fd, atol=1e-7)
    for kn in KNOTS[k]:
        assert abs(bspline(k, kn - 1e-9)[0] - bspline(k, kn + 1e-9)[0]) < 1e-8


# ------------------------------------------------------------ interpolation --
def brute_interp(k, f, bl, p, dx):
    """Sum over every node of f[clamp(i)] * prod_d N(s_d - i_d) with scipy's
    basis, after the reference's coordinate clamp (no tap-window logic)."""
    dim = f.ndim
    s = np.array([np.clip(p[d] / dx - bl[d], 1, f.shape[d] - 2 - 1e-6) for d in range(dim)])
    fl = np.floor(s).astype(int)
    lo = {2: -1, 3: -2, 4: -3}[k]
    tot = 0.0
    import itertools

    for off in itertools.product(range(lo, -lo + 2), repeat=dim):
        idx = tuple(int(np.clip(fl[d] + off[d], 0, f.shape[d] - 1)) for d in range(dim))
        w = np.prod([scipy_basis(k, s[d] - (fl[d] + off[d]))[0] for d in range(dim)])
        tot += f[idx] * w
    return tot


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("dim", [2, 3])
def test_interp_brute_force_with_clamping(k, dim):
    rng = np.random.default_rng(k * 10 + dim)
    shape = (9, 7) if dim == 2 else (7, 6, 8)
    f = rng.standard_normal(shape)
    dx = 0.1
    bl = [0.5, 0.0, 0.5][:dim]
    # include points far outside the domain (exercise the clamp) and near edges
    pts = rng.uniform(-0.3, 1.2, (25, dim))
    val = O.interp(k, f, bl, pts, dx)
    ref = np.array([brute_interp(k, f, bl, p, dx) for p in pts])
    np.testing.assert_allclose(val, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_interp_reproduces_linear_fields_and_gradients(k):
    n = (20, 22, 18)
    dx = 1.0 / 20
    bl = [0.0, 0.5, 0.5]
    a = np.array([0.3, -1.2, 2.0])
    idx = np.stack(np.meshgrid(*[np.arange(m) for m in n], indexing="ij"), -1)
    pos = (idx + np.array(bl)) * dx
    f = pos @ a + 0.7
    rng = np.random.default_rng(1)
    # interior points (away from the clamp band by the tap reach)
    reach = {2: 2, 3: 3, 4: 4}[k]
    lo = (np.array(bl) + reach) * dx
    hi = (np.array(n) - 1 - reach + np.array(bl)) * dx
    pts = rng.uniform(lo, hi, (50, 3))
    val, g = O.interp(k, f, bl, pts, dx, grad=True)
    np.testing.assert_allclose(val, pts @ a + 0.7, atol=1e-12)
    np.testing.assert_allclose(g, np.broadcast_to(a, g.shape), atol=1e-10)


def test_mac_offsets():
    """u = (x, 2y, 3z) sampled on the MAC faces is reproduced with J = diag(1,2,3)."""
    n = (16, 16, 16)
    dx = 1 / 16
    u = [O.face_coords(n, a, dx)[..., a] * (a + 1) for a in range(3)]
    rng = np.random.default_rng(2)
    pts = rng.uniform(0.25, 0.75, (40, 3))
    v, J = O.mac_eval(u, pts, dx, grad=True)
    np.testing.assert_allclose(v, pts * np.array([1, 2, 3]), atol=1e-12)
    np.testing.assert_allclose(J, np.broadcast_to(np.diag([1.0, 2, 3]), J.shape), atol=1e-10)


# -------------------------------------------------------------------- march --
def linear_velocity(n, dx, A):
    return [O.face_coords(n, a, dx) @ A[a] for a in range(len(n))]


@pytest.mark.parametrize("order", [2, 4])
def test_march_uniform_flow_exact(order):
    n = (16, 16)
    dx = 1 / 16
    c = np.array([0.4, -0.7])
    u = [np.full(O.face_shape(n, a), c[a]) for a in range(2)]
    psi = O.face_coords(n, 0, dx)[4:12, 4:12].copy()
    T = np.zeros_like(psi)
    T[..., 0] = 1
    psi0 = psi.copy()
    O.march(u, psi, T, dx, 0.05, order)
    np.testing.assert_allclose(psi, psi0 - 0.05 * c, atol=1e-14)
    np.testing.assert_allclose(T[..., 0], 1.0, atol=1e-14)
    np.testing.assert_allclose(T[..., 1], 0.0, atol=1e-14)


@pytest.mark.parametrize("order,expect", [(2, 8.0), (4, 32.0)])
def test_march_linear_flow_convergence_order(order, expect):
    """d psi/dt = -A psi, dT/dt = -A T: one step vs expm, local error ~ dt^(order+1)."""
    from scipy.linalg import expm

    n = (32, 32, 32)
    dx = 1 / 32
    A = np.array([[0.0, 1.3, -0.4], [-1.3, 0.2, 0.9], [0.4, -0.9, -0.2]])
    u = linear_velocity(n, dx, A)
    p0 = np.array([[0.5, 0.45, 0.55], [0.52, 0.5, 0.47]])
    T0 = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    errs = []
    for dt in (0.04, 0.02):
        psi, T = p0.copy(), T0.copy()
        O.march(u, psi, T, dx, dt, order)
        E = expm(-A * dt)
        errs.append(max(np.abs(psi - p0 @ E.T).max(), np.abs(T - T0 @ E.T).max()))
    ratio = errs[0] / errs[1]
    assert expect * 0.75 < ratio < expect * 1.3, (errs, ratio)


def test_march_rotation_preserves_jacobian_determinant():
    n = (32, 32)
    dx = 1 / 32
    A = np.array([[0.0, -2.0], [2.0, 0.0]])
    u = linear_velocity(n, dx, A)
    p = np.array([[0.5, 0.5], [0.45, 0.55]])
    Tx = np.array([[1.0, 0.0], [1.0, 0.0]])
    Ty = np.array([[0.0, 1.0], [0.0, 1.0]])
    for _ in range(10):
        O.march(u, p.copy(), Tx, dx, 0.01, 4)
        O.march(u, p.copy(), Ty, dx, 0.01, 4)
    det = Tx[:, 0] * Ty[:, 1] - Tx[:, 1] * Ty[:, 0]
    np.testing.assert_allclose(det, 1.0, atol=1e-9)


# ------------------------------------------------------ pullback / SL / BFECC --
def test_pullback_identity_map_is_bspline_smoothing():
    """At T = e_a, psi = X_a the pullback samples u_a at its own nodes: the
    quadratic B-spline there is the (1/8, 3/4, 1/8) stencil per axis."""
    n = (16, 16)
    dx = 1 / 16
    rng = np.random.default_rng(3)
    u = [rng.standard_normal(O.face_shape(n, a)) for a in range(2)]
    fm = O.FlowMap(n, dx, np.float64)
    out = O.advect_u(u, fm, dx)
    w = np.array([0.125, 0.75, 0.125])
    K = np.outer(w, w)
    for a in range(2):
        f = u[a]
        ref = sum(K[i, j] * f[i:f.shape[0] - 2 + i, j:f.shape[1] - 2 + j] for i in range(3) for j in range(3))
        np.testing.assert_allclose(out[a][2:-2, 2:-2], ref[1:-1, 1:-1], atol=1e-13)


def test_pullback_affine_map_of_linear_field():
    """psi(x) = M x + b, T_a = column a of M, u linear: T_a . u(psi) exactly."""
    n = (24, 24)
    dx = 1 / 24
    A = np.array([[0.4, -1.0], [0.7, 0.2]])
    u = linear_velocity(n, dx, A)
    M = np.array([[0.9, 0.15], [-0.1, 1.05]])
    b = np.array([0.02, -0.03])
    fm = O.FlowMap(n, dx, np.float64)
    for a in range(2):
        X = O.face_coords(n, a, dx)
        fm.psi_f[a] = np.ascontiguousarray((X - 0.5) @ M.T + 0.5 + b)
        fm.T_f[a] = np.ascontiguousarray(np.broadcast_to(M[:, a], X.shape))
    out = O.advect_u(u, fm, dx)
    for a in range(2):
        X = O.face_coords(n, a, dx)
        P = (X - 0.5) @ M.T + 0.5 + b
        ref = (P @ A.T) @ M[:, a]
        sl = (slice(4, -4), slice(4, -4))
        np.testing.assert_allclose(out[a][sl], ref[sl], atol=1e-12)


@pytest.mark.parametrize("order", [2, 4])
def test_sl_advect_uniform_flow_shifts_linear_field(order):
    n = (20, 20, 20)
    dx = 1 / 20
    c = np.array([0.3, -0.2, 0.5])
    vel = [np.full(O.face_shape(n, a), c[a]) for a in range(3)]
    g = np.array([1.0, 2.0, -0.5])
    s = O.cell_coords(n, dx) @ g
    out = O.sl_advect_s(vel, s, dx, 0.04, order)
    ref = (O.cell_coords(n, dx) - 0.04 * c) @ g
    sl = (slice(4, -4),) * 3
    np.testing.assert_allclose(out[sl], ref[sl], atol=1e-12)


def test_bfecc_exact_inverse_maps_of_linear_field():
    """With an affine psi and its exact inverse phi, BFECC of a linear field is
    exact (the error estimate vanishes)."""
    n = (24, 24)
    dx = 1 / 24
    A = np.array([[0.4, -1.0], [0.7, 0.2]])
    u = linear_velocity(n, dx, A)
    M = np.array([[1.05, 0.1], [-0.05, 0.95]])
    Mi = np.linalg.inv(M)
    fwd = O.FlowMap(n, dx, np.float64)
    bwd = O.FlowMap(n, dx, np.float64)
    for a in range(2):
        X = O.face_coords(n, a, dx)
        fwd.psi_f[a] = np.ascontiguousarray((X - 0.5) @ M.T + 0.5)
        fwd.T_f[a] = np.ascontiguousarray(np.broadcast_to(M[:, a], X.shape))
        bwd.psi_f[a] = np.ascontiguousarray((X - 0.5) @ Mi.T + 0.5)
        bwd.T_f[a] = np.ascontiguousarray(np.broadcast_to(Mi[:, a], X.shape))
    once = O.advect_u(u, fwd, dx)
    out = O.bfecc(u, fwd, bwd, dx)
    sl = (slice(6, -6), slice(6, -6))
    for a in range(2):
        np.testing.assert_allclose(out[a][sl], once[a][sl], atol=1e-11)


def test_accumulate_ab3_weights_sum_to_one():
    """A constant source through an identity map adds dt per step whatever the
    AB3 branch (23 - 16 + 5 = 12; the ind==1 form too)."""
    n = (12, 12)
    dx = 1 / 12
    fm = O.FlowMap(n, dx, np.float64)
    f = [np.full(O.face_shape(n, a), 2.0) for a in range(2)]
    init = [np.zeros(O.face_shape(n, a)) for a in range(2)]
    H1 = [np.zeros_like(c) for c in init]
    H2 = [np.zeros_like(c) for c in init]
    for ind in range(4):
        O.accumulate_ab3(f, init, fm, H1, H2, ind, dx, 0.1)
    for a in range(2):
        np.testing.assert_allclose(init[a], 4 * 0.1 * 2.0, atol=1e-12)


# ------------------------------------------------------------- adj sources --
@pytest.mark.parametrize("kern", [3, 4])
def test_adj_conv_source_linear_u_constant_w(kern):
    n = (24, 24, 24)
    dx = 1 / 24
    A = np.array([[0.3, -0.8, 0.1], [0.5, 0.2, -0.6], [-0.4, 0.9, 0.0]])
    u = linear_velocity(n, dx, A)
    wv = np.array([0.7, -1.1, 0.4])
    w = [np.full(O.face_shape(n, a), wv[a]) for a in range(3)]
    r = O.adj_conv_source(u, w, dx, kern)
    expect = A.T @ wv  # r_a = sum_c dU_c/dx_a w_c
    sl = (slice(5, -5),) * 3
    for a in range(3):
        np.testing.assert_allclose(r[a][sl], expect[a], atol=1e-10)


def test_adj_scalar_source_linear_s():
    n = (20, 20)
    dx = 1 / 20
    g = np.array([1.5, -0.5])
    s = O.cell_coords(n, dx) @ g
    lam = O.cell_coords(n, dx)[..., 0] * 0 + 3.0
    r = O.adj_scalar_source(s, lam, dx)
    sl = (slice(5, -5), slice(5, -5))
    for a in range(2):
        np.testing.assert_allclose(r[a][sl], g[a] * 3.0, atol=1e-10)


# ------------------------------------------------------------------ Poisson --
def neumann_matrix(shape):
    import scipy.sparse as sp

    mats = []
    for m in shape:
        L = sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1]).tolil()
        L[0, 0] = 1
        L[m - 1, m - 1] = 1
        mats.append(L.tocsr())
    A = None
    for d in range(len(shape)):
        term = None
        for e in range(len(shape)):
            M = mats[e] if e == d else sp.identity(shape[e], format="csr")
            term = M if term is None else sp.kron(term, M, format="csr")
        A = term if A is None else A + term
    return A


@pytest.mark.parametrize("shape", [(16, 12), (8, 10, 6)])
def test_dct_solve_against_sparse_operator(shape):
    rng = np.random.default_rng(4)
    b = rng.standard_normal(shape)
    b -= b.mean()
    x = O.neumann_solve_dct(b)
    A = neumann_matrix(shape)
    res = A @ x.ravel() - b.ravel()
    assert np.linalg.norm(res) < 1e-10 * np.linalg.norm(b)
    assert abs(x.mean()) < 1e-12


def test_projection_divergence_free_idempotent_symmetric():
    n = (16, 16, 16)
    rng = np.random.default_rng(5)
    a = [rng.standard_normal(O.face_shape(n, d)) for d in range(3)]
    b = [rng.standard_normal(O.face_shape(n, d)) for d in range(3)]
    pa = [c.copy() for c in a]
    pb = [c.copy() for c in b]
    O.project_exact(pa)
    O.project_exact(pb)
    assert np.abs(O.divergence(pa)).max() < 1e-11
    ppa = [c.copy() for c in pa]
    O.project_exact(ppa)
    for x, y in zip(ppa, pa):
        np.testing.assert_allclose(x, y, atol=1e-11)
    # symmetric on fields with zero boundary normal flux
    for c in a + b:
        pass
    a0 = [c.copy() for c in a]
    b0 = [c.copy() for c in b]
    O.apply_bc(a0)
    O.apply_bc(b0)
    lhs = sum(np.sum(x * y) for x, y in zip(pa, b0))
    rhs = sum(np.sum(x * y) for x, y in zip(a0, pb))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


@pytest.mark.parametrize("shape,levels", [((32, 32), 3), ((16, 16, 16), 2)])
def test_mgpcg_restatement_converges_to_exact(shape, levels):
    rng = np.random.default_rng(6)
    b = rng.standard_normal(shape)
    b -= b.mean()
    x, it, res, st = O.mgpcg_solve(b, levels, 2, 50, 1e-8, True, 400)
    assert st == 0 and it < 40
    xe = O.neumann_solve_dct(b)
    np.testing.assert_allclose(x - x.mean(), xe, atol=1e-6 * np.abs(xe).max())


# ------------------------------------------------------------ obstacle masks --
def _disk_mask(n, c, r):
    X = O.cell_coords(n, 1.0 / n[1])
    return (np.sqrt(np.sum((X - np.asarray(c)) ** 2, axis=-1)) < r).astype(np.float64)


@pytest.mark.parametrize("n", [(24, 20), (12, 10, 8)])
def test_masked_projection_empty_mask_is_exact_projection(n):
    rng = np.random.default_rng(3)
    u = [rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))]
    a, b = [c.copy() for c in u], [c.copy() for c in u]
    O.project_exact(a)
    O.project_masked_exact(b, np.zeros(n, dtype=bool))
    assert max(np.abs(x - y).max() for x, y in zip(a, b)) < 1e-10


@pytest.mark.parametrize("n", [(32, 28), (14, 12, 10)])
def test_masked_projection_properties(n):
    """Divergence-free on fluid cells, zero on walls and on every face of a
    solid cell, idempotent; the fluid operator is symmetric with zero row sums."""
    rng = np.random.default_rng(4)
    c = [0.45] * len(n)
    solid = _disk_mask(n, c, 0.22) > 0
    u = [rng.standard_normal(O.face_shape(n, a)) for a in range(len(n))]
    O.project_masked_exact(u, solid)
    div = O.divergence(u)
    assert np.abs(div[~solid]).max() < 1e-9
    for a in range(len(n)):
        assert np.all(u[a][O._face_solid(solid, a)] == 0)
    again = [x.copy() for x in u]
    O.project_masked_exact(again, solid)
    assert max(np.abs(x - y).max() for x, y in zip(again, u)) < 1e-9
    A, fl = O.masked_laplacian(solid)
    assert abs(A - A.T).max() == 0
    assert np.abs(np.asarray(A.sum(axis=1))).max() == 0


def test_mask_field_kernels_closed_forms():
    n = (10, 9)
    mask = np.zeros(n)
    mask[3:5, 4:6] = 1.0          # solid block
    mask[8, 0] = 0.5              # neither solid nor fluid
    rng = np.random.default_rng(5)
    # zero mask: no-op
    u = [rng.standard_normal(O.face_shape(n, a)) for a in range(2)]
    ref = [c.copy() for c in u]
    O.mask_velocity(u, np.zeros(n))
    assert all(np.array_equal(a, b) for a, b in zip(u, ref))
    # solid-cell faces take the boundary velocity; other faces untouched
    bvel = [np.full(n, 2.0), np.full(n, -3.0)]
    O.mask_velocity(u, mask, bvel)
    assert np.all(u[0][3:6, 4:6] == 2.0) and np.all(u[1][3:5, 4:7] == -3.0)
    assert np.array_equal(u[0][:3], ref[0][:3]) and np.array_equal(u[1][:, :4], ref[1][:, :4])
    # constant scalar stays constant; the half-masked cell is left alone
    s = np.full(n, 1.25)
    s[8, 0] = 7.0
    O.mask_scalar(s, mask)
    assert np.all(s[3:5, 4:6] == 1.25) and s[8, 0] == 7.0
    # a solid cell with no fluid neighbour gets 0; adjoint form zeroes
    m2 = np.ones((3, 3))
    s2 = np.full((3, 3), 4.0)
    O.mask_scalar(s2, m2)
    assert np.all(s2 == 0)
    s3 = rng.standard_normal(n)
    O.mask_scalar(s3, mask, adjoint=True)
    assert np.all(s3[3:5, 4:6] == 0)
    # brute-force mean of the fluid 3x3 neighbours for one solid cell
    s4 = rng.standard_normal(n)
    want = np.mean([s4[i, j] for i in range(2, 6) for j in range(3, 7)
                    if 2 <= i <= 4 + 1 and abs(i - 3) <= 1 and abs(j - 4) <= 1 and mask[i, j] <= 0])
    O.mask_scalar(s4, mask)
    assert abs(s4[3, 4] - want) < 1e-12


# ---------------------------------------------------------- control forces --
def _forces(rng, count, dim, dx, radii=(2, 3, 4, 6)):
    c = rng.uniform(0.3, 0.7, (count, dim))
    r = np.array([1.0 / (rng.choice(radii) * dx) ** 2 for _ in range(count)])
    s = rng.uniform(-1, 1, (count, dim))
    return c, r, s


def test_gaussian_force_integral_node_value_and_additivity():
    n, dx = (96, 80), 1.0 / 80
    r = 1.0 / (4 * dx) ** 2
    c = np.array([[0.6, 0.5]])
    s = np.array([[0.7, -1.3]])
    f = O.gaussian_force(n, dx, c, [r], s, 1)
    # truncated Gaussian: integral over m^2 r < 5 is (pi / r)(1 - e^-5)
    want = (np.pi / r) * (1 - np.exp(-5.0))
    for a in range(2):
        assert abs(np.sum(f[a]) * dx * dx / s[0, a] - want) < 2e-3 * want
    # a centre on an x-face node gives the strength there (e = 1)
    node = np.array([[30 * dx, (20 + 0.5) * dx]])
    fx = O.gaussian_force(n, dx, node, [r], s, 1)[0]
    assert abs(fx[30, 20] - s[0, 0]) < 1e-6 * abs(s[0, 0])
    # forces add; count limits the forces used
    rng = np.random.default_rng(4)
    C, R, S = _forces(rng, 5, 2, dx)
    tot = O.gaussian_force(n, dx, C, R, S, 5)
    parts = [O.gaussian_force(n, dx, C[i:i + 1], R[i:i + 1], S[i:i + 1], 1) for i in range(5)]
    for a in range(2):
        np.testing.assert_allclose(tot[a], sum(p[a] for p in parts), atol=1e-12)
    first3 = O.gaussian_force(n, dx, C, R, S, 3)
    for a in range(2):
        np.testing.assert_allclose(first3[a], sum(p[a] for p in parts[:3]), atol=1e-12)


def test_gaussian_force_cutoff_is_the_fp32_test():
    """Brute force over nodes: a node is inside exactly when the FP32
    expression (m*m) r < 5 holds (which the FP64 test can disagree with)."""
    n, dx = (40, 40), 1.0 / 40
    r = np.float32(1.0 / (3 * dx) ** 2)
    c = np.array([[0.51, 0.47]], np.float32)
    f = O.gaussian_force(n, dx, c, [r], np.ones((1, 2)), 1)
    for a in range(2):
        shp = O.face_shape(n, a)
        for i in range(shp[0]):
            for j in range(shp[1]):
                p = np.array([(i + (0 if a == 0 else 0.5)), (j + (0 if a == 1 else 0.5))], np.float32)
                p = p * np.float32(dx)
                d = p - c[0]
                m = np.sqrt(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1]))
                assert (f[a][i, j] != 0) == bool(np.float32(m * m) * r < np.float32(5)), (a, i, j)


@pytest.mark.parametrize("dim", [2, 3])
def test_gaussian_force_adjoint_is_the_derivative(dim):
    """adj_strength / adj_center are d<f, lam>/d(s, c) and the penalty terms
    d/d(s, c) of sum over forces of sum f_I^2, checked by central
    differences (FP64 cut-off test; no node lies near the cut-off)."""
    n = (24, 20) if dim == 2 else (12, 10, 14)
    dx = 1.0 / n[1]
    rng = np.random.default_rng(11 + dim)
    C, R, S = _forces(rng, 3, dim, dx, radii=(2, 3))
    lam = [rng.standard_normal(O.face_shape(n, a)) for a in range(dim)]
    for I in range(3):  # keep nodes away from the discontinuity at the cut-off
        for a in range(dim):
            _, e, d = O._force_weights(n, a, dx, C[I], R[I], False)
            assert np.min(np.abs(np.sum(d * d, -1) * R[I] - 5.0)) > 5e-5  # h moves q by < 2e-5
    ac, ast, pc, ps = O.gaussian_force_adjoint(n, dx, C, R, S, 3, lam, fp32_test=False)

    def J(C_, S_):
        f = O.gaussian_force(n, dx, C_, R, S_, 3, fp32_test=False)
        return sum(np.sum(f[a] * lam[a]) for a in range(dim))

    def E(C_, S_):
        return sum(sum(np.sum(g ** 2) for g in O.gaussian_force(n, dx, C_[i:i + 1], R[i:i + 1], S_[i:i + 1], 1,
                                                                 fp32_test=False)) for i in range(3))

    h = 1e-6
    for I in range(3):
        for b in range(dim):
            dS = np.zeros_like(S)
            dS[I, b] = h
            dC = np.zeros_like(C)
            dC[I, b] = h
            for fn, want_s, want_c in ((J, ast, ac), (E, ps, pc)):
                gs = (fn(C, S + dS) - fn(C, S - dS)) / (2 * h)
                gc = (fn(C + dC, S) - fn(C - dC, S)) / (2 * h)
                assert abs(gs - want_s[I, b]) <= 1e-6 * max(1.0, abs(gs)), (fn.__name__, I, b)
                assert abs(gc - want_c[I, b]) <= 1e-5 * max(1.0, abs(gc)), (fn.__name__, I, b)


# -------------------------------------------------------- BFECC variants --
def test_clamp_extrema_brute_force():
    rng = np.random.default_rng(5)
    for shape in [(5, 4), (3, 4, 5)]:
        ref = rng.standard_normal(shape)
        c = 3 * rng.standard_normal(shape)
        out = O.clamp_extrema(c, ref)
        for idx in np.ndindex(*shape):
            nb = []
            for a in range(len(shape)):
                for d in (-1, 1):
                    j = list(idx)
                    j[a] = min(max(j[a] + d, 0), shape[a] - 1)
                    nb.append(ref[tuple(j)])
            assert out[idx] == min(max(c[idx], min(nb)), max(nb))


@pytest.mark.parametrize("clamp", [False, True])
def test_short_bfecc_steps_move_fields_the_semi_lagrangian_way(clamp):
    """One forward and one adjoint step with use_short_BFECC change the fields
    in the direction the semi-Lagrangian step does: the cosine between the two
    changes is high, where maps taken in the wrong direction give a negative
    one (measured: passive 0.98 vs -0.90 for reversed maps at dt = 2 dx)."""
    from dfm import scenes

    n = (32, 32)
    sc = scenes.scene_2d(n, seed=2, vortices=4)
    dx = sc["dx"]

    def cos(a, b):
        a = np.concatenate([x.ravel() for x in a])
        b = np.concatenate([x.ravel() for x in b])
        return float(a @ b / np.linalg.norm(a) / np.linalg.norm(b))

    runs = {}
    for short in (False, True):
        o = O.OracleLFM(n, dx, 2 * dx, 4, 2, bfecc_clamp=clamp, short_bfecc=short)  # no reinitialisation
        o.init(sc["u0"], [sc["density"]])
        u0, s0 = [c.copy() for c in o.u], o.s[0].copy()
        o.step()
        runs[short] = o
    du = {k: [a - b for a, b in zip(o.u, u0)] for k, o in runs.items()}
    ds = {k: [o.s[0] - s0] for k, o in runs.items()}
    assert cos(ds[True], ds[False]) > 0.9
    assert cos(du[True], du[False]) > 0.6
    # adjoint: a second step, then one backtrack step (j = 1, no BFECC block)
    # from the same terminal adjoint state
    lam0 = [0.5 * c for c in runs[False].u]
    dl = {}
    for short, o in runs.items():
        o.step()
        o.init_gradient([sc["target"]])
        o.lam_u = [c.copy() for c in lam0]
        o.init_lam_u = [c.copy() for c in lam0]
        o.backtrack_step()
        dl[short] = [a - b for a, b in zip(o.lam_u, lam0)]
    c = cos(dl[True], dl[False])
    assert c > 0.8  # reversed maps: 0.10


# ------------------------------------------------- linear (N_1) variant -----
def test_linear_bspline_matches_scipy_basis_element():
    """N_1 = the degree-1 B-spline (hat on knots -1, 0, 1) inside its support;
    its derivative is -1 on [0, 1) and +1 on (-1, 0) (the reference's dN_1
    takes -1 at x = 0)."""
    x = np.linspace(-0.999, 0.999, 801)
    ref = BSpline.basis_element([-1, 0, 1], extrapolate=False)(x)
    np.testing.assert_allclose(bspline(1, x), ref, atol=1e-15)
    xs = x[np.abs(x) > 1e-9]
    np.testing.assert_allclose(bspline(1, xs, 1), BSpline.basis_element([-1, 0, 1]).derivative()(xs), atol=1e-12)
    assert bspline(1, [0.0], 1)[0] == -1.0


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_interp_matches_scipy_regular_grid(dim):
    """interp(1, ...) is multilinear interpolation on the node lattice at
    BL*dx: scipy's RegularGridInterpolator at interior points; outside the
    box the coordinate clamps to [0, n-1-1e-6] (the nearest face value)."""
    from scipy.interpolate import RegularGridInterpolator

    rng = np.random.default_rng(5 + dim)
    shape = (9, 7, 6)[:dim]
    dx = 0.1
    bl = [0.0, 0.5, 0.5][:dim]
    f = rng.standard_normal(shape)
    axes = [(np.arange(s) + b) * dx for s, b in zip(shape, bl)]
    rgi = RegularGridInterpolator(axes, f, method="linear")
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    pts = lo + (hi - lo) * rng.uniform(0.001, 0.999, size=(200, dim))
    np.testing.assert_allclose(O.interp(1, f, bl, pts, dx), rgi(pts), atol=1e-12)
    # outside: clamped to the box (the reference's s = max(0, min(u, n-1-eps)))
    out = lo + (hi - lo) * rng.uniform(-0.5, 1.5, size=(200, dim))
    np.testing.assert_allclose(O.interp(1, f, bl, out, dx), rgi(np.clip(out, lo, hi)), atol=1e-5)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_interp_reproduces_affine_fields_and_gradient(dim):
    """A multilinear interpolant is exact on affine data, with the exact
    gradient inside a cell."""
    rng = np.random.default_rng(17 + dim)
    shape = (8, 9, 7)[:dim]
    dx = 0.25
    bl = [0.5] * dim
    c = rng.standard_normal(dim + 1)
    X = np.meshgrid(*[(np.arange(s) + 0.5) * dx for s in shape], indexing="ij")
    f = c[0] + sum(c[1 + d] * X[d] for d in range(dim))
    pts = 0.5 * dx + rng.uniform(0.01, 0.99, size=(100, dim)) * (np.array(shape) - 1) * dx
    v, g = O.interp(1, f, bl, pts, dx, grad=True)
    np.testing.assert_allclose(v, c[0] + pts @ c[1:], atol=1e-12)
    np.testing.assert_allclose(g, np.broadcast_to(c[1:], g.shape), atol=1e-12)
