"""Seeded synthetic scenes for the benchmark configurations (host-side set-up,
run once per simulation; not on the hot path).

Velocity profiles restate the reference's Gaussian ("Lamb-Oseen-like")
vortex, ``angular_vel_func`` (2D/vortex/init_conditions.py:141-145)
with the per-vortex form of ``sixteen_vortex_vel_func_with_pos_coef``
(2D/vortex/init_conditions.py:368-383): u(x) = sum_k w_k * s/r^2 *
(1 - exp(-r^2/a_k^2)) * (-p_y, p_x) * 2.5 with s = -0.01, r = |x - c_k| + 1e-6
and a_k = 0.05 rr_k.
Cell-centred velocities are split to the MAC faces with
``split_central_vector`` (2D/vortex/taichi_utils.py:158-167,
3D/taichi_utils.py:170-183) and then projected by the simulator.

Everything here is float64 numpy; callers cast to the simulation dtype.
"""
from __future__ import annotations

import numpy as np

SWIRL_S = -0.01
SWIRL_GAIN = 2.5


def cell_centers(n, dx):
    axes = [(np.arange(m) + 0.5) * dx for m in n]
    return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)


def angular_vel(r, rad, strength=SWIRL_S):
    """angular_vel_func (2D/vortex/init_conditions.py:141-145)."""
    r = r + 1e-6
    return strength / r * (1.0 - np.exp(-(r ** 2) / (rad ** 2))) / r


def split_central_vector(uc):
    """Cell-centred vector field (n..., d) -> list of MAC face fields, each the
    average of the two adjacent cells with clamp-to-edge sampling."""
    dim = uc.shape[-1]
    out = []
    for a in range(dim):
        c = np.moveaxis(uc[..., a], a, 0)
        n = c.shape[0]
        idx_r = np.clip(np.arange(n + 1), 0, n - 1)
        idx_l = np.clip(np.arange(n + 1) - 1, 0, n - 1)
        f = 0.5 * (c[idx_r] + c[idx_l])
        out.append(np.ascontiguousarray(np.moveaxis(f, 0, a)))
    return out


# ------------------------------------------------------------------- 2D -----
# theta follows the reference's control vector layout (4K values, blocks of K):
# [w_0..w_K-1, cx_0.., cy_0.., rr_0..] with vortex radius a = 0.05 rr
# (sixteen_vortex_vel_func_with_pos_coef, 2D/vortex/init_conditions.py:368-383;
# 2D/vortex/optimize_vortex.py:16-50 builds it for K = 16).
RADIUS_SCALE = 0.05


def vortex_params_2d(seed, count=8):
    """Flat theta (4 count): strengths +-U[0.5,1.5], centres U[0.2,0.8]^2,
    radii a = 0.05 rr in U[0.02,0.06]."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, count) * rng.choice([-1.0, 1.0], count)
    c = rng.uniform(0.2, 0.8, (count, 2))
    a = rng.uniform(0.02, 0.06, count)
    return np.concatenate([w, c[:, 0], c[:, 1], a / RADIUS_SCALE])


def unpack_theta_2d(theta):
    """(K, 4) rows (w, cx, cy, a) from the flat reference-layout theta."""
    t = np.asarray(theta, dtype=np.float64).reshape(4, -1)
    return np.stack([t[0], t[1], t[2], RADIUS_SCALE * t[3]], axis=1)


def vortex_velocity_2d(theta, n, dx):
    """Cell-centred velocity of the vortex sum, shape (nx, ny, 2)."""
    X = cell_centers(n, dx)
    u = np.zeros(tuple(n) + (2,))
    for w, cx, cy, a in unpack_theta_2d(theta):
        p = X - np.array([cx, cy])
        r = np.sqrt(p[..., 0] ** 2 + p[..., 1] ** 2)
        f = angular_vel(r, a) * w * SWIRL_GAIN
        u[..., 0] += f * -p[..., 1]
        u[..., 1] += f * p[..., 0]
    return u


def disk_density(n, dx, radius=0.15, center=None, width=None):
    """Smoothed disk/ball: 0.5 (1 - tanh((|x - c| - r) / (2 dx)))."""
    X = cell_centers(n, dx)
    c = np.full(len(n), 0.5) if center is None else np.asarray(center)
    w = 2 * dx if width is None else width
    d = np.sqrt(np.sum((X - c) ** 2, axis=-1))
    return 0.5 * (1.0 - np.tanh((d - radius) / w))


def gaussian_blobs(n, dx, seed, count, sigma=0.08, lo=0.3, hi=0.7):
    """Target density: sum of `count` seeded Gaussian blobs."""
    rng = np.random.default_rng(seed)
    X = cell_centers(n, dx)
    s = np.zeros(tuple(n))
    for _ in range(count):
        c = rng.uniform(lo, hi, len(n))
        amp = rng.uniform(0.5, 1.0)
        s += amp * np.exp(-np.sum((X - c) ** 2, axis=-1) / (2 * sigma ** 2))
    return s


def scene_2d(n=(256, 256), seed=0, vortices=8):
    """Config: 8 Gaussian vortices, smoothed disk density r=0.15, target = 3
    Gaussian blobs (sigma 0.08).  Returns dict with MAC u0 (unprojected),
    density, target, theta, dx."""
    dx = 1.0 / n[1]
    theta = vortex_params_2d(seed, vortices)
    uc = vortex_velocity_2d(theta, n, dx)
    return {
        "dx": dx,
        "theta": theta,
        "u0": split_central_vector(uc),
        "density": disk_density(n, dx),
        "target": gaussian_blobs(n, dx, seed + 1000, 3),
    }


# ------------------------------------------------------------------- 3D -----
def tube_params_3d(seed, count=16):
    """[w, cx, cy, cz, ax, ay, az, a] per tube: centres U[0.25,0.75]^3, unit
    axes uniform on the sphere, radii U[0.03,0.08], strengths +-U[0.5,1.5]."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, count) * rng.choice([-1.0, 1.0], count)
    c = rng.uniform(0.25, 0.75, (count, 3))
    ax = rng.normal(size=(count, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    a = rng.uniform(0.03, 0.08, count)
    return np.concatenate([w[:, None], c, ax, a[:, None]], axis=1)


def vortex_tubes_3d(params, n, dx):
    """Cell-centred velocity of straight vortex tubes: the 2D swirl extruded
    along each tube axis, u = w * ang(r) * (axis x p_perp) * 2.5."""
    X = cell_centers(n, dx)
    u = np.zeros(tuple(n) + (3,))
    for row in np.asarray(params, dtype=np.float64):
        w, c, ax, a = row[0], row[1:4], row[4:7], row[7]
        p = X - c
        along = p @ ax
        perp = p - along[..., None] * ax
        r = np.sqrt(np.sum(perp ** 2, axis=-1))
        f = angular_vel(r, a) * w * SWIRL_GAIN
        u += f[..., None] * np.cross(np.broadcast_to(ax, perp.shape), perp)
    return u


def box_union_density(n, dx, seed, count=3):
    """Seeded union (max) of tanh-smoothed boxes."""
    rng = np.random.default_rng(seed)
    X = cell_centers(n, dx)
    s = np.zeros(tuple(n))
    for _ in range(count):
        c = rng.uniform(0.35, 0.65, 3)
        h = rng.uniform(0.06, 0.14, 3)
        box = np.ones(tuple(n))
        for d in range(3):
            box *= 0.5 * (1.0 - np.tanh((np.abs(X[..., d] - c[d]) - h[d]) / (2 * dx)))
        s = np.maximum(s, box)
    return s


def scene_3d(n=(128, 128, 128), seed=0, tubes=16):
    """Config: 16 vortex tubes, box-union density, target = 4 Gaussians."""
    dx = 1.0 / n[1]
    params = tube_params_3d(seed, tubes)
    uc = vortex_tubes_3d(params, n, dx)
    return {
        "dx": dx,
        "params": params,
        "u0": split_central_vector(uc),
        "density": box_union_density(n, dx, seed + 500),
        "target": gaussian_blobs(n, dx, seed + 1000, 4, lo=0.3, hi=0.7),
    }
