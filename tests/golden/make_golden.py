"""Generate the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no golden vectors and cannot run here (Taichi is absent),
so these fixtures are the oracle's outputs on seeded inputs, committed so
that drift of either the oracle or the GPU path shows up as a test failure.
Inputs are stored as the FP32 values both paths start from.

    python tests/golden/make_golden.py      (from the repository root)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "differentiable-flowmap_amd"))

from dfm import scenes  # noqa: E402
from oracle import lfm_oracle as O  # noqa: E402

CASES = {
    # name: (dim, n, R, steps, seed, count)
    "lfm2d_32": (2, (32, 32), 4, 8, 3, 4),
    "lfm2d_48x40": (2, (48, 40), 3, 6, 8, 5),
    "lfm3d_16": (3, (16, 16, 16), 3, 6, 3, 6),
}


def r32(a):
    return np.asarray(a, dtype=np.float32)


def lfm_case(dim, n, R, steps, seed, count):
    if dim == 2:
        sc = scenes.scene_2d(n, seed=seed, vortices=count)
    else:
        sc = scenes.scene_3d(n, seed=seed, tubes=count)
    dx = sc["dx"]
    dt = 0.5 * dx
    u0 = [r32(c) for c in sc["u0"]]
    s0, tgt = r32(sc["density"]), r32(sc["target"])
    order = 2 if dim == 2 else 4
    orc = O.OracleLFM(n, dx, dt, R, order)
    orc.init([c.astype(np.float64) for c in u0], [s0.astype(np.float64)])
    for _ in range(steps):
        orc.step()
    out = {"u_T_%d" % a: orc.u[a] for a in range(dim)}
    out["s_T"] = orc.s[0]
    orc.init_gradient([tgt.astype(np.float64)])
    out["loss"] = np.float64(orc.loss)
    for _ in range(steps):
        orc.backtrack_step()
    out.update({"lam_u0_%d" % a: orc.lam_u[a] for a in range(dim)})
    out["lam_s0"] = orc.lam_s[0]
    meta = {"dim": dim, "n": np.array(n), "dx": np.float64(dx), "dt": np.float64(dt), "R": R, "steps": steps,
            "order": order, "seed": seed}
    inputs = {"u0_%d" % a: u0[a] for a in range(dim)}
    inputs.update({"s0": s0, "target": tgt})
    return {**meta, **inputs, **out}


def projection_case():
    """Exact (DCT) projection of a seeded FP32 field on a non-cubic 3D grid."""
    n = (24, 20, 16)
    rng = np.random.default_rng(42)
    u = [r32(rng.standard_normal(O.face_shape(n, a))) for a in range(3)]
    p = [c.astype(np.float64) for c in u]
    O.project_exact(p)
    return {"n": np.array(n), **{"u_%d" % a: u[a] for a in range(3)}, **{"Pu_%d" % a: p[a] for a in range(3)}}


def vortex_case():
    """2D vortex initializer faces and the VJP of a seeded lambda (the VJP by
    central differences of the FP64 initializer, eps = 1e-6)."""
    n, dx = (40, 36), 1.0 / 36
    theta = r32(scenes.vortex_params_2d(21, 5)).astype(np.float64)
    faces = scenes.split_central_vector(scenes.vortex_velocity_2d(theta, n, dx))
    rng = np.random.default_rng(5)
    lam = [r32(rng.standard_normal(f.shape)) for f in faces]
    vjp = np.zeros_like(theta)
    for i in range(theta.size):
        e = np.zeros_like(theta)
        e[i] = 1e-6
        up = scenes.split_central_vector(scenes.vortex_velocity_2d(theta + e, n, dx))
        um = scenes.split_central_vector(scenes.vortex_velocity_2d(theta - e, n, dx))
        vjp[i] = sum(np.sum(l.astype(np.float64) * (a - b)) for l, a, b in zip(lam, up, um)) / 2e-6
    return {"n": np.array(n), "dx": np.float64(dx), "theta": r32(theta), "u_0": faces[0], "u_1": faces[1],
            "lam_0": lam[0], "lam_1": lam[1], "vjp": vjp}


def main():
    for name, args in CASES.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **lfm_case(*args))
        print("wrote", name)
    np.savez_compressed(os.path.join(HERE, "projection3d_24x20x16.npz"), **projection_case())
    np.savez_compressed(os.path.join(HERE, "vortex2d_40x36.npz"), **vortex_case())
    print("done")


if __name__ == "__main__":
    main()
