"""``PoissonSolver`` with the interface MGPCG_3 calls (3D/fast_mgpcg.py:5,192,258-273),
backed by the libdfm HIP MGPCG.

The reference's 3D default projection hands the pressure solve to an
external CUDA extension of this module name (AMGPCG_Pybind, not vendored).
With this directory on ``sys.path``, the unmodified ``fast_mgpcg.MGPCG_3``
imports this class instead:

    solver = PoissonSolver(tile_dim_vec, level_num=3, bottom_smoothing=40)
    solver.setup(a_x=..., a_y=..., a_z=..., b=..., a_diag=..., is_dof=...,
                 x_initial=..., is_uniform=False, uniform_coef=1.0,
                 neg_bc_type_vec=[1, 1, 1], pos_bc_type_vec=[1, 1, 1])
    solver.build()
    solver.solve(x)          # x: flat 8^3-tiled FP32 device tensor, written in place

All arrays are flat CUDA tensors in the 8^3-tiled order of
``rearrange_tensor_for_cuda`` (3D/fast_mgpcg.py:121-145).  The operator
MGPCG_3 builds -- the all-Neumann 7-point Laplacian with every cell a degree
of freedom (3D/fast_mgpcg.py:48-74,162-190) -- is the one the HIP MGPCG
solves; ``build`` checks the coefficient arrays against it on the device and
raises for any other operator or boundary type.  ``solve`` solves A x = b
from x = 0 (MGPCG_3 always passes a zero initial guess), b recentred to mean
zero, to the relative residual ``tol`` (default 1e-6); the all-Neumann
solution is defined up to a constant, and MGPCG_3 uses only its gradient.
``level_num`` and ``bottom_smoothing`` configure the
external solver's own hierarchy; this solver keeps its own (levels
log2(min n) - 2, LDS coarse tail with an exact bottom solve), so they are
recorded and not used.
"""
from __future__ import annotations

import numpy as np
import torch

from dfm import _lib as L
from dfm.ops import MGPCG

NEUMANN = 1


class PoissonSolver:
    def __init__(self, tile_dim_vec, level_num=3, bottom_smoothing=40, tol=1e-6, max_iters=400):
        t = [int(v) for v in tile_dim_vec]
        if len(t) != 3 or min(t) < 1:
            raise ValueError(f"tile_dim_vec must hold 3 positive tile counts, got {tile_dim_vec}")
        self.n = tuple(8 * v for v in t)
        self.level_num, self.bottom_smoothing = int(level_num), int(bottom_smoothing)
        self.tol, self.max_iters = float(tol), int(max_iters)
        self._g = L.make_grid(self.n, 1.0 / self.n[1])
        self._mg = None
        self._args = None
        self._built = False
        self._b = None
        self._x = None

    def _flat(self, t, name, dtype):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a CUDA tensor")
        if t.dtype != dtype:
            raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
        if t.numel() != int(np.prod(self.n)) or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous tensor of {int(np.prod(self.n))} elements")
        return t

    def setup(self, a_x, a_y, a_z, b, a_diag, is_dof, x_initial, is_uniform=False, uniform_coef=1.0,
              neg_bc_type_vec=(NEUMANN,) * 3, pos_bc_type_vec=(NEUMANN,) * 3):
        if list(neg_bc_type_vec) != [NEUMANN] * 3 or list(pos_bc_type_vec) != [NEUMANN] * 3:
            raise NotImplementedError("only the all-Neumann box (bc type 1 on every side) is supported")
        if is_uniform and float(uniform_coef) != 1.0:
            raise NotImplementedError("a uniform operator must have uniform_coef 1 (the unit Laplacian)")
        f32 = torch.float32
        self._args = dict(a_x=self._flat(a_x, "a_x", f32), a_y=self._flat(a_y, "a_y", f32),
                          a_z=self._flat(a_z, "a_z", f32), a_diag=self._flat(a_diag, "a_diag", f32),
                          is_dof=self._flat(is_dof, "is_dof", torch.uint8), b=self._flat(b, "b", f32),
                          x_initial=self._flat(x_initial, "x_initial", f32), is_uniform=bool(is_uniform))
        self._built = False

    def build(self):
        if self._args is None:
            raise RuntimeError("PoissonSolver.build() before setup()")
        a = self._args
        if not a["is_uniform"]:
            bad = L.c_int(0)
            L.check(L.lib().dfm_poisson_operator_check(
                self._g, a["a_diag"].data_ptr(), a["a_x"].data_ptr(), a["a_y"].data_ptr(), a["a_z"].data_ptr(),
                a["is_dof"].data_ptr(), 1, L.ctypes.byref(bad), L.stream_handle()), "poisson_operator_check")
            if bad.value:
                raise NotImplementedError(f"{bad.value} cells differ from the all-Neumann 7-point operator; "
                                          "only that operator is supported")
        if self._mg is None:
            self._mg = MGPCG(self.n, 1.0 / self.n[1], tol=self.tol, relative=True, max_iters=self.max_iters,
                             bottom_exact=True)
            self._b = torch.empty(self.n, dtype=torch.float32, device=a["b"].device)
            self._x = torch.empty_like(self._b)
        self._built = True

    def solve(self, x):
        if not self._built:
            raise RuntimeError("PoissonSolver.solve() before build()")
        x = self._flat(x, "x", torch.float32)
        lib, st = L.lib(), L.stream_handle()
        L.check(lib.dfm_tile8_convert(self._g, self._args["b"].data_ptr(), self._b.data_ptr(), 0, st), "tile8")
        self._mg.solve(self._b, self._x)
        L.check(lib.dfm_tile8_convert(self._g, self._x.data_ptr(), x.data_ptr(), 1, st), "tile8")
        return x

    def status(self):
        """(iterations, relative residual) of the last solve."""
        return self._mg.status()
