"""Torch-tensor front end of the libdfm.so kernels.

Fields are float32 CUDA (HIP) tensors.  A MAC field is a list of face tensors
(component a has shape n + e_a); a flow map holds SoA point sets
(psi: (dim, *n), psi_f[a] / T_f[a]: (dim, *face_a)).
"""
from __future__ import annotations

import warnings

import torch

from . import _lib as L
from . import timing


FM_TILED_SOURCES = 0  # DFM_FM_TILED_SOURCES


def set_flowmap_option(option, value):
    """Library-wide kernel choice (dfm_flowmap_set_option); returns the previous value."""
    import ctypes

    lib = L.lib()
    old = ctypes.c_int(0)
    L.check(lib.dfm_flowmap_get_option(option, ctypes.byref(old)), "dfm_flowmap_get_option")
    L.check(lib.dfm_flowmap_set_option(option, int(value)), "dfm_flowmap_set_option")
    return old.value


def face_shape(n, a):
    return tuple(n[b] + (1 if b == a else 0) for b in range(len(n)))


class MacField:
    """One tensor per face axis."""

    def __init__(self, comps):
        self.c = list(comps)
        self._s = None

    @staticmethod
    def empty(n, device, zero=False):
        f = torch.zeros if zero else torch.empty
        return MacField([f(face_shape(n, a), dtype=torch.float32, device=device) for a in range(len(n))])

    @staticmethod
    def from_arrays(arrs, device):
        return MacField([torch.as_tensor(a, dtype=torch.float32).to(device).contiguous() for a in arrs])

    def clone(self):
        return MacField([t.clone() for t in self.c])

    def copy_(self, other):
        for a, b in zip(self.c, other.c):
            a.copy_(b)
        return self

    def zero_(self):
        for t in self.c:
            t.zero_()
        return self

    def numpy(self):
        return [t.detach().cpu().numpy() for t in self.c]

    @property
    def s(self):
        if self._s is None:
            self._s = L.make_mac(self.c)
        return self._s


class FlowMap:
    """Cell map psi plus per-face maps psi_f[a] and Jacobian columns T_f[a]."""

    def __init__(self, n, device, with_cells=True):
        dim = len(n)
        self.n = tuple(n)
        self.psi = torch.empty((dim,) + tuple(n), dtype=torch.float32, device=device) if with_cells else None
        self.psi_f = [torch.empty((dim,) + face_shape(n, a), dtype=torch.float32, device=device) for a in range(dim)]
        self.T_f = [torch.empty((dim,) + face_shape(n, a), dtype=torch.float32, device=device) for a in range(dim)]
        self._s = None

    @staticmethod
    def wrap(psi, psi_f, T_f):
        """A FlowMap over existing tensors (psi may be None; T_f entries may be
        None for a map without Jacobian columns)."""
        fm = FlowMap.__new__(FlowMap)
        face0 = psi_f[0].shape[1:]  # n + e_0
        fm.n = (face0[0] - 1,) + tuple(face0[1:])
        fm.psi, fm.psi_f, fm.T_f = psi, list(psi_f), list(T_f)
        fm._s = None
        return fm

    @property
    def s(self):
        if self._s is None:
            self._s = L.make_map(self.psi, self.psi_f, self.T_f)
        return self._s

    def copy_(self, other):
        if self.psi is not None:
            self.psi.copy_(other.psi)
        for a in range(len(self.psi_f)):
            self.psi_f[a].copy_(other.psi_f[a])
            self.T_f[a].copy_(other.T_f[a])
        return self


class Ops:
    """Kernel launcher bound to one grid; every call is asynchronous on the
    current torch stream."""

    def __init__(self, n, dx):
        self.n = tuple(n)
        self.dim = len(n)
        self.dx = float(dx)
        self.g = L.make_grid(n, dx)
        self.lib = L.lib()

    def _st(self):
        return L.stream_handle()

    def reset_identity(self, fm: FlowMap):
        L.check(self.lib.dfm_reset_identity(self.g, fm.s, self._st()), "reset_identity")

    def march(self, u: MacField, fm: FlowMap, dt, order, blocks_per_cu=0):
        """blocks_per_cu > 0: at most that many march workgroups per CU
        (dfm_march_capped), for a march beside other work on another stream."""
        with timing.region("march"):
            if blocks_per_cu:
                L.check(self.lib.dfm_march_capped(self.g, u.s, fm.s, float(dt), int(order), int(blocks_per_cu),
                                                  self._st()), "march_capped")
            else:
                L.check(self.lib.dfm_march(self.g, u.s, fm.s, float(dt), int(order), self._st()), "march")

    # kern: sampling spline of the advected quantity, 1 linear / 2 quadratic
    # (the reference's) / 3 cubic; the velocity moving points stays quadratic
    def pullback_u(self, src: MacField, fm: FlowMap, out: MacField, alpha=1.0, beta=0.0, add: MacField = None,
                   kern=2):
        L.check(self.lib.dfm_pullback_u_k(self.g, int(kern), src.s, fm.s, float(alpha), float(beta),
                                          add.s if add is not None else None, out.s, self._st()), "pullback_u")

    def pullback_s(self, src, psi, out, alpha=1.0, beta=0.0, add=None, kern=2):
        L.check(self.lib.dfm_pullback_s_k(self.g, int(kern), src.data_ptr(), psi.data_ptr(), float(alpha),
                                          float(beta), L.ptr(add), out.data_ptr(), self._st()), "pullback_s")

    def bfecc_u(self, init: MacField, fwd: FlowMap, bwd: FlowMap, err: MacField, out: MacField, kern=2):
        L.check(self.lib.dfm_bfecc_u_k(self.g, int(kern), init.s, fwd.s, bwd.s, err.s, out.s, self._st()), "bfecc_u")

    def bfecc_s(self, init, psi, phi, err, out, kern=2):
        L.check(self.lib.dfm_bfecc_s_k(self.g, int(kern), init.data_ptr(), psi.data_ptr(), phi.data_ptr(),
                                       err.data_ptr(), out.data_ptr(), self._st()), "bfecc_s")

    def bfecc_u_clamp(self, init: MacField, fwd: FlowMap, bwd: FlowMap, err: MacField, tmp: MacField, out: MacField):
        L.check(self.lib.dfm_bfecc_u_clamp(self.g, init.s, fwd.s, bwd.s, err.s, tmp.s, out.s, self._st()),
                "bfecc_u_clamp")

    def bfecc_s_clamp(self, init, psi, phi, err, tmp, out):
        L.check(self.lib.dfm_bfecc_s_clamp(self.g, init.data_ptr(), psi.data_ptr(), phi.data_ptr(), err.data_ptr(),
                                           tmp.data_ptr(), out.data_ptr(), self._st()), "bfecc_s_clamp")

    def accumulate(self, f: MacField, fm: FlowMap, init: MacField, h1: MacField, h2: MacField, ind, dt,
                   f2: MacField = None, dt2=0.0):
        L.check(self.lib.dfm_accumulate(self.g, f.s, fm.s, init.s, h1.s, h2.s, int(ind), float(dt),
                                        f2.s if f2 is not None else None, float(dt2), self._st()), "accumulate")

    def sl_advect(self, vel: MacField, dt, order, vec_src=(), vec_dst=(), s_src=(), s_dst=(), kern=2):
        nv, ns = len(vec_src), len(s_src)
        src_arr = (L.Mac * max(1, nv))(*[m.s for m in vec_src])
        dst_arr = (L.Mac * max(1, nv))(*[m.s for m in vec_dst])
        with timing.region("sl_advect"):
            L.check(self.lib.dfm_sl_advect_k(self.g, int(kern), vel.s, float(dt), int(order), nv, src_arr, dst_arr, ns,
                                             L.ptr_array(s_src), L.ptr_array(s_dst), self._st()), "sl_advect")

    # vector-Jacobian products (the gradients accumulate into the g_* outputs)
    def pullback_u_vjp(self, g_out: MacField, fm: FlowMap, g_src: MacField, alpha=1.0, src: MacField = None,
                       g_fm: FlowMap = None, kern=2):
        """g_src += transpose of the pullback applied to g_out; with g_fm (and
        src) also the map's psi_f / T_f gradients."""
        L.check(self.lib.dfm_pullback_u_vjp(self.g, int(kern), g_out.s, src.s if src is not None else None, fm.s,
                                            float(alpha), g_src.s, g_fm.s if g_fm is not None else None,
                                            self._st()), "pullback_u_vjp")

    def pullback_s_vjp(self, g_out, psi, g_src, alpha=1.0, src=None, g_psi=None, kern=2):
        L.check(self.lib.dfm_pullback_s_vjp(self.g, int(kern), g_out.data_ptr(), L.ptr(src), psi.data_ptr(),
                                            float(alpha), g_src.data_ptr(), L.ptr(g_psi), self._st()),
                "pullback_s_vjp")

    def sl_advect_vjp(self, vel: MacField, dt, order, g_vec_dst=(), g_vec_src=(), g_s_dst=(), g_s_src=(),
                      vec_src=(), s_src=(), g_vel: MacField = None, kern=2):
        """Transpose of sl_advect w.r.t. its sources; with g_vel (and the
        forward sources vec_src / s_src) also the advecting velocity's."""
        nv, ns = len(g_vec_dst), len(g_s_dst)
        mk = lambda ms: (L.Mac * max(1, nv))(*[m.s for m in ms])  # noqa: E731
        L.check(self.lib.dfm_sl_advect_vjp(self.g, int(kern), vel.s, float(dt), int(order), nv,
                                           mk(vec_src) if vec_src else None, mk(g_vec_dst), mk(g_vec_src), ns,
                                           L.ptr_array(s_src) if s_src else None, L.ptr_array(g_s_dst),
                                           L.ptr_array(g_s_src), g_vel.s if g_vel is not None else None, self._st()),
                "sl_advect_vjp")

    def march_vjp(self, u: MacField, fm0: FlowMap, dt, order, g_fm: FlowMap, g_u: MacField):
        """g_fm: gradient w.r.t. the marched map on entry, w.r.t. fm0 on
        return; g_u accumulates the velocity's."""
        L.check(self.lib.dfm_march_vjp(self.g, u.s, fm0.s, float(dt), int(order), g_fm.s, g_u.s, self._st()),
                "march_vjp")

    def adj_conv_source(self, u: MacField, w: MacField, kern, r: MacField):
        with timing.region("adj_conv"):
            L.check(self.lib.dfm_adj_conv_source(self.g, u.s, w.s, int(kern), r.s, self._st()), "adj_conv_source")

    def adj_scalar_source(self, s_list, lam_list, r: MacField):
        L.check(self.lib.dfm_adj_scalar_source(self.g, len(s_list), L.ptr_array(s_list), L.ptr_array(lam_list),
                                               r.s, self._st()), "adj_scalar_source")

    def mask_velocity(self, mask, u: MacField, bvel=None):
        """mask_velocity_by_boundary (bvel: (dim, *n) cell tensor) or, with
        bvel None, mask_adj_velocity_by_boundary."""
        L.check(self.lib.dfm_mask_velocity(self.g, mask.data_ptr(), L.ptr(bvel), u.s, self._st()), "mask_velocity")

    def mask_scalar(self, mask, s, adjoint=False):
        """mask_passive_by_boundary (adjoint False) / mask_adj_passive_by_boundary."""
        L.check(self.lib.dfm_mask_scalar(self.g, mask.data_ptr(), s.data_ptr(), 1 if adjoint else 0, self._st()),
                "mask_scalar")

    def central_vector(self, u: MacField, central):
        """get_central_vector: central (dim, *n) FP32 = face averages of u."""
        L.check(self.lib.dfm_central_vector(self.g, u.s, central.data_ptr(), self._st()), "central_vector")

    def split_central_vector(self, central, u: MacField):
        """split_central_vector: MAC u from the cell-centred (dim, *n) field."""
        L.check(self.lib.dfm_split_central_vector(self.g, central.data_ptr(), u.s, self._st()), "split_central_vector")

    def count_nonfinite(self, tensors, out):
        """Queue the count of NaN / Inf entries of the FP32 tensors into
        `out` (a uint32-sized device tensor, e.g. int32 of shape (1,));
        asynchronous."""
        import ctypes

        ts = list(tensors)
        ptrs = (ctypes.c_void_p * max(1, len(ts)))(*[t.data_ptr() for t in ts])
        sizes = (ctypes.c_int64 * max(1, len(ts)))(*[t.numel() for t in ts])
        L.check(self.lib.dfm_count_nonfinite(len(ts), ptrs, sizes, out.data_ptr(), self._st()), "count_nonfinite")

    def max_speed(self, u: MacField, out):
        """calc_max_speed into the device scalar `out` (FP32)."""
        L.check(self.lib.dfm_max_speed(self.g, u.s, out.data_ptr(), self._st()), "max_speed")

    def gaussian_force(self, center, inv_radius, strength, count, f: MacField):
        """apply_gausion_force: f = the field of the first `count` Gaussian
        forces (center (N, dim), inv_radius (N,), strength (N, dim); device FP32)."""
        L.check(self.lib.dfm_gaussian_force(self.g, L.ptr(center), L.ptr(inv_radius), L.ptr(strength), int(count),
                                            f.s, self._st()), "gaussian_force")

    def gaussian_force_adjoint(self, center, inv_radius, strength, count, adj_u: MacField, adj_center, adj_strength,
                               penalty_center, penalty_strength):
        """calculate_adj_gausion_force: the first `count` rows of the four
        (N, dim) outputs; rows beyond `count` are left as they are."""
        L.check(self.lib.dfm_gaussian_force_adjoint(
            self.g, L.ptr(center), L.ptr(inv_radius), L.ptr(strength), int(count), adj_u.s, L.ptr(adj_center),
            L.ptr(adj_strength), L.ptr(penalty_center), L.ptr(penalty_strength), self._st()), "gaussian_force_adjoint")

    def axpy(self, x: MacField, alpha, y: MacField, out: MacField, beta=0.0, z: MacField = None):
        L.check(self.lib.dfm_axpy_mac(self.g, x.s, float(alpha), y.s, float(beta),
                                      z.s if z is not None else None, out.s, self._st()), "axpy")


class NoConvergenceWarning(RuntimeWarning):
    """A PCG solve stopped at max_iters above its tolerance (DFM_E_NOCONV);
    the gradient was still subtracted, so the field is only approximately
    divergence-free."""


class MGPCG:
    """Pressure projection on an all-Neumann box (the reference's
    ``solver.Poisson`` / ``Poisson_adj``, 2D/vortex/mgpcg.py:363-368,
    3D/fast_mgpcg.py:241-323), backed by the libdfm MGPCG handle.

    ``on_noconv`` says what a solve that reaches max_iters without converging
    does: "warn" (default) issues a NoConvergenceWarning, "raise" raises the
    DfmError, "count" only counts it.  Every such solve is counted in
    ``noconv_count`` (the reference prints and carries on,
    2D/vortex/mgpcg.py:280-282)."""

    def __init__(self, n, dx, levels=0, pre_post=2, bottom=50, tol=1e-6, relative=True, max_iters=400,
                 on_noconv="warn", bf16_coarse=False, tail=True, graph=True, check_every=1,
                 bottom_exact=False, fused_sweeps=2, split_layout=True, tail_level=0):
        self.n = tuple(n)
        self.g = L.make_grid(n, dx)
        self.lib = L.lib()
        h = L.c_void_p()
        L.check(self.lib.dfm_mgpcg_create(self.g, int(levels), int(pre_post), int(bottom), L.ctypes.byref(h)),
                "mgpcg_create")
        self.h = h
        if not tail:
            self.set_option(L.MG_TAIL, 0)
        if bf16_coarse:
            self.set_option(L.MG_BF16_COARSE, 1)
        if not graph:
            self.set_option(L.MG_GRAPH, 0)
        if check_every != 1:
            self.set_option(L.MG_CHECK_EVERY, check_every)
        if bottom_exact:
            self.set_option(L.MG_BOTTOM_EXACT, 1)
        if fused_sweeps != 2:
            self.set_option(L.MG_FUSED_SWEEPS, fused_sweeps)
        if not split_layout:
            self.set_option(L.MG_SPLIT_LAYOUT, 0)
        if tail_level:
            self.set_option(L.MG_TAIL_LEVEL, tail_level)
        self.tol, self.relative, self.max_iters = tol, relative, max_iters
        if on_noconv not in ("warn", "raise", "count"):
            raise ValueError(f"on_noconv must be 'warn', 'raise' or 'count', not {on_noconv!r}")
        self.on_noconv = on_noconv
        self.last_iters = 0
        self.noconv_count = 0

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.dfm_mgpcg_destroy(h)
            except Exception:
                pass
            self.h = None

    def set_option(self, option, value):
        L.check(self.lib.dfm_mgpcg_set_option(self.h, int(option), int(value)), "mgpcg_set_option")

    def get_option(self, option):
        v = L.c_int(0)
        L.check(self.lib.dfm_mgpcg_get_option(self.h, int(option), L.ctypes.byref(v)), "mgpcg_get_option")
        return v.value

    def set_mask(self, mask):
        """Obstacle mask (cells, solid where > 0) for the projection, or None.
        The solver derives and keeps its own per-level solid flags."""
        L.check(self.lib.dfm_mgpcg_set_mask(self.h, L.ptr(mask), L.stream_handle()), "mgpcg_set_mask")

    def status(self):
        it = L.c_int(0)
        res = L.c_double(0)
        L.check(self.lib.dfm_mgpcg_status(self.h, L.ctypes.byref(it), L.ctypes.byref(res)), "mgpcg_status")
        return it.value, res.value

    def _done(self, rc, what):
        if rc == L.DFM_E_NOCONV:
            self.noconv_count += 1
            timing.add_count("mgpcg_noconv")
            if self.on_noconv != "raise":
                if self.on_noconv == "warn":
                    it, res = self.status()
                    warnings.warn(f"{what}: PCG stopped at {it} iterations with relative residual {res:.3e} "
                                  f"above tol {self.tol:g}", NoConvergenceWarning, stacklevel=3)
                rc = 0
        L.check(rc, what)
        self.last_iters = self.status()[0]

    def Poisson(self, u: MacField):
        with timing.region("mgpcg"):
            rc = self.lib.dfm_mgpcg_project(self.h, u.s, float(self.tol), 1 if self.relative else 0,
                                            int(self.max_iters), L.stream_handle())
        self._done(rc, "mgpcg_project")
        timing.add_count("mgpcg_iters", self.last_iters)

    # the projection is symmetric and idempotent, so its adjoint is itself
    Poisson_adj = Poisson

    def Poisson_begin(self, u: MacField):
        """Queue a projection of u on the current stream without waiting;
        complete it with Poisson_end (independent projections on other
        handles and streams may be begun in between, and then overlap)."""
        rc = self.lib.dfm_mgpcg_project_begin(self.h, u.s, float(self.tol), 1 if self.relative else 0,
                                              int(self.max_iters), L.stream_handle())
        L.check(rc, "mgpcg_project_begin")

    def Poisson_end(self):
        rc = self.lib.dfm_mgpcg_project_end(self.h)
        self._done(rc, "mgpcg_project_end")
        timing.add_count("mgpcg_iters", self.last_iters)

    def solve(self, b, x):
        rc = self.lib.dfm_mgpcg_solve(self.h, b.data_ptr(), x.data_ptr(), float(self.tol),
                                      1 if self.relative else 0, int(self.max_iters), L.stream_handle())
        self._done(rc, "mgpcg_solve")
        return x
