"""Drop-in LFM_Simulator / LFM_Diff_Simulator backed by libdfm.so.

Mirrors the reference's simulator-class API (the boundary its optimisers call):

* ``LFM_Simulator(res_x, res_y[, res_z], dx, dt, reinit_every, log_dir,
  log_target_dir=None)`` (2D/vortex/lfm_midpoint_simulator.py:188,
  3D/lfm_midpoint_simulator.py:169), ``.init(theta=None, forward=True,
  target=False, init_passive_np=None)`` (2D :370), ``.step_midpoint(
  write_flow_map, write_passive, control_force_para)`` -> ``(output_frame,
  final_flag)`` (2D :472-515);
* ``LFM_Diff_Simulator.forward_step_midpoint`` / ``init_gradient`` /
  ``backtrack_step_midpoint`` (2D/vortex/lfm_midpoint_diff_simulator.py:118-258);
* attributes ``u_x, u_y[, u_z]``, ``adj_u_*``, ``passive``, ``adj_passive``,
  ``loss``, ``step_num[None]``, ``frame_idx[None]`` with ``to_numpy()`` /
  ``from_numpy()`` like Taichi fields.

Hyper-parameters that the reference reads from its star-imported
``hyperparameters.py`` (RK_number, total_steps, add_passive_scalar ...) are
keyword arguments here.  Supported configuration (the one every benchmark
uses): viscosity 0, backward_u=True, final-time density loss; the keywords
``BFECC_clamp`` and ``use_short_BFECC`` select the reference's BFECC variants.  Obstacles: ``set_boundary_mask`` switches the
projection to the obstacle solver (MGPCG_2_solid) and applies the reference's
mask_* kernels at its call sites.  Control forces: ``control_num=N`` (the
reference's add_control_force / control_num) enables ``control_force_para``
({"c": centres (N, dim), "r": inverse squared radii (N,), "s": strengths
(N, dim), "n": count}) in ``step_midpoint`` and ``backtrack_step_midpoint``;
the adjoint step leaves the parameter gradients in ``adj_gaussion_force_*``
and ``penalty_gaussion_force_*`` (N, dim), as the reference does.

The reference's disk cache (RW_Cache, 2D/vortex/lfm_midpoint_simulator.py:46-184)
is replaced by device-resident storage.  ``record=True`` (default) keeps the
midpoint velocity of every step, the passive scalars at every
reinitialisation boundary and the forward map (phi, F) of every interval in
HBM, as the reference keeps them on disk.  ``checkpoint=True`` keeps only the
velocity and passive scalars at each interval start; the adjoint sweep
re-runs an interval's forward steps from its checkpoint when it enters that
interval, so adjoint storage is O(reinit interval) instead of O(steps).  The
kernels are deterministic, so the recomputed midpoint velocities and maps
are bitwise those of the first forward pass.
"""
from __future__ import annotations

import contextlib

import numpy as np
import torch

from . import _lib as L
from . import scenes, timing
from .ops import FlowMap, MacField, MGPCG, Ops


class _Scalar:
    """``field[None]`` access like a 0-d Taichi field."""

    def __init__(self, v=0):
        self.v = v

    def __getitem__(self, k):
        return self.v

    def __setitem__(self, k, v):
        self.v = v


class FieldView:
    """Taichi-field-like view of a device tensor (scalar field, or an SoA
    point set shown as AoS (..., dim) like ti.Vector.field)."""

    def __init__(self, owner, getter, soa=False):
        self._owner = owner
        self._get = getter
        self._soa = soa

    @property
    def tensor(self):
        return self._get(self._owner)

    @property
    def shape(self):
        t = self.tensor
        return tuple(t.shape[1:]) if self._soa else tuple(t.shape)

    def to_numpy(self):
        a = self.tensor.detach().cpu().numpy()
        return np.moveaxis(a, 0, -1).copy() if self._soa else a

    def to_torch(self, device=None):
        t = self.tensor
        return torch.movedim(t, 0, -1) if self._soa else t

    def from_numpy(self, arr):
        t = self.tensor
        src = torch.as_tensor(np.asarray(arr), dtype=t.dtype)
        if self._soa:
            src = torch.movedim(src, -1, 0)
        t.copy_(src.to(t.device))

    def fill(self, v):
        self.tensor.fill_(v)


def _parse_ctor(args):
    """(res_x, res_y, dx, dt, R, log_dir[, target]) or (res_x, res_y, res_z, dx, ...)."""
    if len(args) >= 3 and isinstance(args[2], (int, np.integer)):
        n = (int(args[0]), int(args[1]), int(args[2]))
        rest = args[3:]
    else:
        n = (int(args[0]), int(args[1]))
        rest = args[2:]
    dx, dt, R = float(rest[0]), float(rest[1]), int(rest[2])
    log_dir = rest[3] if len(rest) > 3 else None
    log_target_dir = rest[4] if len(rest) > 4 else None
    return n, dx, dt, R, log_dir, log_target_dir


class LFM_Simulator:
    """Forward leapfrog flow-map midpoint simulator (LFM_Simulator,
    2D/vortex/lfm_midpoint_simulator.py:187-661, 3D/lfm_midpoint_simulator.py:167-807)."""

    def __init__(self, *args, order=None, total_steps=None, n_passive=1, adj_kernel=None, mg_levels=0,
                 mg_pre_post=2, mg_bottom=50, tol=1e-6, max_iters=400, record=True, device="cuda",
                 mg_bf16_coarse=False, mg_tail=True, mg_bottom_exact=True, mg_check_every=1, checkpoint=False,
                 overlap_projections=True, control_num=0, BFECC_clamp=False, use_short_BFECC=False,
                 overlap_march=True, mg_fused_sweeps=2, mg_split_layout=True, mg_on_noconv="warn", mg_tail_level=0,
                 interp_kern=2, mg_graph=True, check_finite=True, side_march_blocks=2):
        n, dx, dt, R, log_dir, log_target_dir = _parse_ctor(args)
        self.n = n
        self.dim = len(n)
        self.res_x, self.res_y = n[0], n[1]
        if self.dim == 3:
            self.res_z = n[2]
        self.dx, self.dt, self.reinit_every = dx, dt, R
        self.log_dir, self.log_target_dir = log_dir, log_target_dir
        # RK_number: 2 in 2D, 4 in 3D (2D/vortex/hyperparameters.py:11, 3D/hyperparameters.py:281)
        self.order = order or (2 if self.dim == 2 else 4)
        # calculate_nabla_u_w kernel: quartic in 2D, cubic in 3D (q5)
        self.adj_kernel = adj_kernel or (4 if self.dim == 2 else 3)
        self.total_steps = total_steps
        self.n_passive = n_passive
        self.record = record or checkpoint
        self.checkpoint = checkpoint
        self._replaying = False
        self.device = torch.device(device)
        if not torch.cuda.is_available():
            raise RuntimeError("dfm.LFM_Simulator needs a HIP device (torch.cuda.is_available() is False)")
        self.ops = Ops(n, dx)
        self._mg_kw = dict(levels=mg_levels, pre_post=mg_pre_post, bottom=mg_bottom, tol=tol, max_iters=max_iters,
                           bf16_coarse=mg_bf16_coarse, tail=mg_tail, bottom_exact=mg_bottom_exact,
                           check_every=mg_check_every, fused_sweeps=mg_fused_sweeps,
                           split_layout=mg_split_layout, on_noconv=mg_on_noconv, tail_level=mg_tail_level,
                           graph=mg_graph)
        self.solver = MGPCG(n, dx, **self._mg_kw)
        # independent projections of one adjoint step run overlapped on a
        # second handle and stream
        self.overlap_projections = overlap_projections
        # NaN / Inf in the final forward state or in the adjoint raise
        # DfmError(DFM_E_NONFINITE) (the projections report theirs already)
        self.check_finite = check_finite
        self._nf_count = None
        self._solver2 = None
        self._stream2 = None
        # flow-map marches that only need the midpoint velocity run on a side
        # stream beside the (latency-bound) projections: True / "both",
        # "forward", "adjoint" or False
        self.overlap_march = overlap_march
        # marches on the side stream hold at most this many workgroups per CU
        # (dfm_march_capped; 0 = no cap), so that the projection's
        # one-workgroup coarse levels find a CU without waiting
        self.side_march_blocks = side_march_blocks
        self._side_active = False
        self._side = None
        self._side_done = None
        dev = self.device
        self.u = MacField.empty(n, dev, zero=True)
        self._u_new = MacField.empty(n, dev, zero=True)
        self.init_u = MacField.empty(n, dev, zero=True)
        self.err_u = MacField.empty(n, dev, zero=True)
        self.passive = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(n_passive)]
        self._passive_new = [torch.zeros_like(p) for p in self.passive]
        self.init_passive = [torch.zeros_like(p) for p in self.passive]
        self._err_s = torch.zeros(n, dtype=torch.float32, device=dev)
        self.phi_map = FlowMap(n, dev)
        self.psi_map = FlowMap(n, dev)
        self.step_num = _Scalar(-1)
        self.frame_idx = _Scalar(0)
        self.max_speed = _Scalar(0.0)
        self.mid_u = {}        # step -> MacField (midpoint velocity)
        self.passive_start = {}  # interval-start step -> [passive]
        self.passive_steps = {}  # step -> [passive], kept when step_midpoint(write_passive=True)
        self.flow_maps = {}    # interval index -> FlowMap (phi, F)
        self.u_start = {}      # interval index -> MacField (checkpoint mode)
        # released per-step storage, reused before anything new is allocated
        # (reserve() fills them up front, so a run allocates no device memory)
        self._pool_mid, self._pool_fm, self._pool_pass, self._pool_u = [], [], [], []
        self.boundary_mask_t = None  # obstacle mask on cells (None: no obstacles)
        self.boundary_vel_t = None   # boundary velocity, (dim, *n), or None for 0
        # Gaussian control forces (add_control_force, 2D/vortex/lfm_midpoint_simulator.py:271-278)
        self.control_num = int(control_num)
        if self.control_num > 0:
            N, d = self.control_num, self.dim
            self.control_f = MacField.empty(n, dev, zero=True)
            self._fc = torch.zeros((N, d), dtype=torch.float32, device=dev)
            self._fr = torch.zeros((N,), dtype=torch.float32, device=dev)
            self._fs = torch.zeros((N, d), dtype=torch.float32, device=dev)
        self._force_store = {}  # step -> (centres, radii, strengths, count) for checkpoint replays
        # B-spline order sampling the advected fields (impulse, passives and
        # their adjoints) in the BFECC pullbacks and the semi-Lagrangian
        # advection: 2 = N_2, the reference's; 1 = N_1 (linear) and 3 = N_3
        # (cubic) are optional variants.  The map march, the back-trace
        # velocity and accumulate_init stay quadratic.
        if interp_kern not in (1, 2, 3):
            raise ValueError(f"interp_kern must be 1, 2 or 3, not {interp_kern!r}")
        if BFECC_clamp and interp_kern != 2:
            raise ValueError("BFECC_clamp is implemented for the quadratic sampling kernel only (interp_kern=2)")
        self.interp_kern = int(interp_kern)
        # BFECC variants (hyperparameters BFECC_clamp / use_short_BFECC)
        self.bfecc_clamp = bool(BFECC_clamp)
        self.short_bfecc = bool(use_short_BFECC)
        self._clamp_tmp = None
        self._short_maps = None
        self._views()

    # Taichi-like attribute views -------------------------------------------
    def _views(self):
        names = ["u_x", "u_y", "u_z"][: self.dim]
        for a, nm in enumerate(names):
            setattr(self, nm, FieldView(self, lambda s, a=a: s.u.c[a]))
        if self.n_passive >= 1:
            self.passive1 = FieldView(self, lambda s: s.passive[0])
        self.psi_view = FieldView(self, lambda s: s.psi_map.psi, soa=True)
        self.phi_view = FieldView(self, lambda s: s.phi_map.psi, soa=True)
        if self.control_num > 0:
            for a, nm in enumerate(["control_force_x", "control_force_y", "control_force_z"][: self.dim]):
                setattr(self, nm, FieldView(self, lambda s, a=a: s.control_f.c[a]))

    # side stream ---------------------------------------------------------------
    @contextlib.contextmanager
    def _on_side(self, where="both"):
        """Queue the enclosed launches on the side stream (default priority),
        after everything the current stream has queued so far; the current
        stream waits for them at the next _join_side().  Without
        overlap_march the launches stay on the current stream."""
        if not self.overlap_march or (self.overlap_march != "both" and self.overlap_march is not True
                                      and self.overlap_march != where):
            yield
            return
        main = torch.cuda.current_stream()
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.device)
        ready = torch.cuda.Event()
        ready.record(main)
        with torch.cuda.stream(self._side):
            self._side.wait_event(ready)
            self._side_active = True
            try:
                yield
            finally:
                self._side_active = False
        done = torch.cuda.Event()
        done.record(self._side)
        self._side_done = done

    def _side_cap(self):
        """Workgroups-per-CU cap for a march queued now (only on the side stream)."""
        return self.side_march_blocks if self._side_active else 0

    def _join_side(self):
        if self._side_done is not None:
            torch.cuda.current_stream().wait_event(self._side_done)
            self._side_done = None

    # BFECC variants -----------------------------------------------------------
    def _bfecc_u(self, init, fwd, bwd, err, out):
        """BFECC (2D/vortex/flowmap.py:440-463), clamp_u'd when BFECC_clamp."""
        if not self.bfecc_clamp:
            self.ops.bfecc_u(init, fwd, bwd, err, out, kern=self.interp_kern)
            return
        if self._clamp_tmp is None:
            self._clamp_tmp = (MacField.empty(self.n, self.device), torch.empty(self.n, device=self.device))
        self.ops.bfecc_u_clamp(init, fwd, bwd, err, self._clamp_tmp[0], out)

    def _bfecc_s(self, init, psi, phi, err, out):
        """BFECC_scalar (2D/vortex/flowmap.py:465-481), clamp_u'd when BFECC_clamp."""
        if not self.bfecc_clamp:
            self.ops.bfecc_s(init, psi, phi, err, out, kern=self.interp_kern)
            return
        if self._clamp_tmp is None:
            self._clamp_tmp = (MacField.empty(self.n, self.device), torch.empty(self.n, device=self.device))
        self.ops.bfecc_s_clamp(init, psi, phi, err, self._clamp_tmp[1], out)

    def _one_step_maps(self, vel, dt):
        """use_short_BFECC's psi_tem / phi_tem: identity maps marched one step
        through vel over +dt and -dt."""
        if self._short_maps is None:
            self._short_maps = (FlowMap(self.n, self.device), FlowMap(self.n, self.device))
        fwd, bwd = self._short_maps
        for m, h in ((fwd, dt), (bwd, -dt)):
            self.ops.reset_identity(m)
            self.ops.march(vel, m, h, self.order)
        return fwd, bwd

    # control forces -----------------------------------------------------------
    def _load_forces(self, para):
        """Copy control_force_para {"c", "r", "s", "n"} to the device
        parameter buffers; returns the force count."""
        if self.control_num <= 0:
            raise ValueError("control_force_para needs a simulator built with control_num > 0 "
                             "(the reference's add_control_force / control_num)")
        count = int(para["n"])
        if not 0 <= count <= self.control_num:
            raise ValueError(f"force count {count} outside [0, control_num={self.control_num}]")
        d = self.dim
        for buf, key, shape in ((self._fc, "c", (-1, d)), (self._fr, "r", (-1,)), (self._fs, "s", (-1, d))):
            src = torch.as_tensor(para[key], dtype=torch.float32).reshape(shape)
            if src.shape[0] < count:
                raise ValueError(f"control_force_para[{key!r}] has {src.shape[0]} rows, fewer than n = {count}")
            buf[: src.shape[0]].copy_(src[: buf.shape[0]].to(self.device))
        return count

    def apply_gausion_force(self, *args):
        """apply_gausion_force(f_x, f_y[, f_z], center, inverse_radius,
        strength, num) (2D/vortex/lfm_midpoint_simulator.py:302-307): f = the
        Gaussian force field of the first `num` forces."""
        fields, (c, r, st, num) = args[: self.dim], args[self.dim:]
        count = self._load_forces({"c": c, "r": r, "s": st, "n": num})
        comps = [f.tensor if isinstance(f, FieldView) else f for f in fields]
        self.ops.gaussian_force(self._fc, self._fr, self._fs, count, MacField(comps))

    def _force_field(self, count):
        self.ops.gaussian_force(self._fc, self._fr, self._fs, count, self.control_f)

    # obstacles --------------------------------------------------------------
    def set_boundary_mask(self, mask, boundary_vel=None):
        """Obstacle cells (mask >= 1 solid for the field masks, > 0 for the
        projection, as in the reference) and their boundary velocity (cell
        array (..., dim) like the reference's boundary_vel, or None for 0);
        mask None removes the obstacles."""
        if mask is None:
            self.boundary_mask_t = None
            self.boundary_vel_t = None
            self.solver.set_mask(None)
            if self._solver2 is not None:
                self._solver2.set_mask(None)
            return
        m = torch.as_tensor(np.asarray(mask, dtype=np.float32)).to(self.device).contiguous()
        if tuple(m.shape) != self.n:
            raise ValueError(f"mask shape {tuple(m.shape)} != grid {self.n}")
        self.boundary_mask_t = m
        if self._solver2 is not None:
            self._solver2.set_mask(m)
        if boundary_vel is not None:
            bv = torch.as_tensor(np.asarray(boundary_vel, dtype=np.float32)).to(self.device)
            self.boundary_vel_t = torch.movedim(bv, -1, 0).contiguous()
        else:
            self.boundary_vel_t = None
        self.solver.set_mask(m)

    def _mask_fwd(self, u):
        """mask_velocity_by_boundary + mask_passive_by_boundary after a
        forward projection (2D/vortex/lfm_midpoint_simulator.py:429-433,543-547,612-616,632-636)."""
        if self.boundary_mask_t is None:
            return
        self.ops.mask_velocity(self.boundary_mask_t, u, self.boundary_vel_t)
        for p in self.passive:
            self.ops.mask_scalar(self.boundary_mask_t, p)

    # set-up -----------------------------------------------------------------
    def _initial_velocity(self, theta):
        if self.dim == 2:
            # sixteen_vortex_vel_func_with_pos_coef + split_central_vector on the device
            from .autograd import vortex_velocity_2d

            th = torch.as_tensor(np.asarray(theta, dtype=np.float32).reshape(-1)).to(self.device)
            with torch.no_grad():
                return list(vortex_velocity_2d(th, self.n, self.dx))
        uc = scenes.vortex_tubes_3d(np.asarray(theta, dtype=np.float64).reshape(-1, 8), self.n, self.dx)
        c = torch.as_tensor(np.ascontiguousarray(np.moveaxis(uc, -1, 0), dtype=np.float32)).to(self.device)
        u = MacField.empty(self.n, self.device)
        self.ops.split_central_vector(c, u)
        return u

    def init(self, theta=None, forward=True, target=False, init_passive_np=None, u0=None):
        """LFM_Simulator.init (2D/vortex/lfm_midpoint_simulator.py:370-439): build
        u0 (MAC arrays `u0`, or the vortex generator at `theta`), project it,
        set the passive scalars, reset the forward map and init_u = u."""
        if not forward:
            return
        if u0 is None:
            if theta is None:
                raise ValueError("init needs theta (vortex parameters) or u0 (MAC velocity arrays)")
            u0 = self._initial_velocity(theta)
        comps = u0.c if isinstance(u0, MacField) else u0
        for a in range(self.dim):
            self.u.c[a].copy_(torch.as_tensor(comps[a], dtype=torch.float32).to(self.device))
        self.solver.Poisson(self.u)
        if init_passive_np is not None:
            ps = init_passive_np if isinstance(init_passive_np, (list, tuple)) else [init_passive_np]
            for k, p in enumerate(ps[: self.n_passive]):
                self.passive[k].copy_(torch.as_tensor(p, dtype=torch.float32).to(self.device))
        for k in range(self.n_passive):
            self.init_passive[k].copy_(self.passive[k])
        self._mask_fwd(self.u)
        self.step_num[None] = -1
        self.frame_idx[None] = 0
        self.ops.reset_identity(self.phi_map)
        self.init_u.copy_(self.u)
        self._pool_mid.extend(self.mid_u.values())
        self._pool_fm.extend(self.flow_maps.values())
        self._pool_pass.extend(self.passive_start.values())
        self._pool_u.extend(self.u_start.values())
        self.mid_u.clear()
        self.passive_start.clear()
        self.flow_maps.clear()
        self.u_start.clear()
        self.passive_steps.clear()
        self._force_store.clear()

    @property
    def noconv_count(self):
        """Projections of this simulator that stopped at max_iters above
        tolerance (each also warned or raised, per ``mg_on_noconv``)."""
        return self.solver.noconv_count + (self._solver2.noconv_count if self._solver2 is not None else 0)

    def final_condition(self):
        return self.total_steps is not None and self.step_num[None] == self.total_steps - 1

    # forward step -------------------------------------------------------------
    def step_midpoint(self, write_flow_map=False, write_passive=False, control_force_para=None):
        """step_midpoint (2D/vortex/lfm_midpoint_simulator.py:472-515)."""
        self.step_num[None] += 1
        step = self.step_num[None]
        j = step % self.reinit_every
        if j == 0 and self.record and not self._replaying:
            if self._pool_pass:
                keep = self._pool_pass.pop()
                for dst, src in zip(keep, self.passive):
                    dst.copy_(src)
            else:
                keep = [p.clone() for p in self.passive]
            self.passive_start[step] = keep
            if self.checkpoint:
                if self._pool_u:
                    ku = self._pool_u.pop()
                    ku.copy_(self.u)
                else:
                    ku = self.u.clone()
                self.u_start[step // self.reinit_every] = ku
        if write_passive and self.record and not self._replaying:
            # the reference writes the passive of every step to disk (:487-491)
            self.passive_steps[step] = self.passive_start.get(step) or [p.clone() for p in self.passive]
        forces = None
        if control_force_para is not None:
            forces = self._load_forces(control_force_para)
            if self.checkpoint and not self._replaying:
                self._force_store[step] = (self._fc.clone(), self._fr.clone(), self._fs.clone(), forces)
        elif self._replaying and step in self._force_store:
            c, r, st, forces = self._force_store[step]
            self._fc.copy_(c)
            self._fr.copy_(r)
            self._fs.copy_(st)
        self.step_flow_map(j, self.dt, forces)
        return False, self.final_condition()

    def _mid_buffer(self, step):
        if step in self.mid_u:
            return self.mid_u[step]
        m = self._pool_mid.pop() if self._pool_mid else MacField.empty(self.n, self.device)
        self.mid_u[step] = m
        return m

    def _drop_mid(self, step):
        m = self.mid_u.pop(step, None)
        if m is not None:
            self._pool_mid.append(m)

    def reserve(self, steps):
        """Allocate up front the per-step storage a forward run of `steps`
        steps keeps (midpoint velocities, each interval's (phi, F) and
        passives), so that neither the run nor a later one of at most that
        length allocates device memory; init() returns the storage to the pool."""
        steps = int(steps)
        intervals = -(-steps // self.reinit_every)
        while len(self._pool_mid) + len(self.mid_u) < steps + 1:
            self._pool_mid.append(MacField.empty(self.n, self.device))
        # checkpoint mode keeps one interval's maps at a time (while the
        # adjoint sweep is inside it) and each interval's starting velocity
        want_fm = 1 if self.checkpoint else (intervals if self.record else 0)
        while len(self._pool_fm) + len(self.flow_maps) < want_fm:
            self._pool_fm.append(FlowMap(self.n, self.device))
        if self.record:
            while len(self._pool_pass) + len(self.passive_start) < intervals + 1:
                self._pool_pass.append([torch.empty_like(p) for p in self.passive])
        if self.checkpoint:
            while len(self._pool_u) + len(self.u_start) < intervals + 1:
                self._pool_u.append(MacField.empty(self.n, self.device))

    def step_flow_map(self, j, dt, forces=None):
        """step_flow_map (2D/vortex/lfm_midpoint_simulator.py:517-661); `forces`:
        the count of control forces loaded for this step, or None."""
        ops, order, step = self.ops, self.order, self.step_num[None]
        mid = self._mid_buffer(step)
        # half-step self-advection of u, then projection -> midpoint velocity
        ops.sl_advect(self.u, 0.5 * dt, order, [self.u], [mid], kern=self.interp_kern)
        if forces is not None:
            # half-step control force before the first projection (:527-537)
            self._force_field(forces)
            ops.axpy(mid, 0.5 * dt, self.control_f, mid)
        self.solver.Poisson(mid)
        self._mask_fwd(mid)
        # forward map phi/F marched with the midpoint velocity (dt sign flipped);
        # nothing before the reinitialisation reads it, so it runs beside the
        # advection and the projection below
        with self._on_side("forward"):
            ops.march(mid, self.phi_map, -dt, order, blocks_per_cu=self._side_cap())
        if self.short_bfecc:
            # BFECC from the step's start through one-step maps of the midpoint
            # velocity (use_short_BFECC, 2D/vortex/lfm_midpoint_simulator.py:560-583)
            fwd, bwd = self._one_step_maps(mid, dt)
            for k in range(self.n_passive):
                self._bfecc_s(self.passive[k], fwd.psi, bwd.psi, self._err_s, self._passive_new[k])
            self._bfecc_u(self.u, fwd, bwd, self.err_u, self._u_new)
        else:
            # passive and full-step velocity advected by the midpoint velocity
            ops.sl_advect(mid, dt, order, [self.u], [self._u_new], self.passive, self._passive_new,
                          kern=self.interp_kern)
        self.u, self._u_new = self._u_new, self.u
        self.passive, self._passive_new = self._passive_new, self.passive
        if forces is not None:
            # full-step force and accumulate_init through (F, phi) (:597-602);
            # init_u only feeds the reinitialisation, which a replay skips
            ops.axpy(self.u, dt, self.control_f, self.u)
            if not self._replaying:
                with self._on_side("forward"):  # after the march
                    ops.pullback_u(self.control_f, self.phi_map, self.init_u, alpha=dt, beta=1.0, add=self.init_u)
        self.solver.Poisson(self.u)
        self._mask_fwd(self.u)
        self._join_side()
        if j == self.reinit_every - 1:
            if self._replaying:
                # recomputation for the adjoint only needs the interval's (phi, F)
                self._store_phi(step)
            else:
                self._reinit(step, j, dt)

    def _store_phi(self, step):
        fm = self._pool_fm.pop() if self._pool_fm else FlowMap(self.n, self.device)
        fm.copy_(self.phi_map)
        self.flow_maps[step // self.reinit_every] = fm

    def _reinit(self, step, j, dt):
        """Reinitialisation: re-march the backward map psi/T from identity
        through the stored midpoint velocities in reverse (the last step uses
        the end-of-step velocity, q3; backtrack_psi_grid,
        2D/vortex/lfm_midpoint_simulator.py:338-353), BFECC pullback of init_u,
        project, BFECC of the passives, store (phi, F), reset."""
        ops, order = self.ops, self.order
        ops.reset_identity(self.psi_map)
        ops.march(self.u, self.psi_map, dt, order)
        for st in reversed(range(step - j, step)):
            ops.march(self.mid_u[st], self.psi_map, dt, order)
        self._bfecc_u(self.init_u, self.psi_map, self.phi_map, self.err_u, self.u)
        self.solver.Poisson(self.u)
        self._mask_fwd(self.u)
        for k in range(self.n_passive):
            self._bfecc_s(self.init_passive[k], self.psi_map.psi, self.phi_map.psi, self._err_s, self.passive[k])
            self.init_passive[k].copy_(self.passive[k])
        if self.record and not self.checkpoint:
            self._store_phi(step)
        else:
            for st in range(step - j, step + 1):
                self._drop_mid(st)
        ops.reset_identity(self.phi_map)
        self.init_u.copy_(self.u)

    # the reference's file layout ------------------------------------------------
    def export_cache(self, log_dir=None):
        """Write the stored run in the reference's on-disk layout under
        ``log_dir`` (default: the constructor's log_dir): midpoint velocities
        (midpoint/u_*), each interval's forward map (flowmap_phi, flowmap_phi_xy
        or _xyz, flowmap_F), the passives at the stored steps (passive/s1_*
        or s_*) and the final state (final/u_x_k ..., final_ind_0, s1_k), as
        RW_Cache / write_disk write them (2D/vortex/lfm_midpoint_simulator.py:46-184,
        472-515,617-643).  Returns the directory."""
        import os

        from . import rw_cache as W

        d = log_dir or self.log_dir
        if d is None:
            raise ValueError("export_cache needs a directory (log_dir)")
        lay = W.layout(self.dim, self.n_passive)
        for key in lay:
            os.makedirs(os.path.join(d, lay[key][0]), exist_ok=True)
        path = {k: os.path.join(d, v[0]) for k, v in lay.items()}
        n, cache = self.n, W.RWCache()
        aos = lambda t: np.moveaxis(t.detach().cpu().numpy(), 0, -1)  # noqa: E731
        for st in sorted(self.mid_u):
            cache.write(W.pack_mac(self.mid_u[st].numpy(), n), "u", path["midpoint"], st)
        cache.force_write("u", path["midpoint"])
        xy = lay["phi_xy"][1]
        for m in sorted(self.flow_maps):
            fm = self.flow_maps[m]
            cache.write(aos(fm.psi), "phi", path["phi"], m)
            cache.write(W.pack_face_vectors([aos(t) for t in fm.psi_f], n), xy, path["phi_xy"], m)
            cache.write(W.pack_face_vectors([aos(t) for t in fm.T_f], n), "F", path["F"], m)
        for name, key in (("phi", "phi"), (xy, "phi_xy"), ("F", "F")):
            cache.force_write(name, path[key])
        # the series is read back by position (entry j in slot j - first), so
        # it holds every step: the per-step copies kept under write_passive,
        # else the interval-start copies, else zeros
        last = max(list(self.passive_steps) + list(self.passive_start), default=-1)
        for st in range(last + 1):
            src = self.passive_steps.get(st, self.passive_start.get(st))
            for k, nm in enumerate(lay["passive"][1]):
                arr = src[k].cpu().numpy() if src is not None else np.zeros(n, np.float32)
                cache.write(W.as_channels(arr, self.dim), nm, path["passive"], st)
        for nm in lay["passive"][1]:
            cache.force_write(nm, path["passive"])
        k = self.step_num[None] + 1
        for a, c in enumerate(self.u.numpy()):
            W.write_disk(c, "u_" + "xyz"[a], path["final"], k)
        W.write_disk(int(k), "final_ind", path["final"], 0)
        for p, nm in zip(self.passive, lay["passive"][1]):
            W.write_disk(p.cpu().numpy(), nm, path["final"], k)
        return d

    def import_cache(self, log_dir=None):
        """Load a run stored in the reference's on-disk layout (``export_cache``
        or the reference's own forward pass with write_flow_map and
        write_passive) into device storage and the final state, so that
        init_gradient / backtrack_step_midpoint can run the adjoint sweep from
        it (init_gradient's read_disk of final/, 2D/vortex/lfm_midpoint_diff_simulator.py:126-207,
        and the per-step reads of backtrack_flow_map)."""
        import os

        from . import rw_cache as W

        if self.checkpoint:
            raise ValueError("import_cache fills stored-everything storage; construct with checkpoint=False")
        d = log_dir or self.log_dir
        lay = W.layout(self.dim, self.n_passive)
        path = {k: os.path.join(d, v[0]) for k, v in lay.items()}
        n, dev, cache, R = self.n, self.device, W.RWCache(), self.reinit_every
        final_ind = int(W.read_disk("final_ind", path["final"], 0)[0])
        for a in range(self.dim):
            self.u.c[a].copy_(torch.as_tensor(W.read_disk("u_" + "xyz"[a], path["final"], final_ind)).to(dev))
        for p, nm in zip(self.passive, lay["passive"][1]):
            p.copy_(torch.as_tensor(np.asarray(W.read_disk(nm, path["final"], final_ind)).reshape(n)).to(dev))
        self.mid_u.clear()
        self.flow_maps.clear()
        self.passive_start.clear()
        soa = lambda arr: torch.as_tensor(np.ascontiguousarray(np.moveaxis(arr, -1, 0))).to(dev)  # noqa: E731
        for st in range(final_ind):
            comps = W.unpack_mac(cache.read("u", path["midpoint"], st), n)
            self.mid_u[st] = MacField([torch.as_tensor(c).to(dev) for c in comps])
        xy = lay["phi_xy"][1]
        for m in range(final_ind // R):
            fm = FlowMap(n, dev)
            fm.psi.copy_(soa(np.asarray(cache.read("phi", path["phi"], m))[tuple(slice(0, s) for s in n)]))
            for dst, name, key in ((fm.psi_f, xy, "phi_xy"), (fm.T_f, "F", "F")):
                for a, v in enumerate(W.unpack_face_vectors(cache.read(name, path[key], m), n)):
                    dst[a].copy_(soa(v))
            self.flow_maps[m] = fm
        for st in range(0, final_ind, R):
            self.passive_start[st] = [
                torch.as_tensor(np.ascontiguousarray(np.asarray(cache.read(nm, path["passive"], st))[..., 0])).to(dev)
                for nm in lay["passive"][1]]
        self.step_num[None] = final_ind - 1
        self.init_u.copy_(self.u)

    # diagnostics ---------------------------------------------------------------
    def calc_max_speed(self, *u):
        """calc_max_speed (2D/vortex/lfm_midpoint_simulator.py:328-335): max
        cell-centre speed (at least 1e-3) of the given MAC components, or of u,
        into ``max_speed[None]``; reads one scalar back.  step_midpoint does not
        call it (the reference only prints the CFL number from it)."""
        m = MacField([f.tensor if isinstance(f, FieldView) else f for f in u]) if u else self.u
        out = torch.empty(1, dtype=torch.float32, device=self.device)
        self.ops.max_speed(m, out)
        self.max_speed[None] = float(out.item())
        return self.max_speed[None]

    def get_central_vector(self, u=None):
        """get_central_vector: (dim, *n) cell-centred average of u (default the
        current velocity), on the device."""
        u = self.u if u is None else u
        c = torch.empty((self.dim, *self.n), dtype=torch.float32, device=self.device)
        self.ops.central_vector(u, c)
        return c

    def calculate_w(self):
        """Vorticity at cell centres (get_central_vector + curl,
        2D/vortex/taichi_utils.py:152-156,198-205); the curl is a host-side diagnostic."""
        if self.dim == 2:
            uc = self.get_central_vector().double()
            pad = torch.nn.functional.pad(uc[None], (1, 1, 1, 1), mode="replicate")[0]
            dvy_dx = (pad[1, 2:, 1:-1] - pad[1, :-2, 1:-1]) / (2 * self.dx)
            dvx_dy = (pad[0, 1:-1, 2:] - pad[0, 1:-1, :-2]) / (2 * self.dx)
            return (dvy_dx - dvx_dy).cpu().numpy()
        raise NotImplementedError("3D vorticity diagnostic is not provided")


class LFM_Diff_Simulator(LFM_Simulator):
    """Forward simulator plus the continuous adjoint sweep on flow maps
    (LFM_Diff_Simulator, 2D/vortex/lfm_midpoint_diff_simulator.py:17-503,
    3D/lfm_midpoint_diff_simulator.py:15-594)."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        n, dev = self.n, self.device
        self.adj_u = MacField.empty(n, dev, zero=True)
        self._adj_u_new = MacField.empty(n, dev, zero=True)
        self.init_adj_u = MacField.empty(n, dev, zero=True)
        self.err_adj_u = MacField.empty(n, dev, zero=True)
        self.r_u = MacField.empty(n, dev, zero=True)     # ajd_uw_source
        self.r_s = MacField.empty(n, dev, zero=True)     # sum of adj_passive_f
        self.ab_h1 = MacField.empty(n, dev, zero=True)   # adj_u_*_AB_tem
        self.ab_h2 = MacField.empty(n, dev, zero=True)   # adj_u_*_AB_tem2
        self.adj_passive = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(self.n_passive)]
        self._adj_passive_new = [torch.zeros_like(p) for p in self.adj_passive]
        self.init_adj_passive = [torch.zeros_like(p) for p in self.adj_passive]
        self.target_passive = [torch.zeros_like(p) for p in self.adj_passive]
        if self.control_num > 0:
            # calculate_adj_gausion_force outputs (2D/shape/lfm_midpoint_diff_simulator.py:58-64)
            N, d = self.control_num, self.dim
            self._force_grads = {k: torch.zeros((N, d), dtype=torch.float32, device=dev)
                                 for k in ("adj_center", "adj_strength", "pen_center", "pen_strength")}
            for nm, k in (("adj_gaussion_force_center", "adj_center"), ("adj_gaussion_force_strength", "adj_strength"),
                          ("penalty_gaussion_force_center", "pen_center"),
                          ("penalty_gaussion_force_strength", "pen_strength")):
                setattr(self, nm, FieldView(self, lambda s, k=k: s._force_grads[k]))
        self.loss = 0.0
        self.final_ind = None
        self._interval_targets = None  # step -> [target per passive], or a callable; see set_interval_targets
        names = ["adj_u_x", "adj_u_y", "adj_u_z"][: self.dim]
        for a, nm in enumerate(names):
            setattr(self, nm, FieldView(self, lambda s, a=a: s.adj_u.c[a]))
        if self.n_passive >= 1:
            self.adj_passive1 = FieldView(self, lambda s: s.adj_passive[0])

    def forward_step_midpoint(self, write_flow_map=False, write_passive=False, control_force_para=None):
        return self.step_midpoint(write_flow_map, write_passive, control_force_para)

    def set_target(self, target):
        ts = target if isinstance(target, (list, tuple)) else [target]
        for k, t in enumerate(ts[: self.n_passive]):
            self.target_passive[k].copy_(torch.as_tensor(t, dtype=torch.float32).to(self.device))

    def set_interval_targets(self, targets):
        """Intermediate density targets at reinitialisation-interval starts
        (2D/vortex/lfm_midpoint_diff_simulator.py:478-497, 3D/lfm_midpoint_diff_simulator.py:586-591):
        when the adjoint sweep reaches interval start `step`, lam_s += s - s*
        (gradient factor 1, as the reference has it) and loss += sum (s* - s)^2.
        `targets`: {step: array or [array per passive]}, a callable step ->
        that or None, or None to turn the terms off (the default: final-time
        loss only)."""
        self._interval_targets = targets

    def interval_targets_from_dir(self, log_target_dir=None):
        """Targets read from a target run's passive series in the reference's
        layout (<log_target_dir>/passive, as the reference's read of
        "target_s*" finds them); returns a callable for set_interval_targets."""
        import os

        from . import rw_cache as W

        d = os.path.join(log_target_dir or self.log_target_dir, "passive")
        names = W.layout(self.dim, self.n_passive)["passive"][1]
        cache = W.RWCache()
        return lambda step: [np.asarray(cache.read(nm, d, step))[..., 0] for nm in names]

    def _interval_target(self, step):
        t = self._interval_targets
        if t is None:
            return None
        t = t(step) if callable(t) else t.get(step)
        if t is None:
            return None
        return t if isinstance(t, (list, tuple)) else [t]

    def init_gradient(self, theta=None, target=None, grad_passive=None, grad_u=None):
        """init_gradient (2D/vortex/lfm_midpoint_diff_simulator.py:126-207):
        terminal adjoint lam_s = 2 (s - s*), lam_u = 0, loss = sum (s - s*)^2,
        starting from the final forward state held on the device.

        For a general final-time loss the caller passes dL/ds_T
        (``grad_passive``, one tensor per passive) and/or dL/du_T (``grad_u``,
        MAC tensors); lam_u starts at P grad_u, the adjoint of the final
        projection, and ``loss`` is not computed."""
        if target is not None:
            self.set_target(target)
        self._check_finite(list(self.u.c) + list(self.passive), "init_gradient: the final forward state")
        self.final_ind = self.step_num[None] + 1
        if self.final_ind % self.reinit_every:
            raise ValueError("the adjoint sweep needs a whole number of reinitialisation intervals "
                             f"({self.final_ind} steps, reinit_every={self.reinit_every}): the forward map of "
                             "an interval is stored at its reinitialisation")
        if grad_u is None:
            self.adj_u.zero_()
        else:
            for a in range(self.dim):
                self.adj_u.c[a].copy_(grad_u[a])
            self.solver.Poisson_adj(self.adj_u)
        if grad_passive is None:
            loss = 0.0
            for k in range(self.n_passive):
                d = self.passive[k] - self.target_passive[k]
                self.adj_passive[k].copy_(2.0 * d)
                loss += float(torch.sum(d.double() ** 2))
            self.loss = loss
        else:
            for k in range(self.n_passive):
                self.adj_passive[k].copy_(grad_passive[k])
            self.loss = None
        self._mask_adj()
        self.step_num[None] = self.final_ind
        self.init_adj_u.copy_(self.adj_u)
        for k in range(self.n_passive):
            self.init_adj_passive[k].copy_(self.adj_passive[k])
        self.init_u.copy_(self.u)
        for k in range(self.n_passive):
            self.init_passive[k].copy_(self.passive[k])
        self.ops.reset_identity(self.psi_map)
        self.ab_h1.zero_()
        self.ab_h2.zero_()

    def backtrack_step_midpoint(self, write_gradient=False, write_passive=False, control_force_para=None):
        """backtrack_step_midpoint (2D/vortex/lfm_midpoint_diff_simulator.py:220-258)."""
        if control_force_para is not None and self.control_num <= 0:
            self._load_forces(control_force_para)  # raises: no control forces configured
        self.step_num[None] -= 1
        j = self.step_num[None] % self.reinit_every
        self.backtrack_flow_map(j, self.dt, control_force_para)
        if self.step_num[None] == 0:
            self._check_finite(list(self.adj_u.c) + list(self.adj_passive), "backtrack_step_midpoint: the adjoint")
        return False, self.step_num[None] == 0

    def _check_finite(self, tensors, what):
        """Raise DfmError(DFM_E_NONFINITE) when any of the tensors holds a NaN
        or an Inf (one device count and one small readback, at points where
        the sweep is read back anyway: the start and the end of the adjoint)."""
        if not self.check_finite:
            return
        if self._nf_count is None:
            self._nf_count = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.ops.count_nonfinite(tensors, self._nf_count)
        bad = int(self._nf_count.item())
        if bad:
            raise L.DfmError(L.DFM_E_NONFINITE, f"{what} holds {bad} non-finite values")

    def backtrack_flow_map(self, j, dt, control_force_para=None):
        """backtrack_flow_map (2D/vortex/lfm_midpoint_diff_simulator.py:260-503)
        with use_short_BFECC=False and backward_u=True; with control forces as
        2D/shape/lfm_midpoint_diff_simulator.py:207-343 adds them."""
        ops, order, step = self.ops, self.order, self.step_num[None]
        if step not in self.mid_u:
            if not self.checkpoint:
                raise KeyError(f"no stored midpoint velocity for step {step} (forward run with record=False)")
            self._replay_interval(step // self.reinit_every)
        mid = self.mid_u[step]
        forces = None
        if control_force_para is not None:
            # force field for the backward_u re-march, and the parameter
            # gradients from lam_u as the step begins (:209-213)
            forces = self._load_forces(control_force_para)
            self._force_field(forces)
            for t in self._force_grads.values():
                t.zero_()
            g = self._force_grads
            ops.gaussian_force_adjoint(self._fc, self._fr, self._fs, forces, self.adj_u, g["adj_center"],
                                       g["adj_strength"], g["pen_center"], g["pen_strength"])
        # adjoint sources at the grid faces, from the current (end-of-step)
        # state; they only read fields that the advection below reads too and
        # write r_u / r_s, which the axpy after it consumes, so they run on the
        # side stream beside the advection (the current stream waits for them
        # before the axpy)
        with self._on_side("adjoint"):
            ops.adj_conv_source(self.u, self.adj_u, self.adj_kernel, self.r_u)
            ops.adj_scalar_source(self.passive, self.adj_passive, self.r_s)
        if self.short_bfecc:
            # BFECC of lam_u, lam_s, s and u through one-step maps of mid; the
            # convective source then enters at -2 dt
            # (use_short_BFECC, 2D/vortex/lfm_midpoint_diff_simulator.py:283-338)
            fwd, bwd = self._one_step_maps(mid, -dt)
            self._bfecc_u(self.adj_u, fwd, bwd, self.err_adj_u, self._adj_u_new)
            for k in range(self.n_passive):
                self._bfecc_s(self.adj_passive[k], fwd.psi, bwd.psi, self._err_s, self._adj_passive_new[k])
                self._bfecc_s(self.passive[k], fwd.psi, bwd.psi, self._err_s, self._passive_new[k])
            self._bfecc_u(self.u, fwd, bwd, self.err_adj_u, self._u_new)
            src_dt = -2.0 * dt
        else:
            # one back-trace through mid over -dt serves lam_u, u, s and lam_s
            ops.sl_advect(mid, -dt, order, [self.adj_u, self.u], [self._adj_u_new, self._u_new],
                          self.passive + self.adj_passive, self._passive_new + self._adj_passive_new,
                          kern=self.interp_kern)
            src_dt = -dt
        self.adj_u, self._adj_u_new = self._adj_u_new, self.adj_u
        self.u, self._u_new = self._u_new, self.u
        self.passive, self._passive_new = self._passive_new, self.passive
        self.adj_passive, self._adj_passive_new = self._adj_passive_new, self.adj_passive
        # lam_u += src_dt r_u - dt r_s  (2D/vortex/lfm_midpoint_diff_simulator.py:297-298,359-374)
        self._join_side()
        ops.axpy(self.adj_u, src_dt, self.r_u, self.adj_u, beta=-dt, z=self.r_s)
        # psi/T march and source accumulation (AB3 at -2dt for r_u, -dt for r_s);
        # they write only psi/T, init_adj_u (init_u) and the AB3 history, so
        # they run on the side stream beside the two projections
        with self._on_side("adjoint"):
            ops.march(mid, self.psi_map, dt, order, blocks_per_cu=self._side_cap())
            ops.accumulate(self.r_u, self.psi_map, self.init_adj_u, self.ab_h1, self.ab_h2,
                           self.reinit_every - 1 - j, -2.0 * dt, f2=self.r_s, dt2=-dt)
            if forces is not None:
                # the step's force accumulated through (T, psi) (:296-300)
                ops.pullback_u(self.control_f, self.psi_map, self.init_u, alpha=-dt, beta=1.0, add=self.init_u)
        if forces is not None:
            ops.axpy(self.u, -dt, self.control_f, self.u)  # u takes the step's force back
        self._project_pair(self.adj_u, self.u)
        self._mask_adj()
        self._join_side()
        if j == 0:
            fwd = self.flow_maps[step // self.reinit_every]
            # BFECC of lam_u and u, then their two (independent) projections
            self._bfecc_u(self.init_adj_u, fwd, self.psi_map, self.err_adj_u, self.adj_u)
            self._bfecc_u(self.init_u, fwd, self.psi_map, self.err_adj_u, self.u)
            self._project_pair(self.adj_u, self.u)
            self.init_adj_u.copy_(self.adj_u)
            self.init_u.copy_(self.u)
            self._mask_adj()
            for k in range(self.n_passive):
                if self.control_num > 0:
                    # with control forces the passive is rebuilt through the
                    # interval's maps (2D/shape/lfm_midpoint_diff_simulator.py:339-343)
                    self._bfecc_s(self.init_passive[k], fwd.psi, self.psi_map.psi, self._err_s, self.passive[k])
                else:
                    self.passive[k].copy_(self.passive_start[step][k])
                self.init_passive[k].copy_(self.passive[k])
                self._bfecc_s(self.init_adj_passive[k], fwd.psi, self.psi_map.psi, self._err_s, self.adj_passive[k])
            targets = self._interval_target(step)
            if targets is not None:
                # intermediate targets: lam_s += s - s*, loss += sum (s* - s)^2
                for k, tk in enumerate(targets[: self.n_passive]):
                    d = self.passive[k] - torch.as_tensor(np.asarray(tk), dtype=torch.float32).to(self.device)
                    self.adj_passive[k].add_(d)
                    if self.loss is not None:
                        self.loss += float(torch.sum(d.double() ** 2))
            for k in range(self.n_passive):
                self.init_adj_passive[k].copy_(self.adj_passive[k])
            self.ops.reset_identity(self.psi_map)
            if self.checkpoint:
                m = step // self.reinit_every
                fm = self.flow_maps.pop(m, None)
                if fm is not None:
                    self._pool_fm.append(fm)
                for st in range(step, step + self.reinit_every):
                    self._drop_mid(st)

    # checkpoint recomputation ------------------------------------------------
    _REPLAY_FIELDS = ("u", "_u_new", "passive", "_passive_new", "phi_map")

    def _replay_interval(self, m):
        """Re-run the forward steps of interval m from its checkpoint (velocity
        and passives at the interval start) on scratch state, leaving the
        interval's midpoint velocities and forward map (phi, F) in storage."""
        if m not in self.u_start:
            raise KeyError(f"no checkpoint for interval {m}")
        if not hasattr(self, "_rp"):
            n, dev = self.n, self.device
            self._rp = {"u": MacField.empty(n, dev), "_u_new": MacField.empty(n, dev),
                        "passive": [torch.empty_like(p) for p in self.passive],
                        "_passive_new": [torch.empty_like(p) for p in self.passive],
                        "phi_map": FlowMap(n, dev)}
        R = self.reinit_every
        start = m * R
        saved = {k: getattr(self, k) for k in self._REPLAY_FIELDS}
        saved_step = self.step_num[None]
        for k, v in self._rp.items():
            setattr(self, k, v)
        try:
            self.u.copy_(self.u_start[m])
            for k in range(self.n_passive):
                self.passive[k].copy_(self.passive_start[start][k])
            self.ops.reset_identity(self.phi_map)
            self.step_num[None] = start - 1
            self._replaying = True
            for _ in range(R):
                self.step_midpoint()
        finally:
            self._replaying = False
            self._rp = {k: getattr(self, k) for k in self._REPLAY_FIELDS}
            for k, v in saved.items():
                setattr(self, k, v)
            self.step_num[None] = saved_step

    def _project_pair(self, first, second):
        """Poisson_adj(first) and Poisson(second) (2D/vortex/lfm_midpoint_diff_simulator.py:391,403),
        which do not depend on each other: queued on two MGPCG handles and
        streams before either waits, so their latency-bound kernels overlap.
        The results are those of the two projections in sequence."""
        if not self.overlap_projections:
            self.solver.Poisson_adj(first)
            self.solver.Poisson(second)
            return
        if self._solver2 is None:
            self._solver2 = MGPCG(self.n, self.dx, **self._mg_kw)
            if self.boundary_mask_t is not None:
                self._solver2.set_mask(self.boundary_mask_t)
            self._stream2 = torch.cuda.Stream(device=self.device)
        main = torch.cuda.current_stream()
        with timing.region("mgpcg", calls=2):
            ready = torch.cuda.Event()
            ready.record(main)
            self.solver.Poisson_begin(first)
            with torch.cuda.stream(self._stream2):
                self._stream2.wait_event(ready)
                self._solver2.Poisson_begin(second)
            self.solver.Poisson_end()
            self._solver2.Poisson_end()
            done = torch.cuda.Event()
            done.record(self._stream2)
            main.wait_event(done)

    def _mask_adj(self):
        """The forward masks on u and the passives plus
        mask_adj_velocity_by_boundary / mask_adj_passive_by_boundary on the
        adjoint fields (2D/vortex/lfm_midpoint_diff_simulator.py:179-189,405-413,451-459)."""
        if self.boundary_mask_t is None:
            return
        m = self.boundary_mask_t
        self.ops.mask_velocity(m, self.u, self.boundary_vel_t)
        self.ops.mask_velocity(m, self.adj_u, None)
        for p in self.passive:
            self.ops.mask_scalar(m, p)
        for p in self.adj_passive:
            self.ops.mask_scalar(m, p, adjoint=True)

    def calculate_gw(self):
        u, self.u = self.u, self.adj_u
        try:
            return self.calculate_w()
        finally:
            self.u = u
