"""Many independent forward+adjoint problems of uneven size on one GPU.

At the sizes of the batched configuration (2D, 128^2 .. 512^2) every kernel of
a step is short and latency bound, and an MGPCG solve waits on the host once
for its convergence flag; run one after another, the GPU idles in each of
those gaps.  Here the instances are spread over W worker threads, each with
its own HIP stream (default priority), and every instance keeps its own
simulator and MGPCG handle, so the hardware queues hold other instances'
kernels while one instance's host thread waits.  Per instance the kernels and
results are those of the single-instance path: the batch is bitwise equal to
running the instances one by one.

``autotune`` times a probe run of the batch at several worker counts and
keeps the fastest; ``autotune_tail`` picks, per grid size, how many of the
coarse MGPCG levels run in the one-workgroup LDS tail.  How many hardware
queues HIP gives the process is fixed when the GPU is first used, by
GPU_MAX_HW_QUEUES (4 by default here); ``configure_hw_queues`` sets it
beforehand.
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

SIZES = (128, 192, 256, 384, 512)


def configure_hw_queues(n=8):
    """Ask HIP for n hardware queues (1..32); effective only before the first
    GPU use in this process."""
    n = max(1, min(32, int(n)))
    os.environ["GPU_MAX_HW_QUEUES"] = str(n)


def make_problems(count=32, seed=100, steps=24, reinit_every=12, sizes=SIZES):
    """Seeded uneven 2D problems: grid n from `sizes`, 4-12 vortices, a disk
    density and a 3-blob Gaussian target per instance (instance i uses seed
    seed + i)."""
    from . import scenes

    out = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        n = int(rng.choice(sizes))
        k = int(rng.integers(4, 13))
        dx = 1.0 / n
        out.append({
            "n": (n, n), "dx": dx, "dt": 0.5 * dx, "R": reinit_every, "steps": steps, "seed": seed + i,
            "theta": scenes.vortex_params_2d(seed + i, k).astype(np.float32),
            "density": scenes.disk_density((n, n), dx).astype(np.float32),
            "target": scenes.gaussian_blobs((n, n), dx, seed + 1000 + i, 3).astype(np.float32),
        })
    return out


class BatchedLFM:
    """Forward + adjoint of a list of problems (see ``make_problems``); each
    instance yields (loss, dL/dtheta)."""

    def __init__(self, problems, workers=4, device="cuda", **sim_kw):
        self.problems = list(problems)
        self.device = device
        # the instances already fill the hardware queues; a per-instance side
        # stream for the marches only adds false dependencies between them
        # (measured: 462 vs 568 instance-steps/s), the overlapped projection
        # pair still helps (568 vs 544)
        self.sim_kw = dict(tol=1e-6, max_iters=200, overlap_march=False)
        self.sim_kw.update(sim_kw)
        self.sims = {}
        self.set_workers(workers)

    def set_workers(self, workers):
        import torch

        self.workers = max(1, int(workers))
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.workers)]
        # longest-processing-time assignment of instances to workers (cost ~ cells)
        order = sorted(range(len(self.problems)), key=lambda i: -float(np.prod(self.problems[i]["n"])))
        load = [0.0] * self.workers
        self.assign = [[] for _ in range(self.workers)]
        for i in order:
            w = int(np.argmin(load))
            self.assign[w].append(i)
            load[w] += float(np.prod(self.problems[i]["n"]))

    def _sim(self, i):
        from .simulator import LFM_Diff_Simulator

        sim = self.sims.get(i)
        if sim is None:
            p = self.problems[i]
            sim = LFM_Diff_Simulator(*p["n"], p["dx"], p["dt"], p["R"], None, total_steps=p["steps"],
                                     device=self.device, **self.sim_kw)
            self.sims[i] = sim
        return sim

    def _run_instance(self, i, steps):
        from .autograd import calculate_dtheta

        p = self.problems[i]
        steps = p["steps"] if steps is None else steps
        sim = self._sim(i)
        sim.total_steps = steps
        sim.init(theta=p["theta"], init_passive_np=p["density"])
        for _ in range(steps):
            sim.step_midpoint()
        sim.init_gradient(target=p["target"])
        for _ in range(steps):
            sim.backtrack_step_midpoint()
        return sim.loss, np.asarray(calculate_dtheta(sim, p["theta"]), dtype=np.float64)

    def _run_worker(self, w, steps, results):
        import torch

        with torch.cuda.stream(self.streams[w]):
            for i in self.assign[w]:
                results[i] = self._run_instance(i, steps)
            self.streams[w].synchronize()

    def run(self, steps=None):
        """Run every instance (`steps` forward steps, or each problem's own,
        the terminal adjoint, as many adjoint steps, the theta gradient);
        returns (results, seconds)."""
        import torch

        torch.cuda.synchronize()
        results = [None] * len(self.problems)
        t0 = time.perf_counter()
        if self.workers == 1:
            self._run_worker(0, steps, results)
        else:
            with ThreadPoolExecutor(max_workers=self.workers) as ex:
                for f in [ex.submit(self._run_worker, w, steps, results) for w in range(self.workers)]:
                    f.result()
        torch.cuda.synchronize()
        return results, time.perf_counter() - t0

    def instance_steps(self, steps=None):
        """Forward+adjoint step pairs in one run of the batch."""
        return sum(p["steps"] if steps is None else steps for p in self.problems)

    def autotune_tail(self, solves=6):
        """Per grid size, the first MGPCG level kept in the one-workgroup LDS
        tail (the levels below it run as global kernels): every candidate
        from the largest suffix that fits in LDS up to "no tail" is timed on
        `solves` projections of one instance's initial velocity, and the
        fastest is set on every instance of that size.  Returns
        {n: {level: ms per projection}}."""
        import math

        import torch

        from . import _lib as L
        from .autograd import vortex_velocity_2d
        from .ops import MacField

        first = {}
        for i, p in enumerate(self.problems):
            first.setdefault(p["n"], i)
        out = {}
        for n, i in first.items():
            p = self.problems[i]
            solver = self._sim(i).solver
            levels = max(1, int(math.log2(min(n))) - 2)
            solver.set_option(L.MG_TAIL_LEVEL, 0)
            lo = solver.get_option(L.MG_TAIL_LEVEL)
            with torch.no_grad():
                u0 = [c.clone() for c in vortex_velocity_2d(torch.tensor(p["theta"], device=self.device), n, p["dx"])]
            times = {}
            for lvl in range(lo, levels + 1):
                solver.set_option(L.MG_TAIL_LEVEL, lvl)
                solver.Poisson(MacField([c.clone() for c in u0]))
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(solves):
                    solver.Poisson(MacField([c.clone() for c in u0]))
                torch.cuda.synchronize()
                times[lvl] = (time.perf_counter() - t0) / solves * 1e3
            best = min(times, key=times.get)
            for j, q in enumerate(self.problems):
                if q["n"] == n:
                    self._sim(j).solver.set_option(L.MG_TAIL_LEVEL, best)
            out[n] = times
        return out

    def autotune(self, candidates=(1, 2, 4, 8, 16), probe_steps=None):
        """Time one probe run per worker count (one reinitialisation interval
        per instance by default; a first untimed run at each count builds the
        handles and graphs) and keep the fastest; returns {workers: seconds}."""
        probe = probe_steps or min(p["R"] for p in self.problems)
        times = {}
        for w in candidates:
            self.set_workers(w)
            self.run(probe)
            times[w] = self.run(probe)[1]
        best = min(times, key=times.get)
        self.set_workers(best)
        return times


class BatchEngine:
    """The batched native form of ``BatchedLFM``: every problem in one
    libdfm batch engine (dfm_batch_*), so each stage of a step is one launch
    over all instances, the projections are one batched MGPCG with
    per-instance FP64 reductions and convergence flags, and the launches
    between two convergence checks replay from captured hipGraphs.

    Problems must share reinit_every and steps (as ``make_problems`` makes
    them).  ``run`` returns the same (loss, dL/dtheta) per instance as
    ``BatchedLFM.run``; the first run captures the graphs, later runs replay
    them."""

    def __init__(self, problems, tol=1e-6, max_iters=200, graphs=True, first_iters=None, chunk=None,
                 tail_cells=None, bottom_exact=True, tile=None, device="cuda"):
        import torch

        from . import _lib as L

        self.problems = list(problems)
        R = {p["R"] for p in self.problems}
        steps = {p["steps"] for p in self.problems}
        if len(R) != 1 or len(steps) != 1:
            raise ValueError("BatchEngine needs one reinit_every and one step count for all problems")
        self.R, self.steps = R.pop(), steps.pop()
        self.device = torch.device(device)
        self.lib = L.lib()
        n = len(self.problems)
        arr = (L.BatchInstance * n)()
        for i, p in enumerate(self.problems):
            arr[i].n[0], arr[i].n[1] = int(p["n"][0]), int(p["n"][1])
            arr[i].dx, arr[i].dt = float(p["dx"]), float(p["dt"])
        h = L.c_void_p()
        L.check(self.lib.dfm_batch_create(L.ctypes.cast(arr, L.c_void_p), n, self.R, self.steps, float(tol),
                                          int(max_iters), L.ctypes.byref(h)), "batch_create")
        self.h = h
        if not graphs:
            self.set_option(L.BATCH_GRAPH, 0)
        if first_iters is not None:
            self.set_option(L.BATCH_FIRST_ITERS, first_iters)
        if chunk is not None:
            self.set_option(L.BATCH_CHUNK, chunk)
        if tail_cells is not None:
            self.set_option(L.BATCH_TAIL_CELLS, tail_cells)
        if not bottom_exact:
            self.set_option(L.BATCH_BOTTOM_EXACT, 0)
        if tile is not None:
            self.set_option(L.BATCH_TILE, tile)
        self.options = {"first_iters": first_iters or 8, "chunk": chunk or 4, "tail_cells": tail_cells or 16384,
                        "bottom_exact": bool(bottom_exact), "tile": tile or 32}
        # per-instance device inputs, uploaded once
        self.theta = [torch.tensor(p["theta"], dtype=torch.float32, device=self.device) for p in self.problems]
        self.density = [torch.tensor(p["density"], dtype=torch.float32, device=self.device) for p in self.problems]
        self.target = [torch.tensor(p["target"], dtype=torch.float32, device=self.device) for p in self.problems]
        self.grids = [L.make_grid(p["n"], p["dx"]) for p in self.problems]
        self.loss = np.zeros(n, dtype=np.float64)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.dfm_batch_destroy(h)
            except Exception:
                pass
            self.h = None

    def set_option(self, option, value):
        from . import _lib as L

        L.check(self.lib.dfm_batch_set_option(self.h, int(option), int(value)), "batch_set_option")

    def _mac(self, field, i):
        from . import _lib as L

        c0, c1 = L.c_void_p(), L.c_void_p()
        L.check(self.lib.dfm_batch_field(self.h, field, i, L.ctypes.byref(c0), L.ctypes.byref(c1)), "batch_field")
        m = L.Mac()
        m.c[0], m.c[1] = c0.value, c1.value
        return m

    def get(self, field, i):
        """Copy of instance i's field: [x faces, y faces] for MAC fields, one
        tensor for cell fields."""
        import torch

        from . import _lib as L

        n = self.problems[i]["n"]
        if field in (L.BATCH_U, L.BATCH_ADJ_U):
            out = [torch.empty((n[0] + 1, n[1]), device=self.device), torch.empty((n[0], n[1] + 1), device=self.device)]
            L.check(self.lib.dfm_batch_copy(self.h, field, i, out[0].data_ptr(), out[1].data_ptr(), 0,
                                            L.stream_handle()), "batch_copy")
        else:
            out = torch.empty(tuple(n), device=self.device)
            L.check(self.lib.dfm_batch_copy(self.h, field, i, out.data_ptr(), None, 0, L.stream_handle()),
                    "batch_copy")
        return out

    def put(self, field, i, value):
        from . import _lib as L

        if isinstance(value, (list, tuple)):
            v = [t.float().contiguous() for t in value]
            L.check(self.lib.dfm_batch_copy(self.h, field, i, v[0].data_ptr(), v[1].data_ptr(), 1,
                                            L.stream_handle()), "batch_copy")
        else:
            v = value.float().contiguous()
            L.check(self.lib.dfm_batch_copy(self.h, field, i, v.data_ptr(), None, 1, L.stream_handle()),
                    "batch_copy")

    def start(self, u0=None):
        """Initial state of every instance: u = the vortex field of theta (or
        u0[i], MAC tensors), s = density, then the projected init."""
        from . import _lib as L

        st = L.stream_handle()
        for i, p in enumerate(self.problems):
            if u0 is not None:
                self.put(L.BATCH_U, i, u0[i])
            else:
                L.check(self.lib.dfm_vortex_velocity_2d(self.grids[i], self.theta[i].data_ptr(),
                                                        self.theta[i].numel() // 4, self._mac(L.BATCH_U, i), st),
                        "vortex_velocity_2d")
            self.put(L.BATCH_S, i, self.density[i])
            self.put(L.BATCH_TARGET, i, self.target[i])
        L.check(self.lib.dfm_batch_init(self.h, st), "batch_init")

    def forward(self, steps=None):
        from . import _lib as L

        L.check(self.lib.dfm_batch_forward(self.h, int(self.steps if steps is None else steps), L.stream_handle()),
                "batch_forward")

    def init_gradient(self):
        from . import _lib as L

        L.check(self.lib.dfm_batch_init_gradient(self.h, self.loss.ctypes.data_as(L.c_void_p), L.stream_handle()),
                "batch_init_gradient")
        return self.loss.copy()

    def backward(self, steps=None):
        from . import _lib as L

        L.check(self.lib.dfm_batch_backward(self.h, int(self.steps if steps is None else steps), L.stream_handle()),
                "batch_backward")

    def dtheta(self, i):
        """dL/dtheta of instance i from its adjoint velocity at t = 0 (the
        analytic vortex VJP, as calculate_dtheta)."""
        import torch

        from . import _lib as L

        out = torch.empty_like(self.theta[i])
        L.check(self.lib.dfm_vortex_velocity_2d_vjp(self.grids[i], self.theta[i].data_ptr(),
                                                    self.theta[i].numel() // 4, self._mac(L.BATCH_ADJ_U, i),
                                                    out.data_ptr(), L.stream_handle()), "vortex_velocity_2d_vjp")
        return out

    def run(self, steps=None):
        """Forward, terminal adjoint, backward and the theta gradients of every
        instance; returns ([(loss, dL/dtheta)], seconds)."""
        import torch

        if steps is not None and steps != self.steps:
            raise ValueError("BatchEngine runs the whole horizon it was created for")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.start()
        self.forward()
        loss = self.init_gradient()
        self.backward()
        grads = [self.dtheta(i) for i in range(len(self.problems))]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return [(float(loss[i]), grads[i].double().cpu().numpy()) for i in range(len(self.problems))], dt

    def instance_steps(self, steps=None):
        return len(self.problems) * self.steps

    def autotune(self, tiles=(16, 32), tails=(1024, 4096, 16384), first_iters=(4, 6, 8)):
        """Pick the tile edge of the LDS-tiled smoothing legs, the LDS-tail
        size (the largest level, in cells, that an instance's one-workgroup
        tail holds; finer levels are batched tiled launches) and the PCG
        iterations run before the first convergence check, one knob after
        the other, each by timing a whole run (after one untimed run that
        captures the graphs); returns {knob: {value: seconds}}."""
        from . import _lib as L

        out = {"tile": {}, "tail_cells": {}, "first_iters": {}}
        for t in tiles:
            self.set_option(L.BATCH_TILE, t)
            self.run()
            out["tile"][t] = self.run()[1]
        best_tile = min(out["tile"], key=out["tile"].get)
        self.set_option(L.BATCH_TILE, best_tile)
        for t in tails:
            self.set_option(L.BATCH_TAIL_CELLS, t)
            self.run()
            out["tail_cells"][t] = self.run()[1]
        best_t = min(out["tail_cells"], key=out["tail_cells"].get)
        self.set_option(L.BATCH_TAIL_CELLS, best_t)
        for f in first_iters:
            self.set_option(L.BATCH_FIRST_ITERS, f)
            self.run()
            out["first_iters"][f] = self.run()[1]
        best_f = min(out["first_iters"], key=out["first_iters"].get)
        self.set_option(L.BATCH_FIRST_ITERS, best_f)
        self.options.update(tile=best_tile, tail_cells=best_t, first_iters=best_f)
        self.run()  # recapture with the chosen schedule
        return out

    def stats(self):
        from . import _lib as L

        a, b, c = L.c_int(), L.c_int(), L.c_int()
        L.check(self.lib.dfm_batch_stats(self.h, L.ctypes.byref(a), L.ctypes.byref(b), L.ctypes.byref(c)),
                "batch_stats")
        return {"projections": a.value, "max_iters": b.value, "noconv": c.value}

    def levels(self, i):
        from . import _lib as L

        a, b = L.c_int(), L.c_int()
        L.check(self.lib.dfm_batch_levels(self.h, i, L.ctypes.byref(a), L.ctypes.byref(b)), "batch_levels")
        return a.value, b.value
