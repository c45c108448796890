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
