"""Benchmark: 3D forward+adjoint leapfrog flow-map simulation at 128^3.

One "step" = one forward step_midpoint + one adjoint backtrack_step_midpoint
of the 128^3 MAC-grid configuration (RK4, reinitialisation every 10 steps,
16 seeded vortex tubes, box-union density, 4-Gaussian target, MGPCG with 5
levels and 2+2 red-black Gauss-Seidel sweeps, tolerance 1e-6 relative).
The timed region runs K forward steps, the terminal adjoint and K adjoint
steps on device-resident inputs; value = K / elapsed.

Prints ONE JSON line (rank 0).  Extra objects: "roofline" (dominant kernel,
timed with HIP events on the launch stream), "mgpcg" (solver HBM GB/s in
the run, beside the side-stream marches, and "alone": the same solver
projecting one field with nothing beside it),
"cpu_baseline" (the CPU oracle on a bounded sample), "grad_check" (GPU vs
FP64 oracle on a small grid).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "differentiable-flowmap_amd"))

METRIC = "fwd+adjoint sim steps/s at 128³ MAC grid; MGPCG HBM GB/s; grad rel-err vs FP64"
HBM_PEAK_GBPS = 8000.0
FP32_PEAK_TFLOPS = 157.3

# Algorithmic work per unit (flop / byte models).
# 3D quadratic B-spline interpolation of one scalar in the reference's tap
# form (a 4x4x4 loop, separable weights): value + gradient ~480 flop, value
# only ~220 flop.  This is the per-unit figure `roofline.achieved` uses.
FLOP_INTERP_GRAD_3D = 480.0
FLOP_INTERP_VAL_3D = 220.0
# The same samples over their 27 support taps, as the kernels execute them
# (the 37 other taps have weight exactly 0): value + gradient = 93
# multiply-adds (186 flop) + 3 axes of weights and derivative weights (33) +
# the clamped coordinates (21) + the gradient scale (3) = 243 flop; value
# only = 39 multiply-adds (78) + weights (24) + coordinates (21) = 123 flop.
EXEC_FLOP_INTERP_GRAD_3D = 243.0
EXEC_FLOP_INTERP_VAL_3D = 123.0
# Issue ceiling of the vector ALUs without packed FP32 (which the build
# turns off): one wave64 instruction per CU per clock, 256 CUs at 2.4 GHz.
VALU_ISSUE_PER_S = 256 * 2.4e9
# MGPCG compulsory FP32 traffic: bytes per fine cell per PCG iteration (3D) and
# per solve (divergence, init, gradient subtraction).
MG_BYTES_CELL_ITER_3D = 156.0
MG_BYTES_CELL_SOLVE_3D = 64.0


REINIT_EVERY = 10


def plan_steps(steps, warmup, reinit_every=REINIT_EVERY):
    """Step counts the run actually uses.  The adjoint sweep walks back
    through whole reinitialisation intervals (each interval's forward map is
    stored at its reinitialisation), so both counts are rounded up to whole
    intervals: a warmup of 5 runs one interval of 10.  Returns
    {"steps", "warmup", "steps_requested", "warmup_requested"}."""
    if steps < 1:
        raise ValueError(f"--steps must be at least 1 (got {steps})")
    if warmup < 0:
        raise ValueError(f"--warmup must be at least 0 (got {warmup})")
    up = lambda k: -(-k // reinit_every) * reinit_every  # noqa: E731
    return {"steps": up(steps), "warmup": up(warmup), "steps_requested": steps, "warmup_requested": warmup}


def march_flop(n, order=4, grad=FLOP_INTERP_GRAD_3D, val=FLOP_INTERP_VAL_3D):
    """Flop of one dfm_march launch: face points carry (psi, T), cells psi only."""
    import numpy as np

    cells = float(np.prod(n))
    faces = sum(float(np.prod([n[b] + (b == a) for b in range(3)])) for a in range(3))
    stages = 4 if order == 4 else 2
    per_face = stages * (3 * grad + 2 * 9 + 6 * 2) + 12
    per_cell = stages * (3 * val + 6) + 6
    return faces * per_face + cells * per_cell


def sl_flop(n, nsrc_vec, nsrc_scal, order=4):
    import numpy as np

    cells = float(np.prod(n))
    faces = sum(float(np.prod([n[b] + (b == a) for b in range(3)])) for a in range(3))
    stages = 4 if order == 4 else 2
    back = stages * 3 * FLOP_INTERP_VAL_3D
    f = faces * (back + nsrc_vec * FLOP_INTERP_VAL_3D) if nsrc_vec else 0.0
    c = cells * (back + nsrc_scal * FLOP_INTERP_VAL_3D) if nsrc_scal else 0.0
    return f + c


def measured_valu(kernel, grid):
    """Vector-ALU wave instructions per launch of a kernel at a grid size from
    the committed rocprofv3 SQ_INSTS_VALU summary of the 128^3 bench workload,
    with the kernel's duration when it ran alone in that pass; None when the
    summary does not hold them."""
    path = os.path.join(ROOT, "profiles", "pmc_valu_3d128.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(kernel, {}).get(str(int(grid)))


def measured_traffic(dom, n):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), when one
    exists for this kernel and grid; None otherwise."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_3d128.json")
    if dom != "march" or tuple(n) != (128, 128, 128) or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("traffic_bytes")


def quiet_gc():
    """Collect once, then move every object that exists into the permanent
    generation (gc.freeze), so that the cyclic collector's full passes stop
    walking the heap torch and numpy build at import.  A full collection
    inside the timed loop took tens of ms with the GPU idle (the host issues
    every launch): with the same kernels, runs came out at 155 or 116-120
    steps/s depending only on whether one fell into the timed region."""
    import gc

    gc.collect()
    gc.freeze()


def run_gpu(args):
    import numpy as np
    import torch

    from dfm import LFM_Diff_Simulator, scenes, timing

    n = (args.size,) * 3
    sc = scenes.scene_3d(n, seed=0, tubes=16)
    dx = sc["dx"]
    dt = 0.5 * dx
    R = REINIT_EVERY
    K, W = args.steps, args.warmup
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=max(K, W), mg_levels=args.mg_levels, mg_pre_post=args.mg_pre_post,
                             mg_bottom=50, tol=1e-6, max_iters=200, mg_tail=not args.mg_no_tail,
                             mg_bf16_coarse=args.mg_bf16_coarse, mg_bottom_exact=args.mg_bottom_exact,
                             mg_check_every=args.mg_check_every, overlap_march=False if args.overlap_march == "none" else args.overlap_march,
                             mg_fused_sweeps=args.mg_fused_sweeps, mg_split_layout=not args.mg_natural_layout,
                             side_march_blocks=args.side_march_blocks)
    u0, s0, tgt = sc["u0"], sc["density"], sc["target"]
    # the storage of a K-step run, allocated before the warm-up: a warm-up
    # shorter than K must not leave allocations to the timed region
    sim.reserve(max(K, W))
    quiet_gc()

    def forward_adjoint(steps):
        sim.total_steps = steps
        sim.init(u0=u0, init_passive_np=s0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step_midpoint()
        sim.init_gradient(target=tgt)
        for _ in range(steps):
            sim.backtrack_step_midpoint()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    if W > 0:
        forward_adjoint(W)
    timing.reset()
    timing.enable(True)
    elapsed = forward_adjoint(K)
    timing.enable(False)
    summ = timing.summary()
    cnt = timing.counts()
    lam = torch.cat([c.flatten() for c in sim.adj_u.c])
    assert bool(torch.isfinite(lam).all()), "non-finite adjoint"

    # dominant flow-map kernel by total device time
    faces = sum(float(np.prod([n[b] + (b == a) for b in range(3)])) for a in range(3))
    cells = float(np.prod(n))
    models = {
        "march": (march_flop(n), "dfm_march (k_march<3,4>)", "k_march<3, 4>", faces + cells),
        "sl_advect": (sl_flop(n, 2, 2), "dfm_sl_advect (k_sl_advect<3,4>)", None, None),
        "adj_conv": (faces * 3 * 216 * 4, "dfm_adj_conv_source (k_src_tiled<3>)", None, None),
    }
    dom = max((k for k in models if k in summ), key=lambda k: summ[k][1])
    calls, tot_ms = summ[dom]
    avg_s = tot_ms / calls / 1e3
    flop = models[dom][0]
    achieved = flop / avg_s / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": measured_traffic(dom, n),
                "kernel": models[dom][1], "avg_launch_ms": round(avg_s * 1e3, 4), "calls": calls,
                "flop_per_launch": flop,
                "flop_model": "the reference's tap form: 480 flop per quadratic sample with the gradient, "
                              "220 value only",
                "note": "B-spline samples with no dense contraction, so no MFMA: priced against the published FP32 "
                        "vector peak, which counts packed FP32; the build issues unpacked FP32, whose ceiling is "
                        "half of it.  executed_*: the 27 support taps the kernel evaluates (243 / 123 flop per "
                        "sample).  valu_issue_frac: the kernel's vector-ALU wave instructions per launch "
                        "(rocprofv3 SQ_INSTS_VALU over this workload) over the issue ceiling of one wave64 "
                        "instruction per CU per clock, in this run's launch time (beside the projections) and "
                        "in the counter pass's own (the kernel alone).  The side-stream marches run with at most "
                        "side_march_blocks workgroups per CU (dfm_march_capped), which lengthens their in-run "
                        "launches; alone_ms / frac_alone: the uncapped launch with nothing beside it"}
    if dom == "march":
        ex = march_flop(n, grad=EXEC_FLOP_INTERP_GRAD_3D, val=EXEC_FLOP_INTERP_VAL_3D)
        roofline["executed_flop_per_launch"] = ex
        roofline["executed_frac"] = round(ex / avg_s / 1e12 / FP32_PEAK_TFLOPS, 4)
        # the same launch with nothing beside it and no workgroup cap, after
        # the timed region (on the run's own maps and velocity, which are no
        # longer needed): the kernel's own rate, against the in-run rate of
        # the side-stream marches that hold at most side_march_blocks
        # workgroups per CU so that the projections beside them find free CUs
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        sim.ops.march(sim.u, sim.phi_map, sim.dt, 4)
        ev0.record()
        for _ in range(5):
            sim.ops.march(sim.u, sim.phi_map, sim.dt, 4)
        ev1.record()
        torch.cuda.synchronize()
        alone_s = ev0.elapsed_time(ev1) / 5 / 1e3
        roofline["side_march_blocks"] = args.side_march_blocks
        roofline["alone_ms"] = round(alone_s * 1e3, 4)
        roofline["frac_alone"] = round(flop / alone_s / 1e12 / FP32_PEAK_TFLOPS, 4)
    if models[dom][2] is not None and tuple(n) == (128, 128, 128):
        grid = -(-int(models[dom][3]) // 256) * 256
        v = measured_valu(models[dom][2], grid)
        if v:
            roofline["valu_per_launch"] = v["valu"]
            roofline["valu_issue_frac"] = round(v["valu"] / (VALU_ISSUE_PER_S * avg_s), 3)
            roofline["valu_issue_frac_alone"] = round(v["valu"] / (VALU_ISSUE_PER_S * v["us"] * 1e-6), 3)
    mg_calls, mg_ms = summ.get("mgpcg", (0, 0.0))
    mg_iters = cnt.get("mgpcg_iters", 0)
    cells = float(np.prod(n))
    mg_bytes = cells * (MG_BYTES_CELL_ITER_3D * mg_iters + MG_BYTES_CELL_SOLVE_3D * mg_calls)
    mg_gbps = mg_bytes / (mg_ms / 1e3) / 1e9 if mg_ms > 0 else 0.0
    mgpcg = {"achieved_GBps": round(mg_gbps, 1), "peak_GBps": HBM_PEAK_GBPS, "frac": round(mg_gbps / HBM_PEAK_GBPS, 4),
             "solves": mg_calls, "pcg_iters_total": mg_iters,
             "iters_per_solve": round(mg_iters / max(mg_calls, 1), 2), "ms_per_solve": round(mg_ms / max(mg_calls, 1), 3),
             "not_converged": sim.noconv_count,
             "bytes_model": "156 B/cell/PCG-iter + 64 B/cell/solve (FP32 compulsory)"}
    # The same solver projecting the scene's initial velocity with nothing
    # beside it: inside the run the projections share the GPU with the
    # marches on the side stream, so their wall time there is not the
    # solver's own.
    from dfm.ops import MacField

    src = MacField([torch.as_tensor(np.ascontiguousarray(c), dtype=torch.float32, device="cuda") for c in u0])
    mgpcg["alone"] = solve_alone(sim.solver, src, n)
    mgpcg["alone"]["field"] = "the scene's initial velocity (unprojected), projected alone"
    if not args.no_mg256 and args.size != 256:
        mgpcg["alone_256"] = mgpcg_256_alone(args)
    breakdown = {k: {"calls": v[0], "ms": round(v[1], 2)} for k, v in summ.items()}
    return elapsed, roofline, mgpcg, breakdown


def solve_alone(solver, src, n, reps=10):
    """Event-timed projections of `src` by `solver` on the current stream
    with nothing else queued: 2 untimed, then `reps` timed."""
    import numpy as np
    import torch

    work = src.clone()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms, its, noconv0 = 0.0, 0, solver.noconv_count
    for rep in range(reps + 2):
        work.copy_(src)
        ev0.record()
        solver.Poisson(work)
        ev1.record()
        torch.cuda.synchronize()
        if rep >= 2:
            ms += ev0.elapsed_time(ev1)
            its += solver.last_iters
    cells = float(np.prod(n))
    gbps = cells * (MG_BYTES_CELL_ITER_3D * its + MG_BYTES_CELL_SOLVE_3D * reps) / (ms / 1e3) / 1e9
    return {"achieved_GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4), "ms_per_solve": round(ms / reps, 3),
            "iters_per_solve": round(its / reps, 2), "solves": reps, "not_converged": solver.noconv_count - noconv0}


def mgpcg_256_alone(args, size=256):
    """The solver at size^3 (at 256^3, 6 levels and the bandwidth-bound
    case: one FP32 field is 67 MB, past what the caches hold) projecting a
    seeded random MAC field alone, with the run's solver options."""
    import torch

    from dfm.ops import MGPCG, MacField

    n = (size, size, size)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1)
    src = MacField.empty(n, "cuda")
    for c in src.c:
        c.copy_(torch.rand(c.shape, generator=gen, device="cuda") - 0.5)
    solver = MGPCG(n, 1.0 / 256, levels=0, pre_post=args.mg_pre_post, bottom=50, tol=1e-6, max_iters=200,
                   bf16_coarse=args.mg_bf16_coarse, tail=not args.mg_no_tail, bottom_exact=args.mg_bottom_exact,
                   check_every=args.mg_check_every, fused_sweeps=args.mg_fused_sweeps,
                   split_layout=not args.mg_natural_layout)
    out = solve_alone(solver, src, n)
    out["field"] = f"seeded uniform random MAC field in [-0.5, 0.5), {size}^3, projected alone"
    del solver, src
    torch.cuda.empty_cache()
    return out


def run_batched_vs_plain(args):
    """The mixed-precision batched form against the plain FP32 form at
    args.size^3, forward+adjoint over one reinitialisation interval per
    problem (seeded vortex tubes).  Plain: one problem after another, FP32
    multigrid levels, eager launches, no side streams.  Mixed batched: k
    problems on k worker threads and streams, each simulator with BF16
    storage of the multigrid levels >= 2, captured PCG-iteration graphs and
    its marches and independent projections on side streams; k from
    --batched-instances (1..k all timed, the best is the value).  Both solve to
    the same tolerance, so the BF16 levels only change the preconditioner;
    the gradients of problem 0 under the two forms are compared (rel-L2)."""
    import numpy as np
    import torch

    from concurrent.futures import ThreadPoolExecutor

    from dfm import LFM_Diff_Simulator, scenes

    n = (args.size,) * 3
    R = REINIT_EVERY
    kmax = max(1, args.batched_instances)
    scs = [scenes.scene_3d(n, seed=s, tubes=16) for s in range(kmax)]
    dx = scs[0]["dx"]

    def make(mixed):
        kw = dict(total_steps=R, mg_levels=args.mg_levels, tol=1e-6, max_iters=200)
        if mixed:
            kw.update(mg_bf16_coarse=True, overlap_march="both", overlap_projections=True)
        else:
            kw.update(mg_bf16_coarse=False, overlap_march=False, overlap_projections=False, mg_graph=False)
        sim = LFM_Diff_Simulator(*n, dx, 0.5 * dx, R, None, **kw)
        sim.reserve(R)
        return sim

    def fwd_adj(sim, sc):
        sim.init(u0=sc["u0"], init_passive_np=sc["density"])
        for _ in range(R):
            sim.step_midpoint()
        sim.init_gradient(target=sc["target"])
        for _ in range(R):
            sim.backtrack_step_midpoint()


    plain = make(False)
    quiet_gc()
    for _ in range(2):  # warm-up: allocator, handles
        fwd_adj(plain, scs[0])
    plain_runs = []
    for _ in range(args.batched_reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd_adj(plain, scs[0])
        torch.cuda.synchronize()
        plain_runs.append(R / (time.perf_counter() - t0))
    plain_sps = max(plain_runs)
    g_plain = torch.cat([c.flatten() for c in plain.adj_u.c]).double().cpu().numpy()
    del plain
    sims = [make(True) for _ in scs]
    streams = [torch.cuda.Stream() for _ in scs]
    quiet_gc()

    def work(k):
        with torch.cuda.stream(streams[k]):
            fwd_adj(sims[k], scs[k])
            streams[k].synchronize()

    mixed, mixed_runs = {}, {}
    for k in range(1, kmax + 1):
        with ThreadPoolExecutor(max_workers=k) as ex:
            for _ in range(2):  # warm-up: graphs, handles
                list(ex.map(work, range(k)))
            runs = []
            for _ in range(args.batched_reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                list(ex.map(work, range(k)))
                torch.cuda.synchronize()
                runs.append(k * R / (time.perf_counter() - t0))
        mixed[k] = max(runs)
        mixed_runs[k] = runs
    best = max(mixed, key=mixed.get)
    g_mixed = torch.cat([c.flatten() for c in sims[0].adj_u.c]).double().cpu().numpy()
    return {
        "metric": f"mixed-precision batched fwd+adjoint steps/s / plain FP32 steps/s at {args.size}^3",
        "value": round(mixed[best] / plain_sps, 3), "unit": "ratio", "n_gpus": 1,
        "mixed_batched_steps_per_s": {str(k): round(v, 3) for k, v in mixed.items()}, "best_instances": best,
        "plain_fp32_steps_per_s": round(plain_sps, 3), "steps_per_instance": R,
        "runs": {"plain": [round(v, 2) for v in plain_runs],
                 "mixed": {str(k): [round(v, 2) for v in r] for k, r in mixed_runs.items()},
                 "statistic": f"best of {args.batched_reps} timed runs after 2 warm-up runs (the plain form's eager launches make its runs vary with host jitter; the best run of each form is the stable figure)"},
        "grad_rel_l2_mixed_vs_plain": float(np.linalg.norm(g_mixed - g_plain) / np.linalg.norm(g_plain)),
        "noconv": int(sum(s.noconv_count for s in sims)),
        "higher_is_better": True, "dtype": "f32 (BF16 storage of MG levels >= 2 in the mixed form)",
        "data": "synthetic (seeded vortex tubes / box-union density / Gaussian targets)",
        "config": {"workload": f"3D forward+adjoint leapfrog flow-map, {args.size}^3, RK4, one reinitialisation "
                               f"interval ({R} steps) per instance, MGPCG tol 1e-6"},
    }


def run_grad_check():
    """GPU forward+adjoint vs the FP64 oracle on a 32^3 grid (2 intervals)."""
    import numpy as np
    import torch

    from dfm import LFM_Diff_Simulator, scenes
    from oracle import lfm_oracle as O

    n, R, steps = (32, 32, 32), 5, 10
    sc = scenes.scene_3d(n, seed=7, tubes=8)
    dx, dt = sc["dx"], 0.5 * sc["dx"]
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    u0 = [f32(c) for c in sc["u0"]]
    s0, tgt = f32(sc["density"]), f32(sc["target"])
    sim = LFM_Diff_Simulator(*n, dx, dt, R, None, total_steps=steps, tol=1e-7, max_iters=200)
    sim.init(u0=u0, init_passive_np=s0)
    orc = O.OracleLFM(n, dx, dt, R, 4, adj_kernel=3)
    orc.init(u0, [s0])
    for _ in range(steps):
        sim.step_midpoint()
        orc.step()

    def rel(dev_list, ref_list):
        num = sum(float(np.sum((d.cpu().numpy() - r) ** 2)) for d, r in zip(dev_list, ref_list))
        return float(np.sqrt(num / sum(float(np.sum(r ** 2)) for r in ref_list)))

    e_u = rel(sim.u.c, orc.u)
    e_s = rel(sim.passive, orc.s)
    sim.init_gradient(target=tgt)
    orc.init_gradient([tgt])
    for _ in range(steps):
        sim.backtrack_step_midpoint()
        orc.backtrack_step()
    torch.cuda.synchronize()
    return {"grid": list(n), "steps": steps, "reinit_every": R, "rel_l2_u_final": e_u, "rel_l2_density_final": e_s,
            "rel_l2_grad_u0": rel(sim.adj_u.c, orc.lam_u), "rel_l2_grad_s0": rel(sim.adj_passive, orc.lam_s),
            "oracle": "FP64 CPU restatement, exact (DCT) projection"}


def run_cpu_baseline(size):
    """The oracle in FP32 mode with the restated MGPCG, OpenMP on the host."""
    import numpy as np

    from dfm import scenes
    from oracle import lfm_oracle as O

    threads = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16))
    n = (size,) * 3
    sc = scenes.scene_3d(n, seed=0, tubes=16)
    dx = sc["dx"]
    orc = O.OracleLFM(n, dx, 0.5 * dx, 10, 4, adj_kernel=3, project="mgpcg", dtype=np.float32, mg_levels=5,
                      mg_tol=1e-6)
    orc.init(sc["u0"], [sc["density"]])
    orc.step()  # untimed: step 0
    t0 = time.perf_counter()
    orc.step()  # step 1 (no reinitialisation)
    t1 = time.perf_counter()
    orc.init_gradient([sc["target"].astype(np.float32)])
    t2 = time.perf_counter()
    orc.backtrack_step()  # adjoint of step 1 (no reinitialisation)
    t3 = time.perf_counter()
    el = (t1 - t0) + (t3 - t2)
    return {"value": round(1.0 / el, 5), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"1 forward + 1 adjoint step at {size}^3 (a step without reinitialisation; the "
                      f"1-in-10 reinitialisation cost is not in the sample), FP32, C++/OpenMP oracle with the "
                      f"restated MGPCG; {el:.1f} s"}


def run_cpu_interval(size, steps):
    """The CPU baseline over whole steps: the oracle (FP32, restated MGPCG,
    OpenMP over OMP_NUM_THREADS threads) runs `steps` forward steps from t =
    0 and then the adjoint sweep back (with `steps` = 10, one full
    reinitialisation interval: the forward reinitialisation at step 9 and
    the adjoint one at step 0 included); every step is timed."""
    import numpy as np

    from dfm import scenes
    from oracle import lfm_oracle as O

    threads = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))
    n = (size,) * 3
    sc = scenes.scene_3d(n, seed=0, tubes=16)
    dx = sc["dx"]
    R = REINIT_EVERY
    if (steps % R and steps >= R) or steps < 2:
        raise SystemExit("bench.py: --cpu-interval steps must be 2..interval-1 or a multiple of the interval")
    orc = O.OracleLFM(n, dx, 0.5 * dx, R, 4, adj_kernel=3, project="mgpcg", dtype=np.float32, mg_levels=5,
                      mg_tol=1e-6)
    t_start = time.perf_counter()
    orc.on_stage = lambda name: print(f"cpu-interval: {name} done at {time.perf_counter() - t_start:.1f} s",
                                      file=sys.stderr, flush=True)
    orc.init(sc["u0"], [sc["density"]])
    fwd, adj = [], []
    for k in range(steps):
        t0 = time.perf_counter()
        orc.step()
        fwd.append(time.perf_counter() - t0)
        print(f"cpu-interval: forward step {k} {fwd[-1]:.1f} s", file=sys.stderr, flush=True)
    if steps % R == 0:
        t0 = time.perf_counter()
        orc.init_gradient([sc["target"].astype(np.float32)])
        t_init = time.perf_counter() - t0
        for k in range(steps):
            t0 = time.perf_counter()
            orc.backtrack_step()
            adj.append(time.perf_counter() - t0)
            print(f"cpu-interval: adjoint step {steps - 1 - k} {adj[-1]:.1f} s", file=sys.stderr, flush=True)
    else:
        # fewer steps than an interval: the adjoint of the last forward step
        # only (not step 0, whose adjoint reinitialises from the interval's
        # stored forward map)
        t0 = time.perf_counter()
        orc.init_gradient([sc["target"].astype(np.float32)])
        t_init = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.backtrack_step()
        adj.append(time.perf_counter() - t0)
        print(f"cpu-interval: adjoint step {steps - 1} {adj[-1]:.1f} s", file=sys.stderr, flush=True)
    total = sum(fwd) + sum(adj) + t_init
    # whole intervals: step pairs over the whole time; otherwise one pair of
    # mean forward and adjoint steps (no reinitialisation in either)
    rate = len(fwd) / total if len(adj) == len(fwd) else 1.0 / (sum(fwd) / len(fwd) + sum(adj) / len(adj))
    return {"cpu_interval": {"value": round(rate, 6), "unit": "fwd+adjoint steps/s", "threads": threads,
                             "kind": "port", "size": size, "steps": steps,
                             "forward_s": [round(x, 2) for x in fwd], "adjoint_s": [round(x, 2) for x in adj],
                             "init_gradient_s": round(t_init, 2), "total_s": round(total, 1)}}


def run_batch(args):
    """Config 5: 32 independent 2D forward+adjoint problems, grids from
    {128,192,256,384,512}^2 by seed, 4-12 vortices each, theta gradient per
    instance.  The batched native engine (one launch per stage over all
    instances, batched MGPCG, graph-replayed segments; LDS-tail size and
    first-check iterations autotuned), timed beside the per-instance
    simulators run one after another and on concurrent worker threads; the
    per-instance gradients are compared with the one-after-another run."""
    import numpy as np

    from dfm.batch import BatchedLFM, BatchEngine, make_problems

    probs = make_problems(count=32, seed=100, steps=args.batch_steps, reinit_every=12)
    eng = BatchEngine(probs, tol=1e-6, max_iters=200)
    quiet_gc()
    tune = eng.autotune()
    res, t = eng.run()
    st = eng.stats()
    steps_total = eng.instance_steps()
    # per-instance simulators: one after another (the reference gradients),
    # then on worker threads
    b = BatchedLFM(probs, workers=1)
    b.run(12)  # untimed: builds every handle and graph
    seq, t_seq = b.run()
    b.set_workers(8)
    b.run(12)
    _, t_thr = b.run()
    per = []
    for p, (l0, g0), (l1, g1) in zip(probs, seq, res):
        per.append({"n": p["n"][0], "vortices": int(p["theta"].size // 4), "loss": round(float(l1), 6),
                    "grad_norm": round(float(np.linalg.norm(g1)), 6),
                    "grad_rel_err_vs_single": float(np.linalg.norm(g1 - g0) / max(np.linalg.norm(g0), 1e-30))})
    return {
        "metric": "aggregate fwd+adjoint instance-steps/s, 32 uneven 2D instances (128^2..512^2)",
        "value": round(steps_total / t, 2), "unit": "steps/s", "n_gpus": 1, "steps": args.batch_steps,
        "ms_total": round(t * 1e3, 1),
        "sequential_value": round(steps_total / t_seq, 2), "thread_pool_value": round(steps_total / t_thr, 2),
        "speedup_vs_sequential": round(t_seq / t, 3), "speedup_vs_thread_pool": round(t_thr / t, 3),
        "engine": {"options": eng.options, "projections": st["projections"], "max_pcg_iters": st["max_iters"],
                   "noconv": st["noconv"],
                   "autotune_s": {k: {str(a): round(v, 4) for a, v in d.items()} for k, d in tune.items()}},
        "max_grad_rel_err_vs_single": max(q["grad_rel_err_vs_single"] for q in per),
        "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded vortices / disk / Gaussian blobs)",
        "config": {"workload": "32 independent 2D forward+adjoint LFM problems, RK2, reinit every 12, "
                               "MGPCG tol 1e-6, theta gradient", "instances": len(probs)},
        "per_instance": per,
    }


def run_copy_calibration():
    """Plain device-to-device copy of a known size (256 MiB, above the 256 MB
    MALL), the reference point that scales the FETCH_SIZE / WRITE_SIZE
    counters of a profiled kernel to HBM bytes."""
    import torch

    nbytes = 256 << 20
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 5}), flush=True)


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad-check", action="store_true")
    ap.add_argument("--no-mg256", action="store_true", help="skip the 256^3 MGPCG-alone measurement")
    ap.add_argument("--mg-no-tail", action="store_true", help="one launch per sweep on every MG level")
    ap.add_argument("--mg-bf16-coarse", action="store_true", help="BF16 storage of MG levels >= 2")
    ap.add_argument("--mg-rbgs-bottom", dest="mg_bottom_exact", action="store_false",
                    help="the reference's 50 RBGS sweeps on the coarsest level instead of the exact DCT solve")
    ap.add_argument("--mg-check-every", type=int, default=1, help="PCG iterations per convergence poll")
    ap.add_argument("--mg-pre-post", type=int, default=2, help="RBGS sweeps before and after each coarse correction")
    ap.add_argument("--mg-fused-sweeps", type=int, default=2, choices=[0, 1, 2],
                    help="coarse-level smoothing: 2 LDS-tiled, 1 one launch per sweep, 0 two per sweep")
    ap.add_argument("--mg-natural-layout", action="store_true",
                    help="finest MG level in natural cell order (default: colour-split red/black halves)")
    ap.add_argument("--overlap-march", default="both", choices=["none", "forward", "adjoint", "both"],
                    help="which steps run their flow-map marches on a side stream beside the projections")
    ap.add_argument("--mg-levels", type=int, default=0,
                    help="MG levels (0: log2(n) - 2, the reference's count: 5 at 128^3, 6 at 256^3)")
    ap.add_argument("--batch", action="store_true",
                    help="the batched configuration: 32 uneven 2D instances, aggregate steps/s")
    ap.add_argument("--batch-steps", type=int, default=24)
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for --batch (1..32; 0 keeps the platform's setting: 16 measured slower than 4)")
    ap.add_argument("--batched-vs-plain", action="store_true",
                    help="mixed-precision batched fwd+adjoint vs the plain FP32 form at --size^3 (steps/s ratio)")
    ap.add_argument("--batched-instances", type=int, default=4, help="instances of the --batched-vs-plain batched form")
    ap.add_argument("--batched-reps", type=int, default=3, help="timed runs per form in --batched-vs-plain (best reported)")
    ap.add_argument("--cpu-interval", type=int, default=0, metavar="STEPS",
                    help="only the CPU baseline: the oracle over STEPS forward steps and (for a whole number of "
                         "intervals) the adjoint sweep back, each step timed (OMP_NUM_THREADS threads)")
    ap.add_argument("--side-march-blocks", type=int, default=2,
                    help="workgroups per CU a side-stream march may hold (0: no cap)")
    ap.add_argument("--mg-alone", action="store_true",
                    help="only the MGPCG solver projecting a seeded random field at --size^3 (10 timed solves)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 256 MiB copy used to calibrate the HBM traffic counters")
    return ap.parse_args(argv)


def main(argv=None):
    args = parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: the simulator runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is not implemented.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        run_copy_calibration()
        return
    if args.cpu_interval:
        print(json.dumps(run_cpu_interval(args.size, args.cpu_interval)), flush=True)
        return
    if args.mg_alone:
        out = mgpcg_256_alone(args, args.size)
        out.update({"metric": f"MGPCG projection alone at {args.size}^3", "bytes_model": "156 B/cell/PCG-iter + 64 B/cell/solve"})
        print(json.dumps(out), flush=True)
        return
    if args.batched_vs_plain:
        print(json.dumps(run_batched_vs_plain(args)), flush=True)
        return
    if args.batch:
        from dfm.batch import configure_hw_queues

        if args.hw_queues > 0:
            configure_hw_queues(args.hw_queues)
        print(json.dumps(run_batch(args)), flush=True)
        return
    try:
        plan = plan_steps(args.steps, args.warmup)
    except ValueError as e:
        sys.exit(f"bench.py: {e}")
    args.steps, args.warmup = plan["steps"], plan["warmup"]
    elapsed, roofline, mgpcg, breakdown = run_gpu(args)
    K = args.steps
    out = {
        "metric": METRIC if args.size == 128 else METRIC.replace("128³", f"{args.size}³"),
        "value": round(K / elapsed, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "steps_requested": plan["steps_requested"],
        "warmup_requested": plan["warmup_requested"],
        "ms_per_step": round(elapsed / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded vortex tubes / box-union density / Gaussian target)",
        "config": {"workload": "3D forward+adjoint leapfrog flow-map, %d^3 MAC grid, RK4, reinit every 10, "
                               "MGPCG %s levels 2+2 RBGS tol 1e-6%s" % (
                                   args.size, args.mg_levels or "auto", ", BF16 coarse levels" if args.mg_bf16_coarse else ""),
                   "grid": [args.size] * 3, "reinit_every": 10, "steps_timed": K},
        "roofline": roofline,
        "mgpcg": mgpcg,
        "breakdown_ms": breakdown,
    }
    if not args.no_grad_check:
        out["grad_check"] = run_grad_check()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_baseline(args.size)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
