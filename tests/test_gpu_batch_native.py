"""The native batched engine (dfm_batch_*): uneven 2D instances stepped
together, one launch per stage and a batched MGPCG.

* Every instance's forward fields, loss, adjoint fields and theta gradient
  match the FP64 oracle run on that instance alone (rel-L2 <= 1e-5; the
  engine's batched PCG and the oracle's exact projection agree to the solve
  tolerance).
* Replayed graphs and eager launches give bitwise the same results, and a
  second run replays the first run's graphs to the same bits (the
  per-instance reductions are deterministic).
* At the batched configuration's sizes (128^2 .. 512^2, 32 instances) every
  solve converges, and the gradients match the single-instance GPU
  simulator's (rel-L2 <= 1e-3, the gradient bar)."""
import numpy as np
import pytest
import torch

from gpu_util import ADJ_TOL, FWD_TOL, rel_l2, rel_l2_fields, round32, within

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def small_problems(sizes, steps=8, R=4, seed=40):
    from dfm import batch

    probs = []
    for k, n in enumerate(sizes):
        probs += batch.make_problems(count=1, seed=seed + k, steps=steps, reinit_every=R, sizes=(n,))
    return probs


@pytest.mark.parametrize("bottom_exact", [True, False])
def test_instances_match_the_oracle(bottom_exact, device):
    """bottom_exact: the coarsest level solved by DCT (default) or by the
    reference's 50 red-black sweeps; either preconditioner gives the
    projection to the solve tolerance."""
    from dfm import _lib as L
    from dfm import scenes
    from dfm.batch import BatchEngine

    probs = small_problems([32, 48, 64, 32])
    eng = BatchEngine(probs, tol=1e-7, max_iters=200, bottom_exact=bottom_exact)
    eng.start()
    eng.forward()
    loss = eng.init_gradient()
    fwd = [([t.cpu().numpy() for t in eng.get(L.BATCH_U, i)], eng.get(L.BATCH_S, i).cpu().numpy())
           for i in range(len(probs))]
    eng.backward()
    st = eng.stats()
    assert st["noconv"] == 0, st
    for i, p in enumerate(probs):
        n, dx = p["n"], p["dx"]
        u0 = round32(scenes.split_central_vector(scenes.vortex_velocity_2d(p["theta"].astype(np.float64), n, dx)))
        orc = O.OracleLFM(n, dx, p["dt"], p["R"], 2)
        orc.init(u0, [round32(p["density"])])
        for _ in range(p["steps"]):
            orc.step()
        within(f"inst {i} u", rel_l2_fields(fwd[i][0], orc.u), FWD_TOL)
        within(f"inst {i} s", rel_l2(fwd[i][1], orc.s[0]), FWD_TOL)
        orc.init_gradient([round32(p["target"])])
        assert abs(loss[i] - orc.loss) <= 1e-5 * orc.loss, (i, loss[i], orc.loss)
        for _ in range(p["steps"]):
            orc.backtrack_step()
        within(f"inst {i} lam_u", rel_l2_fields([t.cpu().numpy() for t in eng.get(L.BATCH_ADJ_U, i)], orc.lam_u),
               ADJ_TOL)
        within(f"inst {i} lam_s", rel_l2(eng.get(L.BATCH_ADJ_S, i).cpu().numpy(), orc.lam_s[0]), ADJ_TOL)


def test_graphs_match_eager_and_replays_are_bitwise(device):
    from dfm.batch import BatchEngine

    probs = small_problems([64, 32, 48], steps=8, R=4, seed=70)
    runs = []
    for graphs in (True, False):
        eng = BatchEngine(probs, tol=1e-6, graphs=graphs)
        res, _ = eng.run()
        runs.append(res)
        if graphs:
            again, _ = eng.run()  # replays the captured graphs
            for (la, ga), (lb, gb) in zip(res, again):
                assert la == lb and np.array_equal(ga, gb)
    for (la, ga), (lb, gb) in zip(*runs):
        assert la == lb and np.array_equal(ga, gb)


def test_batched_configuration_converges_and_matches_single_instance(device):
    """The full batched workload: 32 instances drawn from 128^2 .. 512^2; every solve
    converges and the theta gradients of a sample of instances match the
    single-instance GPU simulator's."""
    from dfm.batch import BatchedLFM, BatchEngine, make_problems

    probs = make_problems(count=32, seed=100, steps=12, reinit_every=12)
    eng = BatchEngine(probs, tol=1e-6, max_iters=200)
    res, _ = eng.run()
    st = eng.stats()
    assert st["noconv"] == 0, st
    assert all(np.isfinite(l) and np.all(np.isfinite(g)) for l, g in res)
    sample = sorted({int(np.argmin([abs(p["n"][0] - s) for p in probs])) for s in (128, 192, 256, 384, 512)})
    ref = BatchedLFM([probs[i] for i in sample], workers=1)
    rres, _ = ref.run()
    for k, i in enumerate(sample):
        assert abs(res[i][0] - rres[k][0]) <= 1e-4 * abs(rres[k][0]), (i, res[i][0], rres[k][0])
        within(f"dtheta inst {i} (n={probs[i]['n'][0]})", rel_l2(res[i][1], rres[k][1]), 1e-3)


def test_tile_sizes_are_bitwise_identical(device):
    """The LDS-tiled smoothing legs reproduce the global red-black sweeps
    exactly (the tile halo covers the legs' dependency cone), so every tile
    edge gives bitwise the same projections; a small LDS tail puts several
    levels on the tiled path.  (An in-place up leg, whose halo reads race
    with the neighbouring tiles' writes, fails this.)"""
    from dfm.batch import BatchEngine

    probs = small_problems([256, 192, 128, 64], steps=4, R=4, seed=90)
    out = {}
    for tile in (16, 32):
        eng = BatchEngine(probs, tol=1e-6, tile=tile, tail_cells=256)
        res, _ = eng.run()
        st = eng.stats()
        assert st["noconv"] == 0 and st["max_iters"] <= 12, (tile, st)
        out[tile] = res
    for (la, ga), (lb, gb) in zip(out[16], out[32]):
        assert la == lb and np.array_equal(ga, gb)
