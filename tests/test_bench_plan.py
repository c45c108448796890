"""bench.py's argument handling, on the CPU: the step plan it derives from
--steps / --warmup and the one-GPU guard."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_driver_command_plan():
    # the round-end command: --gpus 1 --steps 20 --warmup 5
    args = bench.parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    plan = bench.plan_steps(args.steps, args.warmup)
    assert plan == {"steps": 20, "warmup": 10, "steps_requested": 20, "warmup_requested": 5}


@pytest.mark.parametrize("steps,warmup,exp_steps,exp_warmup", [
    (60, 10, 60, 10), (1, 0, 10, 0), (10, 1, 10, 10), (25, 19, 30, 20), (20, 20, 20, 20)])
def test_plan_rounds_up_to_whole_intervals(steps, warmup, exp_steps, exp_warmup):
    plan = bench.plan_steps(steps, warmup)
    assert (plan["steps"], plan["warmup"]) == (exp_steps, exp_warmup)
    assert plan["steps"] % bench.REINIT_EVERY == 0 and plan["warmup"] % bench.REINIT_EVERY == 0
    assert plan["steps"] >= steps and plan["steps"] - steps < bench.REINIT_EVERY


@pytest.mark.parametrize("steps,warmup", [(0, 5), (20, -1)])
def test_plan_rejects_nonsense(steps, warmup):
    with pytest.raises(ValueError):
        bench.plan_steps(steps, warmup)


def test_defaults_finish_quickly():
    args = bench.parse_args([])
    assert args.gpus == 1 and args.size == 128
    plan = bench.plan_steps(args.steps, args.warmup)
    assert plan["steps"] <= 60 and plan["warmup"] <= 20


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits(n, capsys):
    with pytest.raises(SystemExit) as e:
        bench.main(["--gpus", str(n)])
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err
