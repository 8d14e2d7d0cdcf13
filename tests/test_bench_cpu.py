"""bench.py's host logic (no GPU): the batched-method ladder, rung selection,
the one-GPU refusal, and that every function the script calls is defined."""
import ast
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_every_called_name_is_defined():
    src = open(os.path.join(ROOT, "bench.py")).read()
    tree = ast.parse(src)
    defined = {n.name for n in ast.walk(tree) if isinstance(n, (ast.FunctionDef, ast.ClassDef))}
    called = {n.func.id for n in ast.walk(tree)
              if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    import builtins
    missing = sorted(c for c in called - defined if not hasattr(builtins, c))
    assert not missing, missing


def test_ladder_rungs_are_valid_options():
    import bench
    from cvt_amd import _lib
    names = {f for f, _ in _lib.Options._fields_}
    for rung in bench.BATCHED_LADDER:
        assert set(rung) - {"iterations"} <= names, rung
        o = bench.rung_options(_lib, rung)
        assert o.coarse_levels == rung["coarse_levels"] and o.grad_steps == rung["grad_steps"]
        assert o.coarse_pd_iters == 30 and o.use_graph == 1
        assert o.random_refine == rung["random_refine"] and o.coarse_random_refine == 2
        assert o.random_neighbours == rung["random_neighbours"]
        assert o.coarse_random_neighbours == 6
        assert f"{2 * rung['coarse_patch'] + 1}x" in bench.describe_rung(rung)
        assert f"{rung['random_refine']} randomized" in bench.describe_rung(rung)


def test_pick_rung_prefers_fastest_at_or_below_plain_error():
    import bench
    rungs = [{"seconds": 1.0, "bad_1px_percent": 0.9}, {"seconds": 2.0, "bad_1px_percent": 0.5},
             {"seconds": 3.0, "bad_1px_percent": 0.4}]
    pick, met = bench.pick_rung(rungs, 0.6)
    assert met and pick["seconds"] == 2.0
    pick, met = bench.pick_rung(rungs, 0.1)
    assert not met and pick["bad_1px_percent"] == 0.4


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
