"""bench.py's host logic (no GPU): the fixed batched method and the ablation
rows are valid options, the one-GPU refusal, and that every function the
script calls is defined."""
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
    # parameters (a callable passed in) are defined names too
    defined |= {a.arg for n in ast.walk(tree) if isinstance(n, (ast.FunctionDef, ast.Lambda))
                for a in n.args.args}
    # and names bound by assignment or a loop target (a callable from a table)
    defined |= {n.id for n in ast.walk(tree)
                if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Store)}
    called ={n.func.id for n in ast.walk(tree)
              if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    import builtins
    missing = sorted(c for c in called - defined if not hasattr(builtins, c))
    assert not missing, missing


def test_batched_method_and_ablation_are_valid_options():
    import bench
    from cvt_amd import _lib
    names = {f for f, _ in _lib.Options._fields_}
    m = bench.BATCHED_METHOD
    assert set(m) - {"iterations"} <= names, m
    o = bench.method_options(_lib, m)
    assert (o.coarse_levels, o.coarse_iters, o.coarse_patch) == (3, 6, 4)
    assert (o.grad_steps, o.random_refine, o.random_neighbours) == (2, 0, 0)
    assert (o.coarse_random_refine, o.coarse_random_neighbours, o.coarse_pd_iters) == (2, 6, 30)
    assert o.coarse_joint_views == 1 and "both views per launch" in bench.describe_method(m)
    assert o.use_graph == 1 and o.sweep == _lib.SWEEP_REDBLACK and o.maxdispdiff == 1.0
    assert "9x9 window" in bench.describe_method(m)
    rows = bench.ablation(4)
    assert rows[-1][1] == m  # the last row is the method itself
    plain = bench.method_options(_lib, rows[0][1])
    assert (plain.coarse_levels, plain.grad_steps, plain.random_refine) == (0, 0, 2)
    assert rows[0][1]["iterations"] == 4 and rows[1][1]["iterations"] == 2
    assert bench.method_options(_lib, rows[1][1]).coarse_levels == 0
    nograd = bench.method_options(_lib, rows[2][1])
    assert nograd.grad_steps == 0 and nograd.coarse_levels == 3
    assert bench.ablation(8)[1][1]["iterations"] == 4


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
