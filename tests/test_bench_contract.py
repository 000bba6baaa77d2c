"""bench.py's contract, checked without a GPU: its helpers exist, its byte
tables cover the stages it times, every name its functions call resolves,
and a multi-GPU request exits with a message."""
import ast
import builtins
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_helpers_and_tables():
    b = load_bench()
    for name in ("cpu_baseline", "load_traffic", "run_plain", "run_fused", "run_ensemble",
                 "run_serving", "run_slab", "copy_calibration", "mixed_errors", "form_errors",
                 "oracle_run", "plain_state", "host_info", "run_one_form", "main"):
        assert callable(getattr(b, name)), name
    for s in b.STEP_STAGES + b.TWO_KERNEL_STAGES + b.SPLIT_STAGES:
        assert s in b.PLAIN_BYTES and s in b.KERNEL_NAMES, s
    assert sum(b.PLAIN_BYTES[s] for s in b.SPLIT_STAGES) == 304
    assert sum(b.PLAIN_BYTES[s] for s in b.STEP_STAGES) == 168
    assert sum(b.PLAIN_BYTES[s] for s in b.TWO_KERNEL_STAGES) == 200


def test_every_global_name_resolves():
    """Names loaded in bench.py's functions are module globals, builtins,
    locals or parameters (a removed helper fails here, not on the GPU)."""
    b = load_bench()
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    known = set(dir(b)) | set(dir(builtins))
    for fn in [n for n in tree.body if isinstance(n, ast.FunctionDef)]:
        local = {a.arg for a in fn.args.args + fn.args.kwonlyargs}
        for extra in (fn.args.vararg, fn.args.kwarg):
            if extra is not None:
                local.add(extra.arg)
        for n in ast.walk(fn):
            if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Store):
                local.add(n.id)
            elif isinstance(n, (ast.FunctionDef, ast.Lambda)):
                local |= {a.arg for a in n.args.args}
                if isinstance(n, ast.FunctionDef):
                    local.add(n.name)
            elif isinstance(n, (ast.Import, ast.ImportFrom)):
                local |= {(a.asname or a.name).split(".")[0] for a in n.names}
            elif isinstance(n, ast.ExceptHandler) and n.name:
                local.add(n.name)
            elif isinstance(n, ast.comprehension):
                for t in ast.walk(n.target):
                    if isinstance(t, ast.Name):
                        local.add(t.id)
        for n in ast.walk(fn):
            if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load):
                assert n.id in known or n.id in local, (fn.name, n.id)


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
