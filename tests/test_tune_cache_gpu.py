"""The tuned GEMM plan persists across processes: the first process of a
system size times the tiles and the trial factorisations and writes the plan
to the JSON plan file (BEM_TUNE_CACHE); a second process of the same size
reads it and factors at once, with the same plan and no trials.  Each process
is a child of the test (one context each, one after the other)."""
import json
import os
import subprocess
import sys

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

_CHILD = r"""
import json, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
import bem_amd
n = 8192
rs = np.random.RandomState(12)
Z = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
with bem_amd.Context(0) as c:
    c.set_Z(Z)
    t0 = time.perf_counter()
    info = c.factor("fp32")
    t1 = time.perf_counter()
    print(json.dumps({"info": info, "plan": c.gemm_plan_info("fp32"), "first_factor_s": t1 - t0}))
"""


def _child(path):
    env = dict(os.environ, BEM_TUNE_CACHE=path)
    r = subprocess.run([sys.executable, "-c", _CHILD, PKG], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_plan_file_replaces_tuning_in_a_second_process(tmp_path):
    path = str(tmp_path / "plans" / "gemm_plans.json")
    print("first process (tunes and trials)", flush=True)
    a = _child(path)
    assert a["info"] == 0 and a["plan"]["trialed"], a
    text = open(path).read()
    assert json.loads(text)["plans"], text  # the file is valid JSON with the plan in it
    print("second process (reads the plan file)", flush=True)
    b = _child(path)
    assert b["info"] == 0
    assert b["plan"] == a["plan"], (a["plan"], b["plan"])
    print(f"first factorisation: {a['first_factor_s']:.3f} s with tuning, {b['first_factor_s']:.3f} s from the file")
    assert b["first_factor_s"] < 0.5 * a["first_factor_s"]
