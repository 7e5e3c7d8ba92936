"""The dense command-line driver end to end on the GPU (main.cpp:660-854 with
output.cpp:5-64 and the per-orientation file of main.cpp:640-653): the
result file against the FP64 oracle, the per-orientation file against the
result file, and a cluster run (orientation slices merged as
cluster_bem.py:98-128 does) against the full run.  The driver runs as a
child process."""
import os
import subprocess

import numpy as np
import pytest

import bem_amd
from bem_amd import formats

pytestmark = pytest.mark.gpu

ARGS = ["--ka", "1.0", "--ri", "1.3116", "0", "--shape", "sphere", "--ref", "1", "--orient", "3", "2", "1",
        "--ntheta", "19"]


def _run(tmp_path, name, extra=()):
    if not os.path.exists(bem_amd.CLI_PATH):
        pytest.fail(f"{bem_amd.CLI_PATH} not built")
    out = str(tmp_path / name)
    print(f"bem_dense {' '.join(ARGS + list(extra))} --out {out}", flush=True)
    r = subprocess.run([bem_amd.CLI_PATH] + ARGS + list(extra) + ["--out", out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return out


@pytest.fixture(scope="module")
def full(tmp_path_factory):
    return _run(tmp_path_factory.mktemp("cli"), "full.json")


def test_cli_result_matches_oracle(full, oracle):
    d = formats.read_result(full)
    M = np.array(d["mueller"])
    p = oracle.Problem.from_shape("sphere", 1)
    _, Mo, _ = p.dense_orient(1.0, 1.3116 + 0j, 3, 2, 1, ntheta=19)
    assert (np.abs(M - Mo) / Mo[0, 0][None, None, :]).max() < 1e-4
    assert d["orient"] == [3, 2, 1] and d["ntheta"] == 19 and d["refinements"] == 1
    assert np.allclose(d["theta"], np.linspace(0, 180, 19), atol=5e-5)


def test_cli_per_orient_file_sums_to_the_average(full):
    n, nt, na, nb, w, Mo = formats.read_per_orient(full + ".per_orient")
    assert (n, nt, na, nb) == (6, 19, 3, 2)
    assert abs(w.sum() - 1.0) < 1e-12
    M = np.array(formats.read_result(full)["mueller"]).reshape(16, nt)
    avg = np.einsum("o,oet->et", w, Mo) / 1.0 ** 2  # |k_ext|^2 = ka^2 = 1
    assert np.abs(avg - M).max() < 2e-8 * np.abs(M).max()


def test_cli_orientation_slices_merge_to_the_full_run(full, tmp_path):
    a = _run(tmp_path, "a.json", ["--orient-range", "0", "2"])
    b = _run(tmp_path, "b.json", ["--orient-range", "2", "6"])
    merged = np.array(formats.merge_results([a, b])["mueller"])
    M = np.array(formats.read_result(full)["mueller"])
    assert np.abs(merged - M).max() < 1e-7 * np.abs(M).max()
    # the slices' per-orientation files are the full run's, in order
    fa, fb, ff = (formats.read_per_orient(x + ".per_orient") for x in (a, b, full))
    assert fa[0] == 2 and fb[0] == 4
    assert np.array_equal(np.concatenate([fa[4], fb[4]]), ff[4])
    Mo = np.concatenate([fa[5], fb[5]])
    assert np.abs(Mo - ff[5]).max() < 1e-7 * np.abs(ff[5]).max()
