"""Result files of the dense path (CPU): the library's JSON writer against the
oracle's byte for byte (output.cpp:5-64), the per-orientation file against
its layout (main.cpp:640-653) and the reader of
orient_convergence_analysis.py:17-25, and the cluster merge
(cluster_bem.py:98-128) on partial results of orientation slices."""
import json
import os
import struct

import numpy as np
import pytest

import bem_amd
from bem_amd import formats


@pytest.fixture(scope="module")
def lib_built():
    if not os.path.exists(bem_amd.LIB_PATH):
        pytest.skip("libbem_amd.so not built")
    return bem_amd.lib()


def _M(seed, nt):
    rs = np.random.RandomState(seed)
    return rs.randn(4, 4, nt) * 10.0 ** rs.randint(-6, 3, size=(4, 4, 1))


def test_json_writer_matches_oracle_bytes(lib_built, oracle, tmp_path):
    th = np.linspace(0, np.pi, 7)
    M = _M(1, 7)
    a, b = str(tmp_path / "lib.json"), str(tmp_path / "orc.json")
    timing = (1.234, 2.0, 3.005, 6.239)
    bem_amd.write_json(a, M, th, 5.0, 1.3116, 0.01, 3, (8, 8, 1), timing)
    oracle.write_json(b, M, th, 5.0, 1.3116, 0.01, 3, (8, 8, 1), timing)
    assert open(a, "rb").read() == open(b, "rb").read()
    d = formats.read_result(a)
    assert np.allclose(np.array(d["mueller"]), M, rtol=1e-8, atol=0)


def test_per_orient_layout_and_reader(lib_built, tmp_path):
    n, nt = 5, 9
    rs = np.random.RandomState(2)
    w = rs.rand(n)
    Mo = rs.randn(n, 4, 4, nt)
    path = str(tmp_path / "r.json.per_orient")
    bem_amd.write_per_orient(path, w, Mo, 4, 3)
    raw = open(path, "rb").read()
    # layout of main.cpp:640-653, built independently of the writer
    want = struct.pack("4i", n, nt, 4, 3) + w.astype("<f8").tobytes() + Mo.reshape(n, 16, nt).astype("<f8").tobytes()
    assert raw == want
    got = formats.read_per_orient(path)
    assert got[:4] == (n, nt, 4, 3)
    assert np.array_equal(got[4], w) and np.array_equal(got[5], Mo.reshape(n, 16, nt))


def test_per_orient_reader_rejects_truncated_file(lib_built, tmp_path):
    path = str(tmp_path / "t.per_orient")
    bem_amd.write_per_orient(path, np.ones(2), np.zeros((2, 16, 3)), 2, 1)
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-8])
    with pytest.raises(ValueError):
        formats.read_per_orient(path)


def test_writers_report_unwritable_paths(lib_built, tmp_path):
    bad = str(tmp_path / "no_such_dir" / "r.json")
    with pytest.raises(bem_amd.BemError):
        bem_amd.write_json(bad, np.zeros((4, 4, 3)), np.zeros(3), 1.0, 1.0, 0.0, 1)
    with pytest.raises(bem_amd.BemError):
        bem_amd.write_per_orient(bad + ".per_orient", np.ones(1), np.zeros((1, 16, 3)), 1, 1)


def test_merge_sums_mueller_and_keeps_longest_timings(lib_built, tmp_path):
    th = np.linspace(0, np.pi, 11)
    parts, Ms = [], []
    timings = [(1.0, 4.0, 0.5, 6.0), (1.1, 3.0, 0.7, 5.5), (0.9, 3.5, 0.6, 7.25)]
    for k, tm in enumerate(timings):
        Ms.append(_M(10 + k, th.size))
        p = str(tmp_path / f"part{k}.json")
        bem_amd.write_json(p, Ms[-1], th, 4.0, 1.31, 0.0, 2, (4, 3, 1), tm)
        parts.append(p)
    out = str(tmp_path / "merged.json")
    assert formats.main(["merge", out] + parts) == 0
    d = json.load(open(out))
    assert np.allclose(np.array(d["mueller"]), sum(Ms), rtol=1e-7, atol=1e-7 * np.abs(sum(Ms)).max())
    assert d["timing"] == {"assembly_s": 1.0, "solve_s": 4.0, "farfield_s": 0.7, "total_s": 7.25}
    assert d["ka"] == 4.0 and d["orient"] == [4, 3, 1] and d["theta"][-1] == 180.0


def test_merge_of_oracle_orientation_slices_equals_full_run(lib_built, oracle, tmp_path):
    """--orient-range slices write partial weighted sums (main.cpp:728-734);
    their merge is the full average to the file's print precision."""
    p = oracle.Problem.from_shape("sphere", 1)
    nt = 19
    th = np.pi * np.arange(nt) / (nt - 1)
    _, Mf, _ = p.dense_orient(1.0, 1.3116 + 0j, 3, 2, 1, ntheta=nt)
    paths = []
    for i0, i1 in [(0, 1), (1, 4), (4, 6)]:
        _, Mp, _ = p.dense_orient(1.0, 1.3116 + 0j, 3, 2, 1, i0=i0, i1=i1, ntheta=nt)
        paths.append(str(tmp_path / f"s{i0}.json"))
        bem_amd.write_json(paths[-1], Mp, th, 1.0, 1.3116, 0.0, 1, (3, 2, 1))
    merged = np.array(formats.merge_results(paths)["mueller"])
    assert np.abs(merged - Mf).max() < 3e-8 * np.abs(Mf).max()


def test_merge_rejects_mismatched_parts(tmp_path):
    a = {"mueller": [[[0.0] * 3] * 4] * 4, "timing": {"solve_s": 0, "farfield_s": 0, "total_s": 0}}
    b = {"mueller": [[[0.0] * 5] * 4] * 4, "timing": {"solve_s": 0, "farfield_s": 0, "total_s": 0}}
    with pytest.raises(ValueError):
        formats.merge_results([a, b])
    with pytest.raises(ValueError):
        formats.merge_results([])
