"""The reference's own C++ dense-path signatures, called by a compiled program.

bem-cpp_amd/bin/ref_api_dense (csrc/cli/ref_api_dense.cpp) is written against
bem_host.h, the reference's interface (pmchwt.h:12-19, solver.h:11-22,
rhs.h:11-16, farfield.h:9-71, orient.h), and makes the calls of the
reference's dense branch in its order: the single-orientation branch
(main.cpp:670-719: assemble_pmchwt, lu_factorize_cuda, compute_rhs_planewave,
lu_solve_cuda, compute_far_field, amplitude_to_mueller) and the orientation
average (main.cpp:720-846: lu_solve_full, FFCacheGPU::upload,
compute_farfield_batch_cuda).  Its Mueller matrices are compared with the
reference twin's golden vectors; the solver shims' return conventions
(solver.cu:7-84) are checked by its `shims` mode.  It runs as a child process.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "bem-cpp_amd", "bin", "ref_api_dense")
MUELLER_GATE = 1e-4  # max_ij,theta |dM_ij| / M11(theta)


def _run(args, tmp_path=None, ntheta=None):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built")
    print("ref_api_dense " + " ".join(args), flush=True)
    r = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


@pytest.mark.parametrize("fix,shape,ref,param,orient", [
    ("c1_mueller.npz", "sphere", 2, 1.0, (0, 0, 0)),
    ("ico1_orient.npz", "sphere", 1, 1.0, (2, 2, 1)),
    ("hex0_orient.npz", "hex", 0, 1.4286, (3, 2, 2)),
])
def test_reference_dense_branch(tmp_path, fix, shape, ref, param, orient):
    g = golden(fix)
    nt = len(g["theta"])
    out = str(tmp_path / "M.bin")
    m = complex(g["m"])
    _run([shape, str(ref), repr(param), repr(float(g["ka"])), repr(m.real), repr(m.imag),
          *[str(o) for o in orient], str(nt), out])
    M = np.fromfile(out, dtype=np.float64).reshape(4, 4, nt)
    err = float((np.abs(M - g["M"]) / g["M"][0, 0][None, None, :]).max())
    print(f"{fix}: Mueller error against the reference twin {err:.2e}")
    assert err < MUELLER_GATE


def test_solver_shim_conventions():
    r = _run(["shims"])
    print(r.stdout)
    assert "FAILED" not in r.stdout
