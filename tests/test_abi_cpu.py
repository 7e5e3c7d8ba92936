"""CPU checks of the C ABI library: it loads without a GPU, exports exactly the
entry points include/bem_abi.h declares, and its host-only generators
(meshes, orientation grids) agree bit for bit with the oracle.  No device
work is started here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bem_abi.h")
LIB = os.path.join(ROOT, "bem-cpp_amd", "lib", "libbem_amd.so")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bem_[A-Za-z0-9_]+)\s*\(", text)))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if re.search(r"\sT\sbem_", ln)})


@pytest.fixture(scope="module")
def so():
    if not os.path.exists(LIB):
        pytest.skip("libbem_amd.so not built")
    return ctypes.CDLL(LIB)


def test_header_declares_the_boundary():
    names = _declared()
    for must in ["bem_create", "bem_set_mesh", "bem_assemble", "bem_factor", "bem_solve", "bem_rhs_planewave",
                 "bem_farfield", "bem_orient_mueller", "bem_mueller_grad", "bem_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol(so):
    declared = _declared()
    for name in declared:
        assert hasattr(so, name), name
    # and nothing undeclared under the bem_ prefix
    assert _exported() == declared


def test_binding_covers_the_header(so):
    import bem_amd
    lib = bem_amd.lib()
    for name in _declared():
        fn = getattr(lib, name)
        assert fn.restype is not None or name == "bem_destroy", name


@pytest.mark.parametrize("shape,ref,param", [("sphere", 0, 1.0), ("sphere", 2, 1.0), ("hex", 1, 0.7),
                                             ("hex", 3, 1.4286)])
def test_meshes_match_oracle(oracle, shape, ref, param):
    import bem_amd
    v, t = bem_amd.make_mesh(shape, ref, param)
    vo, to = oracle.make_mesh(shape, ref, param)
    assert np.array_equal(t, to)
    assert np.array_equal(v, vo)


@pytest.mark.parametrize("grid", [(8, 8, 1, 1, 1, False, True), (3, 4, 2, 2, 1, True, False),
                                  (5, 3, 4, 1, 2, False, True)])
def test_orientations_match_oracle(oracle, grid):
    import bem_amd
    RT, w = bem_amd.orientations(*grid[:3], beta_sym=grid[3], gamma_sym=grid[4], mirror=grid[5], sort=grid[6])
    RTo, wo = oracle.orientations(*grid[:3], beta_sym=grid[3], gamma_sym=grid[4], mirror=grid[5], sort=grid[6])
    assert np.array_equal(RT, RTo) and np.array_equal(w, wo)
    assert abs(w.sum() - 1.0) < 1e-13


def test_mesh_counts_rejects_bad_shape(so):
    nv, nt = ctypes.c_int(), ctypes.c_int()
    assert so.bem_mesh_counts(7, 1, ctypes.byref(nv), ctypes.byref(nt)) < 0
    assert so.bem_mesh_counts(0, -1, ctypes.byref(nv), ctypes.byref(nt)) < 0
