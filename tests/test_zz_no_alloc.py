"""bem_farfield_noalloc (no device allocation: every device buffer in the
library's static per-GPU array) against bem_farfield, through the C ABI.
Both run the same kernels on the same inputs, so the far fields agree bit for
bit; two calls in a row, the second smaller, check that the shared array's
leftover contents of the first call do not leak into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rs, N, n_calls, n_orient, ndir):
    cJ = rs.randn(n_calls, N) + 1j * rs.randn(n_calls, N)
    cM = rs.randn(n_calls, N) + 1j * rs.randn(n_calls, N)
    rh = rs.randn(n_orient, ndir, 3)
    rh /= np.linalg.norm(rh, axis=2)[:, :, None]
    return cJ, cM, rh


def test_noalloc_farfield_matches_farfield(bem):
    print("farfield_noalloc against farfield: sphere ref 2, then a smaller call", flush=True)
    rs = np.random.RandomState(42)
    v, t = bem.make_mesh("sphere", 2)
    with bem.Context(0) as ctx:
        N = ctx.set_mesh(v, t, 7)
        k = 1.7 + 0.0j
        for n_calls, n_orient, ndir in [(4, 2, 37), (1, 1, 5)]:
            print(f"n_calls={n_calls} n_orient={n_orient} ndir={ndir}", flush=True)
            cJ, cM, rh = _case(rs, N, n_calls, n_orient, ndir)
            F0 = ctx.farfield(cJ, cM, rh, k, 1.0, n_calls, n_orient)
            F1 = ctx.farfield(cJ, cM, rh, k, 1.0, n_calls, n_orient, noalloc=True)
            assert np.all(np.isfinite(F1))
            assert np.array_equal(F0, F1), np.abs(F0 - F1).max()


def test_noalloc_farfield_rejects_oversized_calls(bem):
    # 64 calls on 8960 quadrature points need ~55 MB of staging, above the
    # array: refused before anything is launched
    print("farfield_noalloc size limit", flush=True)
    v, t = bem.make_mesh("sphere", 3)
    rs = np.random.RandomState(1)
    with bem.Context(0) as ctx:
        N = ctx.set_mesh(v, t, 7)
        cJ, cM, rh = _case(rs, N, 64, 32, 3)
        with pytest.raises(bem.BemError) as e:
            ctx.farfield(cJ, cM, rh, 1.0, 1.0, 64, 32, noalloc=True)
        assert e.value.code == bem.ERR_ARG
