"""Edge cases of the factorisation and the solves on the GPU: systems much
smaller than one 256 block (identity padding), more right-hand sides than
one tile row, zero right-hand sides, a singular matrix in FP32, and the
dense GMRES on the same cases, all against LAPACK (numpy.linalg.solve)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx(bem):
    c = bem.Context(0)
    yield c
    c.close()


def _system(n, nrhs, seed, shift=0.0):
    rs = np.random.RandomState(seed)
    Z = rs.randn(n, n) + 1j * rs.randn(n, n) + shift * np.eye(n)
    B = rs.randn(n, nrhs) + 1j * rs.randn(n, nrhs)
    return Z, B


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n", [1, 2, 5])
def test_tiny_systems(ctx, prec, n):
    print(f"tiny n={n} {prec}", flush=True)
    Z, B = _system(n, 3, 100 + n, shift=2.0)
    ctx.set_Z(Z)
    assert ctx.factor(prec) == 0
    X, st = ctx.solve(B, ir_tol=1e-13, ir_max=10 if prec == "fp32" else 0)
    Xr = np.linalg.solve(Z, B)
    assert np.abs(X - Xr).max() <= 1e-11 * np.abs(Xr).max()
    Xg, stg = ctx.solve_gmres(B, tol=1e-13, restart=5, max_iter=20)
    assert np.abs(Xg - Xr).max() <= 1e-11 * np.abs(Xr).max()


def test_many_rhs_more_than_one_tile_row(ctx):
    print("n=600, 200 RHS", flush=True)
    n = 600
    Z, B = _system(n, 200, 7, shift=4 * np.sqrt(n))
    ctx.set_Z(Z)
    assert ctx.factor("fp32") == 0
    X, st = ctx.solve(B, ir_tol=1e-12, ir_max=20)
    assert st["converged"] == 1
    Xr = np.linalg.solve(Z, B)
    assert np.abs(X - Xr).max() <= 1e-10 * np.abs(Xr).max()
    Xg, stg = ctx.solve_gmres(B, tol=1e-12, restart=10, max_iter=30)
    assert np.abs(Xg - Xr).max() <= 1e-10 * np.abs(Xr).max()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_zero_right_hand_sides(ctx, prec):
    print(f"zero RHS columns {prec}", flush=True)
    n = 400
    Z, B = _system(n, 4, 9, shift=3 * np.sqrt(n))
    B[:, 1] = 0.0
    B[:, 3] = 0.0
    ctx.set_Z(Z)
    assert ctx.factor(prec) == 0
    X, st = ctx.solve(B, ir_tol=1e-12, ir_max=10)
    assert np.all(np.isfinite(X)) and st["converged"] == 1
    assert np.all(X[:, 1] == 0) and np.all(X[:, 3] == 0)
    Xr = np.linalg.solve(Z, B)
    assert np.abs(X - Xr).max() <= 1e-10 * np.abs(Xr).max()
    Xg, stg = ctx.solve_gmres(B, tol=1e-12, restart=10, max_iter=30)
    assert np.all(np.isfinite(Xg)) and np.all(Xg[:, 1] == 0) and np.all(Xg[:, 3] == 0)
    assert np.abs(Xg - Xr).max() <= 1e-10 * np.abs(Xr).max()


@pytest.mark.parametrize("n,col", [(64, 10), (700, 300)])
def test_fp32_singular_info(ctx, n, col):
    # a zero column gives the first zero pivot exactly there, as in FP64
    print(f"singular fp32 n={n} col={col}", flush=True)
    rs = np.random.RandomState(3)
    Z = rs.randn(n, n) + 0j
    Z[:, col] = 0.0
    ctx.set_Z(Z)
    assert ctx.factor("fp32") == col + 1
