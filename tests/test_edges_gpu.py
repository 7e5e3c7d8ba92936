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


def _sphere_pipeline(ctx, bem, orient, th):
    ka, m = 1.0, 1.3116 + 0.01j
    RT, w = bem.orientations(*orient)
    kh, e0 = bem.planewave_orient_rhs(RT)
    ctx.rhs_planewave(ka, 1.0, kh, e0)
    return ka, RT, w, len(kh)


def test_new_right_hand_sides_drop_the_solution(ctx, bem):
    # orient_mueller and mueller_grad refuse solutions that belong to earlier
    # right-hand sides (no kernel is launched for the refused calls)
    print("stale solutions after new right-hand sides", flush=True)
    v, t = bem.make_mesh("sphere", 1)
    ctx.set_mesh(v, t, 7)
    ka, m = 1.0, 1.3116 + 0.01j
    ctx.assemble(ka, ka * m, 1.0, 1.0 / abs(m))
    th = np.linspace(0.0, np.pi, 7)
    ka, RT, w, nr = _sphere_pipeline(ctx, bem, (2, 2, 1), th)
    assert ctx.factor("fp32") == 0
    ctx.solve(None, ir_tol=1e-12, ir_max=20, return_X=False, nrhs=nr)
    M0 = ctx.orient_mueller(ka, 1.0, RT, w, th)
    # 32 right-hand sides (another stride), no solve
    ka, RT2, w2, nr2 = _sphere_pipeline(ctx, bem, (4, 4, 1), th)
    for call in (lambda: ctx.orient_mueller(ka, 1.0, RT2, w2, th), lambda: ctx.mueller_grad(np.ones((4, 4, 7)))):
        with pytest.raises(bem.BemError) as e:
            call()
        assert e.value.code == bem.ERR_STATE
    # the first right-hand sides again (same stride), still no solve
    _sphere_pipeline(ctx, bem, (2, 2, 1), th)
    with pytest.raises(bem.BemError) as e:
        ctx.orient_mueller(ka, 1.0, RT, w, th)
    assert e.value.code == bem.ERR_STATE
    ctx.solve(None, ir_tol=1e-12, ir_max=20, return_X=False, nrhs=nr)
    M1 = ctx.orient_mueller(ka, 1.0, RT, w, th)
    assert np.abs(M1 - M0).max() <= 1e-12 * np.abs(M0).max()


def test_adjoint_noconv_reported_and_forward_state_restored(ctx, bem, monkeypatch):
    # an unreachable adjoint tolerance (set from the host) makes the adjoint
    # refinement stop at its sweep limit: the call reports NOCONV, still
    # writes the gradient, and leaves the forward solution in place
    print("adjoint refinement without convergence", flush=True)
    v, t = bem.make_mesh("sphere", 1)
    ctx.set_mesh(v, t, 7)
    ka, m = 1.0, 1.3116 + 0.01j
    ctx.assemble(ka, ka * m, 1.0, 1.0 / abs(m))
    th = np.linspace(0.0, np.pi, 7)
    ka, RT, w, nr = _sphere_pipeline(ctx, bem, (2, 2, 1), th)
    assert ctx.factor("fp32") == 0
    ctx.solve(None, ir_tol=1e-12, ir_max=20, return_X=False, nrhs=nr)
    M0 = ctx.orient_mueller(ka, 1.0, RT, w, th)
    G = np.ones((4, 4, 7))
    g_ok = ctx.mueller_grad(G)
    monkeypatch.setenv("BEM_ADJOINT_TOL", "1e-40")
    with pytest.raises(bem.BemError) as e:
        ctx.mueller_grad(G)
    assert e.value.code == bem.ERR_NOCONV
    g = ctx.mueller_grad(G, allow_noconv=True)
    assert np.all(np.isfinite(g)) and np.abs(g - g_ok).max() <= 1e-6 * np.abs(g_ok).max()
    monkeypatch.delenv("BEM_ADJOINT_TOL")
    M1 = ctx.orient_mueller(ka, 1.0, RT, w, th)
    assert np.array_equal(M1, M0)
    assert np.array_equal(ctx.mueller_grad(G), g_ok)
