"""Restarted flexible GMRES on the FP64 matrix with LU preconditioning
(block_gmres.h:52-61, python/bem_core.py:1071-1123) on the GPU: against
LAPACK (numpy.linalg.solve) on a seeded random system, the reference's A_ext
preconditioner (the factors of Z at m = 1 reused for another refractive
index) against a direct FP64 solve of the same system, and the state rules
for factors of an earlier matrix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx(bem):
    c = bem.Context(0)
    yield c
    c.close()


def test_gmres_with_own_fp32_factors_matches_lapack(ctx):
    rs = np.random.RandomState(21)
    n = 900
    Z = rs.randn(n, n) + 1j * rs.randn(n, n) + 40 * np.eye(n)
    B = rs.randn(n, 12) + 1j * rs.randn(n, 12)
    print("random n=900, 12 RHS", flush=True)
    ctx.set_Z(Z)
    assert ctx.factor("fp32") == 0
    X, st = ctx.solve_gmres(B, tol=1e-12, restart=20, max_iter=60)
    Xr = np.linalg.solve(Z, B)
    assert st["converged"] == 1 and st["residual"] <= 1e-12
    assert st["iterations"] <= 6  # FP32 factors of Z itself: a few iterations to FP64 accuracy
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) < 1e-10
    R = B - Z @ X
    assert (np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)).max() <= 1e-12


def test_gmres_aext_preconditioner_reused_for_another_index(ctx, oracle, bem):
    """python/bem_core.py:1071-1123: A_ext = Z(m = 1) / 2 for eta_ext = 1;
    its LU preconditions GMRES on Z(m) for any m."""
    p = oracle.Problem.from_shape("sphere", 1)
    ka, m = 1.0, 1.3116 + 0.01j
    print("sphere ref 1, A_ext preconditioner", flush=True)
    ctx.set_mesh(p.verts, p.tris, 7)
    RT, w = bem.orientations(2, 2, 1)
    kh, e0 = bem.planewave_orient_rhs(RT)
    # factors of the exterior-only operator (m = 1: k_int = k_ext, eta_int = 1)
    ctx.assemble(complex(ka), complex(ka), 1.0, 1.0)
    assert ctx.factor("fp32") == 0
    # the dielectric matrix, FP64 only: the FP32 factors stay
    ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m), z64=True, z32=False)
    ctx.rhs_planewave(ka, 1.0, kh, e0)
    with pytest.raises(bem.BemError):
        ctx.solve(None, ir_tol=1e-10, nrhs=len(kh))  # stale factors: the direct solve refuses them
    X, st = ctx.solve_gmres(None, tol=1e-10, restart=60, max_iter=120, nrhs=len(kh))
    assert st["converged"] == 1 and st["residual"] <= 1e-10
    assert st["iterations"] <= 60, st
    # the same system solved directly in FP64
    ctx.factor("fp64")
    Xd, std = ctx.solve(None, ir_tol=0.0, ir_max=0, nrhs=len(kh))
    assert np.linalg.norm(X - Xd) / np.linalg.norm(Xd) < 1e-8
    # and against the FP64 oracle's solution of the same plane waves (the
    # device Z agrees with the oracle's to the assembly's accuracy, ~1e-8)
    info, Mo, Xo = p.dense_orient(ka, m, 2, 2, 1, ntheta=7)
    assert info == 0
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-6
    print("A_ext-preconditioned GMRES iterations:", st["iterations"], flush=True)


def test_stale_factor_rules(ctx, bem):
    rs = np.random.RandomState(5)
    n = 300
    Z = rs.randn(n, n) + 1j * rs.randn(n, n) + 20 * np.eye(n)
    B = rs.randn(n, 3) + 0j
    with pytest.raises(bem.BemError):
        ctx.set_Z(Z) or ctx.solve_gmres(B)  # no factors yet
    assert ctx.factor("fp64") == 0
    Z2 = Z + 0.01 * rs.randn(n, n)
    ctx.set_Z(Z2)  # factors of Z kept for GMRES on Z2
    with pytest.raises(bem.BemError):
        ctx.solve(B)
    X, st = ctx.solve_gmres(B, tol=1e-13, restart=10, max_iter=50)
    assert st["converged"] == 1
    assert np.linalg.norm(X - np.linalg.solve(Z2, B)) / np.linalg.norm(X) < 1e-11
    # a limit too small to converge reports it (and keeps the last iterate)
    X1, st1 = ctx.solve_gmres(B, tol=1e-15, restart=1, max_iter=1, allow_noconv=True)
    assert st1["converged"] == 0 and st1["iterations"] == 1
    with pytest.raises(bem.BemError):
        ctx.solve_gmres(B, tol=1e-15, restart=1, max_iter=1)
    ctx.factor("fp64")  # factors of Z2: the direct solve works again
    X2, _ = ctx.solve(B, ir_max=0)
    assert np.linalg.norm(X2 - X) / np.linalg.norm(X) < 1e-11
