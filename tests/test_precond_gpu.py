"""Block-Jacobi / RAS preconditioner on the GPU (precond.cu:407-895 build,
1303-1410 apply; here block inverses applied by one batched GEMM): the
device apply against numpy (each extended block's sub-matrix of Z solved by
LAPACK, own unknowns kept), the one-RWG blocks
(the 2 x 2 self-interaction preconditioner), GMRES with it against a direct
solve, and an exactly block-diagonal matrix where it is the inverse."""
import numpy as np
import pytest

import bem_amd

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx(bem):
    c = bem.Context(0)
    yield c
    c.close()


def _setup(ctx, oracle, shape="sphere", ref=1, param=None, ka=1.0, m=1.3116 + 0.01j):
    verts, tris = oracle.make_mesh(shape, ref, param)
    N = ctx.set_mesh(verts, tris, 7)
    ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m), z64=True, z32=False)
    return verts, tris, N, ctx.get_Z()


def _numpy_apply(Z, N, parts, R):
    out = np.zeros_like(R)
    for own, ext in parts:
        rows = np.concatenate([ext, N + ext])
        y = np.linalg.solve(Z[np.ix_(rows, rows)], R[rows])
        B = len(own)
        keep = np.concatenate([np.arange(B), len(ext) + np.arange(B)])
        out[rows[keep]] = y[keep]
    return out


@pytest.mark.parametrize("max_block,layers", [(100, 1), (60, 0), (1, 0)])
def test_apply_matches_blockwise_lapack(ctx, oracle, max_block, layers):
    print(f"block apply max_block={max_block} layers={layers}", flush=True)
    verts, tris, N, Z = _setup(ctx, oracle)
    nb, t = ctx.block_jacobi(2.0, max_block, layers)
    parts = bem_amd.block_partition(verts, tris, 7, 2.0, max_block, layers)
    assert nb == len(parts)
    rs = np.random.RandomState(4)
    R = rs.randn(2 * N, 5) + 1j * rs.randn(2 * N, 5)
    Zo = ctx.precond_apply(R, "block")
    want = _numpy_apply(Z, N, parts, R)
    assert np.abs(Zo - want).max() <= 1e-10 * np.abs(want).max()


def test_gmres_with_block_preconditioner(ctx, oracle, bem):
    print("gmres block-Jacobi RAS", flush=True)
    verts, tris, N, Z = _setup(ctx, oracle, "sphere", 2, ka=2.0)
    nb, t = ctx.block_jacobi(2.0, 120, 1)
    RT, w = bem.orientations(2, 2, 1)
    kh, e0 = bem.planewave_orient_rhs(RT)
    ctx.rhs_planewave(2.0, 1.0, kh, e0)
    X, st = ctx.solve_gmres(None, tol=1e-9, restart=80, max_iter=400, nrhs=len(kh), precond="block")
    print("block-Jacobi GMRES:", nb, "blocks, build %.1f ms," % t, st, flush=True)
    assert st["converged"] == 1
    Xd = np.linalg.solve(Z, ctx.rhs_planewave(2.0, 1.0, kh, e0, return_B=True))
    assert np.linalg.norm(X - Xd) / np.linalg.norm(Xd) < 1e-7
    # unpreconditioned GMRES needs more iterations for the same tolerance
    _, st0 = ctx.solve_gmres(None, tol=1e-9, restart=80, max_iter=400, nrhs=len(kh), precond="none",
                             allow_noconv=True)
    assert st0["iterations"] > st["iterations"]


def test_block_diagonal_matrix_is_solved_at_once(ctx, oracle):
    print("block-diagonal matrix", flush=True)
    verts, tris, N, Z = _setup(ctx, oracle)
    parts = bem_amd.block_partition(verts, tris, 7, 2.0, 80, 0)
    Zb = np.zeros_like(Z)
    for own, _ in parts:
        rows = np.concatenate([own, N + own])
        Zb[np.ix_(rows, rows)] = Z[np.ix_(rows, rows)]
    ctx.set_Z(Zb)
    ctx.block_jacobi(2.0, 80, 0)
    rs = np.random.RandomState(6)
    B = rs.randn(2 * N, 3) + 1j * rs.randn(2 * N, 3)
    X, st = ctx.solve_gmres(B, tol=1e-12, restart=5, max_iter=10, precond="block")
    assert st["iterations"] <= 2 and st["converged"] == 1
    assert np.linalg.norm(X - np.linalg.solve(Zb, B)) / np.linalg.norm(X) < 1e-11
