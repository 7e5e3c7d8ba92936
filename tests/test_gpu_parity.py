"""GPU parity tests: the HIP library (through the C ABI) against the FP64
oracle and the reference twin's golden vectors, on inputs the oracle finishes
in seconds.  Tolerances are stated per test."""
import numpy as np
import pytest

from conftest import golden, python_to_sorted

pytestmark = pytest.mark.gpu

# normwise backward error bound of an unrefined FP64 solve, max|ZX - B| /
# (max|Z| max|X| n): about 10 eps, where LAPACK reaches 0.5-3.3e-16 on the
# seeded systems of these tests (see test_lu_fp64_vs_numpy)
BE_FP64 = 2e-15


@pytest.fixture(scope="module")
def ctx(bem):
    c = bem.Context(0)
    yield c
    c.close()


def _perm2(keys):
    order = python_to_sorted(keys)
    N = len(order)
    return np.concatenate([order, order + N])


# Z: far triangle pairs use FP32 sin/cos/exp of an FP64-reduced phase (relative
# error ~1e-7 per point pair), near pairs are FP64.  Entry-wise bound:
# |dZ| <= 2e-6 |Z| + 1e-7 max|Z|.
def _z_close(Z, Zr, rtol=2e-6, atol=1e-7):
    scale = np.abs(Zr).max()
    bad = np.abs(Z - Zr) > rtol * np.abs(Zr) + atol * scale
    return not bad.any(), np.abs(Z - Zr).max() / scale


@pytest.mark.parametrize("shape,ref,param,ka,m", [
    ("sphere", 0, 1.0, 1.0, 1.3116 + 0j),
    ("sphere", 1, 1.0, 2.0, 1.33 + 0.02j),
    ("sphere", 2, 1.0, 1.0, 1.3116 + 0j),
    ("hex", 1, 1.4286, 3.0, 1.3116 + 0.01j),
    ("hex", 0, 0.7, 1.5, 1.3116 + 0j),
])
def test_Z_vs_oracle(oracle, ctx, shape, ref, param, ka, m):
    p = oracle.Problem.from_shape(shape, ref, param)
    ctx.set_mesh(p.verts, p.tris, 7)
    args = (complex(ka), ka * m, 1.0, 1.0 / abs(m))
    ctx.assemble(*args)
    Z = ctx.get_Z()
    Zo = p.assemble_pmchwt(*args)
    ok, err = _z_close(Z, Zo)
    print(f"{shape} ref={ref}: max|dZ|/max|Z| = {err:.2e}")
    assert ok, err
    S = np.concatenate([np.ones(p.N), -np.ones(p.N)])
    assert np.abs(Z.T - S[:, None] * Z * S[None, :]).max() < 1e-12 * np.abs(Z).max()


@pytest.mark.parametrize("quad", [4, 13])
def test_Z_other_quadratures(oracle, ctx, quad):
    p = oracle.Problem.from_shape("sphere", 1, quad_order=quad)
    ctx.set_mesh(p.verts, p.tris, quad)
    args = (1.5 + 0j, 1.5 * (1.4 + 0.01j), 1.0, 1.0 / abs(1.4 + 0.01j))
    ctx.assemble(*args)
    ok, err = _z_close(ctx.get_Z(), p.assemble_pmchwt(*args))
    assert ok, err


@pytest.mark.parametrize("tag", ["ico0_real", "hex0"])
def test_Z_vs_reference_golden(ctx, tag):
    g = golden(f"{tag}_Z.npz")
    ctx.set_mesh(g["verts"], g["tris"], 7)
    ka, m = float(g["ka"]), complex(g["m"])
    ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
    P = _perm2(g["keys"])
    ok, err = _z_close(ctx.get_Z(), g["Z"][np.ix_(P, P)])
    assert ok, err


def test_LK_vs_oracle(oracle, ctx):
    p = oracle.Problem.from_shape("sphere", 1)
    ctx.set_mesh(p.verts, p.tris, 7)
    k = 2.5 + 0.03j
    L, K = ctx.assemble_LK(k)
    Lo, Ko = p.assemble_LK(k)
    assert _z_close(L, Lo)[0] and _z_close(K, Ko, atol=1e-6)[0]


# ---------------------------------------------------------------- LU
@pytest.mark.parametrize("n,nrhs", [(40, 1), (300, 5), (700, 17), (1024, 3), (2500, 2)])
def test_lu_fp64_vs_numpy(ctx, n, nrhs):
    rs = np.random.RandomState(42)
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    B = rs.randn(n, nrhs) + 1j * rs.randn(n, nrhs)
    ctx.set_Z(Z)
    assert ctx.factor("fp64") == 0
    X, st = ctx.solve(B, ir_max=0)
    Xr = np.linalg.solve(Z, B)
    cond = np.linalg.cond(Z)
    assert np.abs(X - Xr).max() <= 1e-13 * cond * np.abs(Xr).max()
    # normwise backward error; a stable pivoted LU gives O(n eps).  The FP64
    # factorisation and solves substitute with the diagonal blocks (no
    # explicit inverse), as Zgetrf / Zgetrs do
    be = np.abs(Z @ X - B).max() / (np.abs(Z).max() * np.abs(X).max() * n)
    print(f"n={n}: FP64 backward error {be:.2e}")
    assert be < BE_FP64, be
    # one FP64 refinement sweep brings it to LAPACK's level
    X1, st1 = ctx.solve(B, ir_tol=0.0, ir_max=1, allow_noconv=True)
    be1 = np.abs(Z @ X1 - B).max() / (np.abs(Z).max() * np.abs(X1).max() * n)
    assert be1 < 5e-16, be1


@pytest.mark.parametrize("n,nrhs", [(300, 4), (1100, 32)])
def test_lu_fp32_refined_vs_numpy(ctx, n, nrhs):
    rs = np.random.RandomState(42)
    Z = rs.randn(n, n) + 1j * rs.randn(n, n) + 4 * np.sqrt(n) * np.eye(n)
    B = rs.randn(n, nrhs) + 1j * rs.randn(n, nrhs)
    ctx.set_Z(Z)
    assert ctx.factor("fp32") == 0
    X, st = ctx.solve(B, ir_tol=1e-12, ir_max=20)
    assert st["converged"] == 1 and st["residual"] <= 1e-12
    assert st["residual0"] > st["residual"]
    # the reported residual is max_j ||B_j - Z X_j|| / ||B_j|| of the returned X
    # (its own rounding, ~eps |Z| |X| / |B|, is far below the value itself)
    res_np = (np.linalg.norm(Z @ X - B, axis=0) / np.linalg.norm(B, axis=0)).max()
    assert abs(st["residual"] - res_np) <= 0.05 * res_np + 1e-15, (st["residual"], res_np)
    Xr = np.linalg.solve(Z, B)
    assert np.abs(X - Xr).max() <= 1e-10 * np.abs(Xr).max()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lu_zero_diagonal_needs_pivoting(ctx, prec):
    # every leading minor of order 1 is zero: no pivot-free elimination exists
    rs = np.random.RandomState(11)
    n = 600
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    Z[np.arange(n), np.arange(n)] = 0.0
    B = rs.randn(n, 3) + 1j * rs.randn(n, 3)
    ctx.set_Z(Z)
    assert ctx.factor(prec) == 0
    X, st = ctx.solve(B, ir_tol=1e-12, ir_max=30 if prec == "fp32" else 0)
    be = np.abs(Z @ X - B).max() / (np.abs(Z).max() * np.abs(X).max() * n)
    assert be < BE_FP64
    # the pivots form a permutation
    ip = ctx.get_pivots(n) - 1
    perm = np.arange(n)
    for i in range(n):
        perm[[i, ip[i]]] = perm[[ip[i], i]]
    assert sorted(perm.tolist()) == list(range(n))


@pytest.mark.parametrize("n,col", [(64, 10), (700, 300), (700, 699)])
def test_lu_singular_info(ctx, n, col):
    # a zero column gives the first zero pivot exactly there (LAPACK info = col + 1)
    rs = np.random.RandomState(3)
    Z = rs.randn(n, n) + 0j
    Z[:, col] = 0.0
    ctx.set_Z(Z)
    assert ctx.factor("fp64") == col + 1


def test_lu_get_set_roundtrip(ctx):
    rs = np.random.RandomState(5)
    n = 90
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    b = rs.randn(n) + 1j * rs.randn(n)
    ctx.set_Z(Z)
    assert ctx.factor("fp64") == 0
    LU, ip = ctx.get_lu(n), ctx.get_pivots(n)
    P = np.eye(n)[:, :]
    perm = np.arange(n)
    for i in range(n):
        j = ip[i] - 1
        perm[[i, j]] = perm[[j, i]]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.abs(L @ U - Z[perm]).max() < 1e-12 * np.abs(Z).max()
    ctx.set_lu(LU, ip)
    x, _ = ctx.solve(b, ir_max=0)
    assert np.abs(Z @ x[:, 0] - b).max() < 1e-11 * np.abs(b).max() * np.linalg.cond(Z)


# ---------------------------------------------------------------- RHS / far field
def test_rhs_vs_oracle(oracle, ctx, bem):
    p = oracle.Problem.from_shape("hex", 1, 1.4286)
    ctx.set_mesh(p.verts, p.tris, 7)
    RT, w = bem.orientations(3, 2, 1)
    kh, e0 = bem.planewave_orient_rhs(RT)
    k = 2.0 + 0j
    B = ctx.rhs_planewave(k, 1.0, kh, e0, return_B=True)
    Bo = np.stack([p.rhs_planewave(k, 1.0, e0[j], kh[j]) for j in range(len(kh))], axis=1)
    assert np.abs(B - Bo).max() < 1e-13 * np.abs(Bo).max()


def test_farfield_vs_oracle(oracle, ctx):
    p = oracle.Problem.from_shape("sphere", 1)
    ctx.set_mesh(p.verts, p.tris, 7)
    rs = np.random.RandomState(42)
    N = p.N
    cJ = rs.randn(3, N) + 1j * rs.randn(3, N)
    cM = rs.randn(3, N) + 1j * rs.randn(3, N)
    rh = rs.randn(2, 7, 3)
    rh /= np.linalg.norm(rh, axis=2)[:, :, None]
    k = 1.7 + 0j
    Fv = ctx.farfield(cJ, cM, rh, k, 1.0, n_calls=3, n_orient=2)
    for c in range(3):
        Fo = p.farfield_vec(cJ[c], cM[c], k, 1.0, rh[c // 2])
        assert np.abs(Fv[c] - Fo).max() < 1e-12 * np.abs(Fo).max()


# ---------------------------------------------------------------- pipelines
MUELLER_GATE = 1e-4  # max_ij,theta |dM_ij| / M11(theta)


def _mueller_err(M, Mr):
    return (np.abs(M - Mr) / Mr[0, 0][None, None, :]).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c1_single_vs_reference(ctx, bem, precision):
    g = golden("c1_mueller.npz")
    v, t = bem.make_mesh("sphere", 2)
    ctx.set_mesh(v, t, 7)
    m = 1.3116 + 0j
    ctx.assemble(1.0 + 0j, 1.0 * m, 1.0, 1.0 / abs(m))
    RT, w = bem.lab_single_orientation()
    kh, e0 = bem.planewave_orient_rhs(RT)
    ctx.rhs_planewave(1.0, 1.0, kh, e0)
    assert ctx.factor(precision) == 0
    X, st = ctx.solve(None, ir_tol=1e-10, ir_max=10, return_X=False, nrhs=2)
    M = ctx.orient_mueller(1.0, 1.0, RT, w, g["theta"])
    err = _mueller_err(M, g["M"])
    print(f"C1 {precision}: Mueller error {err:.2e}, stats {st}")
    assert err < MUELLER_GATE


@pytest.mark.parametrize("fix,shape,ref,param,ka,orient", [
    ("ico1_orient.npz", "sphere", 1, 1.0, 1.0, (2, 2, 1)),
    ("hex0_orient.npz", "hex", 0, 1.4286, 1.5, (3, 2, 2)),
])
def test_orientation_average_vs_reference(ctx, bem, fix, shape, ref, param, ka, orient):
    g = golden(fix)
    v, t = bem.make_mesh(shape, ref, param)
    ctx.set_mesh(v, t, 7)
    m = 1.3116 + 0j
    ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
    RT, w = bem.orientations(*orient)
    kh, e0 = bem.planewave_orient_rhs(RT)
    ctx.rhs_planewave(ka, 1.0, kh, e0)
    assert ctx.factor("fp32") == 0
    ctx.solve(None, ir_tol=1e-10, ir_max=10, return_X=False, nrhs=len(kh))
    M = ctx.orient_mueller(ka, 1.0, RT, w, g["theta"])
    assert _mueller_err(M, g["M"]) < MUELLER_GATE


def test_orient_vs_oracle_lossy_hex(oracle, ctx, bem):
    p = oracle.Problem.from_shape("hex", 1, 1.4286)
    m = 1.3116 + 0.01j
    ka = 2.0
    info, Mo, Xo = p.dense_orient(ka, m, 3, 3, 1, ntheta=37)
    ctx.set_mesh(p.verts, p.tris, 7)
    ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
    RT, w = bem.orientations(3, 3, 1)
    kh, e0 = bem.planewave_orient_rhs(RT)
    ctx.rhs_planewave(ka, 1.0, kh, e0)
    assert ctx.factor("fp32") == 0
    X, st = ctx.solve(None, ir_tol=1e-10, ir_max=10, nrhs=len(kh))
    th = np.linspace(0, np.pi, 37)
    M = ctx.orient_mueller(ka, 1.0, RT, w, th)
    assert _mueller_err(M, Mo) < MUELLER_GATE
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-5


def test_autotune_reports_and_factorisation_stays_exact(ctx):
    # the tuner times every tile configuration at this size; whatever it
    # picks, the factorisation keeps its backward-error bound
    rs = np.random.RandomState(8)
    n = 900
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    B = rs.randn(n, 16) + 1j * rs.randn(n, 16)
    ctx.set_Z(Z)
    rep = ctx.autotune(16)
    for key in ("fp32_trailing", "fp32_solve", "fp64_trailing", "fp64_solve"):
        assert all(v > 0 for v in rep[key].values()), rep
    assert rep["fp32_choice"][0] in ctx.GEMM_TILES and rep["fp64_choice"][1] in ctx.GEMM_TILES64
    for prec in ("fp64", "fp32"):
        assert ctx.factor(prec) == 0
        X, st = ctx.solve(B, ir_tol=1e-12, ir_max=30 if prec == "fp32" else 0)
        be = np.abs(Z @ X - B).max() / (np.abs(Z).max() * np.abs(X).max() * n)
        assert be < BE_FP64


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
def test_every_gemm_tile_and_capped_grid_factor_exactly(ctx, cfg):
    # each tile configuration as the only one, with the capped trailing
    # updates on a 7-block grid, so every block loops over several tiles and
    # the last round is partial; FP64 and FP32 + refinement keep the
    # backward-error bound of an unpivoted-growth-free LU
    rs = np.random.RandomState(11)
    n = 1000
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    B = rs.randn(n, 8) + 1j * rs.randn(n, 8)
    ctx.set_Z(Z)
    X_ref = np.linalg.solve(Z, B)
    for prec in ("fp64", "fp32"):
        ctx.gemm_plan(prec, cfg, cfg, cap=7)
        assert ctx.factor(prec) == 0
        X, st = ctx.solve(B, ir_tol=1e-12, ir_max=30 if prec == "fp32" else 0)
        be = np.abs(Z @ X - B).max() / (np.abs(Z).max() * np.abs(X).max() * n)
        assert be < BE_FP64, (prec, cfg, be)
        assert np.abs(X - X_ref).max() <= 1e-9 * np.abs(X_ref).max()
    with pytest.raises(RuntimeError):
        ctx.gemm_plan("fp32", len(ctx.GEMM_TILES), 0)


def test_factor_trials_pick_a_tile_and_keep_results(ctx):
    # from 8192 unknowns on, the first factorisation of each precision times
    # one whole trial factorisation per tile configuration (on the context's
    # own matrix) and keeps the fastest (ties within 1 % go to the tile that
    # is faster alone); the factors are the same either way
    rs = np.random.RandomState(12)
    n = 8192
    Z = (rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))).astype(np.complex128)
    B = rs.standard_normal((n, 4)) + 1j * rs.standard_normal((n, 4))
    ctx.set_Z(Z)
    for prec in ("fp32", "fp64"):
        assert ctx.factor(prec) == 0
        info = ctx.gemm_plan_info(prec)
        assert info["trialed"], info
        tiles = ctx.UPDATE_TILES32 if prec == "fp32" else ctx.GEMM_TILES64[:4]  # the 32 x 32 tile is not tried
        assert sorted(info["trial_factor_ms"]) == sorted(tiles), info
        # the fastest trial, or one within 1 % of it
        ms = info["trial_factor_ms"]
        assert ms[info["update"]] <= 1.01 * min(ms.values()) + 1e-9, info
        X, st = ctx.solve(B, ir_tol=1e-12, ir_max=30 if prec == "fp32" else 0)
        be = np.abs(Z @ X - B).max() / (np.abs(Z).max() * np.abs(X).max() * n)
        # partial pivoting's growth on a random matrix rises with n: LAPACK
        # gives 7.6e-16 on this system
        print(f"n={n} {prec}: backward error {be:.2e}")
        assert be < (3e-15 if prec == "fp64" else 1e-14), (prec, be)
        # a second factorisation reuses the choice (no further trials)
        assert ctx.factor(prec) == 0
        assert ctx.gemm_plan_info(prec) == info


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_overlapped_streams_match_one_stream_bitwise(bem, prec, monkeypatch):
    # the three-stream factorisation applies every update in the same order
    # as the one-stream one, so the solutions are bitwise equal; a missing
    # cross-stream dependency shows up as a difference.  The 7-block grid of
    # the trailing updates stretches the overlap.
    rs = np.random.RandomState(13)
    n = 1800
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    B = rs.randn(n, 6) + 1j * rs.randn(n, 6)
    X = {}
    for layout in ("0", "1"):
        monkeypatch.setenv("BEM_LU_STREAMS", layout)
        with bem.Context(0) as c:
            c.set_Z(Z)
            c.gemm_plan(prec, 2, 2, cap=7)
            assert c.factor(prec) == 0
            X[layout], _ = c.solve(B, ir_max=0)
    assert np.array_equal(X["0"], X["1"])
