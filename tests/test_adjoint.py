"""Adjoint gradient of a Mueller-matrix loss with respect to (Re m, Im m, ka).

The reference has no derivatives, so the pin is a central finite difference of
the FP64 oracle's whole dense pipeline (assembly -> solve -> far field ->
orientation-averaged Mueller matrix) at the same inputs.  The CPU test checks
the convention the kernels implement, dL/dp = 2 Re[lambda^T (dB/dp - dZ/dp x)]
with Z^T lambda = dL/dx (holomorphic derivative) and Z^T = S Z S for the
PMCHWT block pattern, on a small dense toy against finite differences.
"""
import numpy as np
import pytest


def _pmchwt_like(rs, n, p):
    """Z(p) = [[A, -C], [C, D]] with A, C, D complex symmetric, smooth in p."""
    def sym(M):
        return M + M.T
    A0, A1, C0, C1, D0, D1 = [sym(rs.randn(n, n) + 1j * rs.randn(n, n)) for _ in range(6)]
    A = A0 + 4 * n * np.eye(n) + p * A1 + 0.3 * p * p * A1 @ A1 / n
    C = C0 + np.sin(p) * C1
    D = D0 + 4 * n * np.eye(n) + np.exp(0.2j * p) * D1
    return np.block([[A, -C], [C, D]])


def _toy_loss(x, wts):
    return float(np.sum(wts * np.abs(x) ** 2) + np.real(np.sum(np.conj(x[:2]) * x[2:4])))


def test_adjoint_convention_on_toy_system():
    n = 3
    p0 = 0.37
    wts = np.linspace(0.5, 2.0, 2 * n)

    def setup(p):
        rs = np.random.RandomState(7)
        Z = _pmchwt_like(rs, n, p)
        B = rs.randn(2 * n, 2) + 1j * rs.randn(2 * n, 2) + p * np.exp(1j * np.arange(2 * n))[:, None]
        return Z, B

    def loss(p):
        Z, B = setup(p)
        X = np.linalg.solve(Z, B)
        return sum(_toy_loss(X[:, c], wts) for c in range(2))

    Z, B = setup(p0)
    X = np.linalg.solve(Z, B)
    # holomorphic derivative dL/dx of the real loss (dL = 2 Re sum g dx)
    G = np.empty_like(X)
    for c in range(2):
        x = X[:, c]
        g = wts * np.conj(x)
        g[:2] += 0.5 * np.conj(x[2:4])
        g[2:4] += 0.5 * np.conj(x[:2])
        G[:, c] = g
    # Z^T = S Z S
    S = np.diag(np.r_[np.ones(n), -np.ones(n)])
    assert np.allclose(Z.T, S @ Z @ S)
    Lam = S @ np.linalg.solve(Z, S @ G)
    assert np.allclose(Z.T @ Lam, G)
    h = 1e-6
    dZ = (setup(p0 + h)[0] - setup(p0 - h)[0]) / (2 * h)
    dB = (setup(p0 + h)[1] - setup(p0 - h)[1]) / (2 * h)
    W = Lam @ X.T
    grad = 2 * np.real(np.sum(Lam * dB)) - 2 * np.real(np.sum(dZ * W))
    fd = (loss(p0 + h) - loss(p0 - h)) / (2 * h)
    assert abs(grad - fd) <= 1e-7 * max(1.0, abs(fd))
    # a mistake the kernels could make: dropping the factor 2 or using the conjugate
    wrong = np.real(np.sum(np.conj(Lam) * dB)) - np.real(np.sum(dZ * np.conj(W)))
    assert abs(wrong - fd) > 1e-3 * max(1.0, abs(fd))


def _oracle_loss(oracle, prob, ka, m, Gw, grid, ntheta, mirror):
    _, M, _ = prob.dense_orient(ka, m, *grid, mirror=mirror, ntheta=ntheta)
    return float(np.sum(Gw * M.reshape(16, ntheta)))


CASES = [
    # shape, ref, param, ka, m, orientation grid, mirror
    ("sphere", 1, 1.0, 1.5, 1.33 + 0.01j, (2, 2, 1), False),
    ("hex", 0, 0.7, 2.0, 1.3116 + 0.0j, (2, 1, 2), True),
]


@pytest.mark.gpu
@pytest.mark.parametrize("shape,ref,param,ka,m,grid,mirror", CASES)
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_mueller_grad_vs_oracle_fd(oracle, bem, shape, ref, param, ka, m, grid, mirror, prec):
    ntheta = 19
    rs = np.random.RandomState(42)
    Gw = rs.randn(16, ntheta)
    prob = oracle.Problem.from_shape(shape, ref, param)
    # central differences of the FP64 oracle pipeline
    h = 1e-5
    fd = []
    for dp in [(h, 0, 0), (0, h, 0), (0, 0, h)]:
        up = _oracle_loss(oracle, prob, ka + dp[2], m + dp[0] + 1j * dp[1], Gw, grid, ntheta, mirror)
        dn = _oracle_loss(oracle, prob, ka - dp[2], m - dp[0] - 1j * dp[1], Gw, grid, ntheta, mirror)
        fd.append((up - dn) / (2 * h))
    fd = np.array(fd)
    with bem.Context(0) as ctx:
        ctx.set_mesh(prob.verts, prob.tris, 7)
        ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
        RT, w = bem.orientations(*grid, mirror=mirror)
        kh, e0 = bem.planewave_orient_rhs(RT)
        ctx.rhs_planewave(ka, 1.0, kh, e0)
        assert ctx.factor(prec) == 0
        ctx.solve(None, ir_tol=1e-12, ir_max=30, return_X=False, nrhs=len(kh))
        th = np.linspace(0, np.pi, ntheta)
        M = ctx.orient_mueller(ka, 1.0, RT, w, th, gamma_mirror=mirror)
        L_gpu = float(np.sum(Gw * M.reshape(16, ntheta)))
        L_orc = _oracle_loss(oracle, prob, ka, m, Gw, grid, ntheta, mirror)
        assert abs(L_gpu - L_orc) <= 1e-6 * abs(L_orc)
        grad = ctx.mueller_grad(Gw.reshape(4, 4, ntheta))
        # the state of the forward pass is left as it was
        M2 = ctx.orient_mueller(ka, 1.0, RT, w, th, gamma_mirror=mirror)
        assert np.array_equal(M, M2)
    scale = np.abs(fd).max()
    err = np.abs(grad - fd)
    assert np.all(err <= 1e-5 * scale), (grad, fd, err / scale)


@pytest.mark.gpu
def test_mueller_grad_c4_vs_oracle_fd_fixture(bem):
    """The adjoint workload: icosphere ref 3 (2N = 3840), ka = 3,
    m = 1.33 + 0.005i, lab-frame orientation, 181 angles, central FD (step
    1e-6) of the FP64 oracle from tests/golden/make_fd_c4.py, for the loss
    L = sum_t w_t log M11(theta_t) (w uniform in cos theta) and for a loss
    reaching all 16 elements."""
    from conftest import golden
    g = golden("c4_grad_fd.npz")
    ka, m, nt = float(g["ka"]), complex(g["m"]), int(g["ntheta"])
    # the fixture's own truncation error is far below the tolerance used here
    for key in ("log", "G"):
        a, b = g[f"fd_{key}_h1e6"], g[f"fd_{key}_h1e5"]
        assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
    v, t = bem.make_mesh("sphere", 3, 1.0)
    with bem.Context(0) as ctx:
        ctx.set_mesh(v, t, 7)
        ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
        RT, w = bem.lab_single_orientation()
        kh, e0 = bem.planewave_orient_rhs(RT)
        ctx.rhs_planewave(ka, 1.0, kh, e0)
        assert ctx.factor("fp32") == 0
        ctx.solve(None, ir_tol=1e-12, ir_max=30, return_X=False, nrhs=2)
        M = ctx.orient_mueller(ka, 1.0, RT, w, np.linspace(0, np.pi, nt))
        assert np.abs(M - g["M0"]).max() <= 1e-6 * np.abs(g["M0"][0, 0]).max()
        G_log = np.zeros((4, 4, nt))
        G_log[0, 0] = g["w"] / M[0, 0]
        grad_log = ctx.mueller_grad(G_log)
        grad_G = ctx.mueller_grad(g["G"])
    for grad, fd in ((grad_log, g["fd_log_h1e6"]), (grad_G, g["fd_G_h1e6"])):
        assert np.all(np.abs(grad - fd) <= 1e-5 * np.abs(fd).max()), (grad, fd)


@pytest.mark.gpu
def test_mueller_grad_follows_the_forward_tolerance(oracle, bem):
    """The adjoint solve is refined to the tolerance of the forward solve:
    after a forward solve refined to 1e-8 the gradient stays within 1e-6 of
    the one after a 1e-12 solve (which the FD tests above pin)."""
    ka, m, ntheta = 1.5, 1.33 + 0.01j, 19
    rs = np.random.RandomState(7)
    Gw = rs.randn(4, 4, ntheta)
    prob = oracle.Problem.from_shape("sphere", 1, 1.0)
    grads = {}
    with bem.Context(0) as ctx:
        ctx.set_mesh(prob.verts, prob.tris, 7)
        ctx.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
        RT, w = bem.orientations(2, 2, 1)
        kh, e0 = bem.planewave_orient_rhs(RT)
        ctx.rhs_planewave(ka, 1.0, kh, e0)
        assert ctx.factor("fp32") == 0
        th = np.linspace(0, np.pi, ntheta)
        for tol in (1e-12, 1e-8):
            ctx.solve(None, ir_tol=tol, ir_max=30, return_X=False, nrhs=len(kh))
            ctx.orient_mueller(ka, 1.0, RT, w, th)
            grads[tol] = ctx.mueller_grad(Gw)
    scale = np.abs(grads[1e-12]).max()
    assert np.all(np.abs(grads[1e-8] - grads[1e-12]) <= 1e-6 * scale), grads
