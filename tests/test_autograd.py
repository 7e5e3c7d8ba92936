"""torch.autograd binding of the dense pipeline (bem_amd.autograd): the
backward pass of L = sum(G * M(params)) equals the central difference of the
FP64 oracle pipeline, and re-running the forward at another point between
forward and backward does not corrupt the gradient."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_fd(oracle, prob, params, G, grid, ntheta, h=1e-5):
    m0, ka0 = params[0] + 1j * params[1], params[2]

    def L(ka, m):
        _, M, _ = prob.dense_orient(ka, m, *grid, ntheta=ntheta)
        return float(np.sum(G * M))

    out = []
    for dp in [(h, 0, 0), (0, h, 0), (0, 0, h)]:
        up = L(ka0 + dp[2], m0 + dp[0] + 1j * dp[1])
        dn = L(ka0 - dp[2], m0 - dp[0] - 1j * dp[1])
        out.append((up - dn) / (2 * h))
    return np.array(out)


def test_torch_backward_vs_oracle_fd(oracle):
    import torch
    from bem_amd.autograd import MuellerProblem
    prob = oracle.Problem.from_shape("sphere", 1, 1.0)
    grid, nt = (2, 2, 1), 19
    G = np.random.RandomState(3).randn(4, 4, nt)
    mp = MuellerProblem(prob.verts, prob.tris, orient=grid, ntheta=nt)
    try:
        p0 = [1.33, 0.01, 1.5]
        params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
        M = mp(params)
        L = (torch.from_numpy(G) * M).sum()
        # another forward at a different point before the backward of the first graph
        other = torch.tensor([1.5, 0.0, 2.0], dtype=torch.float64, requires_grad=True)
        (torch.from_numpy(G) * mp(other)).sum().backward()
        L.backward()
        fd = _oracle_fd(oracle, prob, p0, G, grid, nt)
        g = params.grad.numpy()
        assert np.all(np.abs(g - fd) <= 1e-5 * np.abs(fd).max()), (g, fd)
        fd2 = _oracle_fd(oracle, prob, [1.5, 0.0, 2.0], G, grid, nt)
        assert np.all(np.abs(other.grad.numpy() - fd2) <= 1e-5 * np.abs(fd2).max())
    finally:
        mp.close()
