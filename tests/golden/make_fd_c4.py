"""Central finite differences of the FP64 oracle at the adjoint workload.

Writes tests/golden/c4_grad_fd.npz: for the icosphere of 3 refinements
(T = 1280, N = 1920, 2N = 3840), ka = 3, m = 1.33 + 0.005i, the lab-frame
orientation with two polarisations (the --single path, main.cpp:672-720) and
181 angles, the loss L = sum G * M (G seeded normal, 4 x 4 x 181) and its
central differences in (Re m, Im m, ka) with step 1e-5, plus the step-1e-4
differences as an estimate of the truncation error.

The oracle is this repository's FP64 restatement of the reference (oracle/);
the reference itself has no derivatives.  Run in the build container:
    python tests/golden/make_fd_c4.py        (about 13 oracle pipelines)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

KA, M, NTHETA = 3.0, 1.33 + 0.005j, 181


def loss(prob, G, ka, m):
    info, Mm, _ = prob.dense_single(ka, m, NTHETA)
    assert info == 0
    return float(np.sum(G * Mm)), Mm


def main():
    prob = orc.Problem.from_shape("sphere", 3, 1.0)
    G = np.random.RandomState(42).randn(4, 4, NTHETA)
    t0 = time.time()
    L0, M0 = loss(prob, G, KA, M)
    print(f"L0 = {L0!r}  ({time.time() - t0:.1f} s per pipeline)", flush=True)
    out = {}
    for h, key in [(1e-5, "fd"), (1e-4, "fd_h1e4")]:
        fd = []
        for dp in [(h, 0.0, 0.0), (0.0, h, 0.0), (0.0, 0.0, h)]:
            up, _ = loss(prob, G, KA + dp[2], M + dp[0] + 1j * dp[1])
            dn, _ = loss(prob, G, KA - dp[2], M - dp[0] - 1j * dp[1])
            fd.append((up - dn) / (2 * h))
            print(key, len(fd), fd[-1], flush=True)
        out[key] = np.array(fd)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_grad_fd.npz"), G=G, L0=L0, M0=M0,
             ka=KA, m=M, ntheta=NTHETA, **out)
    print("written", out)


if __name__ == "__main__":
    main()
