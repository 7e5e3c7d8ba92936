"""Central finite differences of the FP64 oracle at the adjoint workload.

Writes tests/golden/c4_grad_fd.npz: for the icosphere of 3 refinements
(T = 1280, N = 1920, 2N = 3840), ka = 3, m = 1.33 + 0.005i, the lab-frame
orientation with two polarisations (the --single path, main.cpp:672-720) and
181 angles, two losses and their central differences in (Re m, Im m, ka) with
step 1e-6 (and step 1e-5 as an estimate of the truncation error):
  * L_log = sum_t w_t log M11(theta_t), w uniform in cos(theta):
    w_t = sin(theta_t) pi / (ntheta - 1) (the trapezoid weights of
    d(cos theta), zero at both ends);
  * L_G = sum G * M with G seeded normal (4 x 4 x 181), which reaches all
    16 elements.

The oracle is this repository's FP64 restatement of the reference (oracle/);
the reference itself has no derivatives.  Run in the build container:
    python tests/golden/make_fd_c4.py        (13 oracle pipelines, a few minutes)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

KA, M, NTHETA = 3.0, 1.33 + 0.005j, 181


def pipeline(prob, ka, m):
    info, Mm, _ = prob.dense_single(ka, m, NTHETA)
    assert info == 0
    return Mm


def main():
    prob = orc.Problem.from_shape("sphere", 3, 1.0)
    G = np.random.RandomState(42).randn(4, 4, NTHETA)
    theta = np.pi * np.arange(NTHETA) / (NTHETA - 1)
    w = np.sin(theta) * np.pi / (NTHETA - 1)

    def losses(Mm):
        return float(np.sum(w * np.log(Mm[0, 0]))), float(np.sum(G * Mm))

    t0 = time.time()
    M0 = pipeline(prob, KA, M)
    L0 = losses(M0)
    print(f"L0 = {L0!r}  ({time.time() - t0:.1f} s per pipeline)", flush=True)
    out = {}
    for h, key in [(1e-6, "h1e6"), (1e-5, "h1e5")]:
        fd = np.zeros((2, 3))
        for j, dp in enumerate([(h, 0.0, 0.0), (0.0, h, 0.0), (0.0, 0.0, h)]):
            up = losses(pipeline(prob, KA + dp[2], M + dp[0] + 1j * dp[1]))
            dn = losses(pipeline(prob, KA - dp[2], M - dp[0] - 1j * dp[1]))
            fd[:, j] = (np.array(up) - np.array(dn)) / (2 * h)
            print(key, j, fd[:, j], flush=True)
        out["fd_log_" + key] = fd[0]
        out["fd_G_" + key] = fd[1]
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_grad_fd.npz"), G=G, w=w, M0=M0,
             L0=np.array(L0), ka=KA, m=M, ntheta=NTHETA, **out)
    print("written", out)


if __name__ == "__main__":
    main()
