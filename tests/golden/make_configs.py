"""FP64 oracle fixtures at the full benchmark workloads.

Writes tests/golden/cfg_<name>.npz for
  * c2: icosphere 3 refinements (T = 1280, 2N = 3840), ka = 5,
        m = 1.3116 + 0.01i, lab-frame orientation, 2 polarisations
        (the --single path, main.cpp:672-720);
  * c3: hexagonal prism H/D = 1.4286, 3 refinements (T = 1536, 2N = 4608),
        ka = 5, m = 1.3116, Gauss-Legendre orientations 8 x 8 x 1 (128 RHS,
        main.cpp:722-853);
  * c5: hexagonal prism H/D = 0.7, 4 refinements (T = 6144, 2N = 18432),
        ka = 10, m = 1.3116, orientations 8 x 8 x 1 (128 RHS);
all with 7-point quadrature and 181 scattering angles.  Each file holds
  * Zr, Zc, Zv: 2000 seeded (seed 42) entries of the PMCHWT matrix Z and
    Zmax = max|Z| (the scale of the entry-wise absolute tolerance), plus the
    full diagonal Zdiag (the singular self terms);
  * M: the (orientation-averaged) Mueller matrix, 4 x 4 x 181;
  * Xcols, X: the surface-current coefficients (solutions of Z X = B) of the
    listed right-hand-side columns, and Xnorm: the 2-norm of every column.

The oracle is this repository's FP64 restatement of the reference (oracle/),
pinned by the other fixtures in this folder.  Run in the build container:
    python tests/golden/make_configs.py [c2 c3 c5]
(c2 and c3 take about a minute each on 8 cores, c5 about 15 minutes and
~12 GB of memory.)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

NTHETA = 181
NSAMPLE = 2000

CONFIGS = {
    "c2": dict(shape="sphere", ref=3, param=1.0, ka=5.0, m=1.3116 + 0.01j, orient=None, xcols=[0, 1]),
    "c3": dict(shape="hex", ref=3, param=1.4286, ka=5.0, m=1.3116 + 0j, orient=(8, 8, 1), xcols=[0, 1, 126, 127]),
    "c5": dict(shape="hex", ref=4, param=0.7, ka=10.0, m=1.3116 + 0j, orient=(8, 8, 1), xcols=[0, 1]),
}


def make(name):
    c = CONFIGS[name]
    t0 = time.time()
    p = orc.Problem.from_shape(c["shape"], c["ref"], c["param"])
    ka, m = c["ka"], c["m"]
    n = 2 * p.N
    Z = p.assemble_pmchwt(complex(ka), ka * m, 1.0, 1.0 / abs(m))
    rs = np.random.RandomState(42)
    Zr = rs.randint(0, n, NSAMPLE).astype(np.int32)
    Zc = rs.randint(0, n, NSAMPLE).astype(np.int32)
    Zv = Z[Zr, Zc].copy()
    Zmax = float(np.abs(Z).max())
    Zdiag = np.diag(Z).copy()
    del Z
    t1 = time.time()
    if c["orient"] is None:
        info, M, X = p.dense_single(ka, m, NTHETA)
    else:
        info, M, X = p.dense_orient(ka, m, *c["orient"], ntheta=NTHETA)
    assert info == 0, info
    t2 = time.time()
    cols = np.asarray(c["xcols"], dtype=np.int32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"cfg_{name}.npz")
    np.savez_compressed(out, verts=p.verts, tris=p.tris, ka=ka, m=m, orient=np.asarray(c["orient"] or (0, 0, 0)),
                        Zr=Zr, Zc=Zc, Zv=Zv, Zmax=Zmax, Zdiag=Zdiag, M=M, Xcols=cols, X=X[:, cols].copy(),
                        Xnorm=np.linalg.norm(X, axis=0))
    print(f"{name}: 2N={n} assembly {t1 - t0:.1f} s, solve + far field {t2 - t1:.1f} s -> {out}", flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or ["c2", "c3", "c5"]:
        make(name)
