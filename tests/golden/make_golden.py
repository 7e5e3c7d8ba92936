"""Write golden input/output vectors from the reference's FP64 Python twin.

Run in the build container only (the reference tree is not present on GPU
machines):  python tests/golden/make_golden.py

It imports python/bem_core.py from the read-only reference checkout and
records, for small deterministic inputs, what the reference computes for the
functions whose exact FP64 values no closed form pins:

  * assemble_pmchwt (bem_core.py:1032-1068) on the icosahedron (ref=0, sphere),
    the hex prism ref=0 and (sampled entries) the ref=1 icosphere;
  * compute_mueller_matrix (bem_core.py:1441-1457) at the validate.py case
    (icosphere ref=2, ka=1, m=1.3116, 181 angles; validate.py:9-38);
  * orientation_average_mueller_batched (bem_core.py:1606-1762) on small meshes;
  * compute_rhs_planewave (bem_core.py:1190-1215) and _compute_far_field_vec
    (bem_core.py:1265-1294) for a tilted plane wave;
  * the Graglia integrals (bem_core.py:517-633) at a few observation points.

Python's RWG order is dict-insertion order (bem_core.py:397-421), the C++
reference's is std::map order (rwg.cpp:8-40); each fixture stores the Python
edge keys so tests can map it onto the sorted order.

Only data (inputs and expected outputs) is written; nothing of the reference
is copied.
"""
import os
import sys
import time

import numpy as np

REF = "/root/reference/python"
OUT = os.path.dirname(os.path.abspath(__file__))


def edge_keys(tris):
    keys = {}
    for ti, tri in enumerate(tris):
        for i in range(3):
            a, b = int(tri[i]), int(tri[(i + 1) % 3])
            k = (min(a, b), max(a, b))
            keys.setdefault(k, 0)
            keys[k] += 1
    return np.array([k for k, c in keys.items() if c == 2], dtype=np.int64)


def main():
    sys.path.insert(0, REF)
    import bem_core as bc  # noqa: E402  (reference FP64 twin; fixtures only)

    t0 = time.time()

    def pmchwt(verts, tris, ka, m):
        rwg = bc.build_rwg(verts, tris)
        k_ext = ka
        k_int = ka * m
        Z, _, _ = bc.assemble_pmchwt(rwg, verts, tris, k_ext, k_int, 1.0, 1.0 / abs(m), quad_order=7,
                                     parallel=False)
        return rwg, Z

    # 1. icosahedron (ref 0), real and lossy interior
    v0, t0_ = bc.icosphere(1.0, 0)
    for tag, ka, m in (("ico0_real", 1.0, 1.3116 + 0j), ("ico0_lossy", 2.0, 1.33 + 0.02j)):
        rwg, Z = pmchwt(v0, t0_, ka, m)
        np.savez_compressed(os.path.join(OUT, f"{tag}_Z.npz"), verts=v0, tris=t0_, ka=ka, m=m,
                            keys=edge_keys(t0_), Z=Z)

    # 2. hex prism ref 0 (flat faces: K vanishes for same-face pairs)
    sys.stdout.flush()
    hv, ht = _hex_prism_ref0(1.4286)
    rwg, Z = pmchwt(hv, ht, 1.5, 1.3116 + 0j)
    np.savez_compressed(os.path.join(OUT, "hex0_Z.npz"), verts=hv, tris=ht, ka=1.5, m=1.3116 + 0j,
                        keys=edge_keys(ht), Z=Z)

    # 3. icosphere ref 1: diagonal and 3000 seeded entries of Z
    v1, t1 = bc.icosphere(1.0, 1)
    rwg1, Z1 = pmchwt(v1, t1, 1.0, 1.3116 + 0j)
    rng = np.random.RandomState(42)
    n = Z1.shape[0]
    ii = np.concatenate([np.arange(n), rng.randint(0, n, 3000)])
    jj = np.concatenate([np.arange(n), rng.randint(0, n, 3000)])
    np.savez_compressed(os.path.join(OUT, "ico1_Zsample.npz"), verts=v1, tris=t1, ka=1.0, m=1.3116 + 0j,
                        keys=edge_keys(t1), i=ii, j=jj, z=Z1[ii, jj])

    # 4. validate.py case: icosphere ref 2, ka 1, m 1.3116, 181 angles
    v2, t2 = bc.icosphere(1.0, 2)
    rwg2, Z2 = pmchwt(v2, t2, 1.0, 1.3116 + 0j)
    theta = np.linspace(0, np.pi, 181)
    M = bc.compute_mueller_matrix(rwg2, v2, t2, 1.0, 1.0, theta, Z=Z2, k_hat=np.array([0, 0, 1.0]), sM=-1,
                                  quad_order=7)
    np.savez_compressed(os.path.join(OUT, "c1_mueller.npz"), ka=1.0, m=1.3116 + 0j, ref=2, theta=theta, M=M)

    # 5. orientation averages on small meshes (sums are order independent)
    theta37 = np.linspace(0, np.pi, 37)
    Mo = bc.orientation_average_mueller_batched(rwg1, v1, t1, 1.0, 1.0, theta37, Z=Z1, sM=-1, quad_order=7,
                                                n_alpha=2, n_beta=2, n_gamma=1)
    np.savez_compressed(os.path.join(OUT, "ico1_orient.npz"), ka=1.0, m=1.3116 + 0j, ref=1, orient=(2, 2, 1),
                        theta=theta37, M=Mo)
    hrwg, hZ = pmchwt(hv, ht, 1.5, 1.3116 + 0j)
    Mh = bc.orientation_average_mueller_batched(hrwg, hv, ht, 1.5, 1.0, theta37, Z=hZ, sM=-1, quad_order=7,
                                                n_alpha=3, n_beta=2, n_gamma=2)
    np.savez_compressed(os.path.join(OUT, "hex0_orient.npz"), ka=1.5, m=1.3116 + 0j, ar=1.4286,
                        orient=(3, 2, 2), theta=theta37, M=Mh)

    # 6. RHS and far-field vector for a tilted plane wave on ref 1
    kh = np.array([0.3, -0.4, np.sqrt(1 - 0.25)])
    E0 = np.cross(kh, [1.0, 0.0, 0.0])
    E0 /= np.linalg.norm(E0)
    b = bc.compute_rhs_planewave(rwg1, v1, t1, 1.7, 1.0, E0=E0, k_hat=kh, quad_order=7)
    rs = np.random.RandomState(7)
    N1 = rwg1["N"]
    cJ = rs.randn(N1) + 1j * rs.randn(N1)
    cM = rs.randn(N1) + 1j * rs.randn(N1)
    rh = rs.randn(5, 3)
    rh /= np.linalg.norm(rh, axis=1)[:, None]
    Fv = np.array([bc._compute_far_field_vec(rwg1, v1, t1, cJ, cM, 1.7, 1.0, r, sM=-1, quad_order=7)
                   for r in rh])
    np.savez_compressed(os.path.join(OUT, "ico1_rhs_ff.npz"), verts=v1, tris=t1, keys=edge_keys(t1), k=1.7,
                        k_hat=kh, E0=E0, b=b, cJ=cJ, cM=cM, r_hats=rh, Fv=Fv)

    # 7. Graglia integrals
    tri = np.array([[0.1, -0.2, 0.3], [1.2, 0.1, 0.25], [0.3, 0.9, -0.1]])
    pts = [tri.mean(0), tri.mean(0) + [0, 0, 0.2], [2.0, 1.0, 3.0], 0.7 * tri[0] + 0.2 * tri[1] + 0.1 * tri[2],
           tri[0] + 0.01 * (tri[1] - tri[0]) + [0.0, 0.05, -0.02]]
    P = np.array([bc.potential_integral_triangle(np.array(p), *tri) for p in pts])
    V = np.array([bc.vector_potential_integral_triangle(np.array(p), *tri) for p in pts])
    np.savez_compressed(os.path.join(OUT, "graglia.npz"), tri=tri, pts=np.array(pts), P=P, V=V)
    print(f"golden fixtures written in {time.time() - t0:.1f}s")


def _hex_prism_ref0(ar):
    """The 24-triangle hexagonal prism of mesh.cpp:102-146 (no subdivision)."""
    a = (8.0 * np.pi / (27.0 * ar)) ** (1.0 / 3.0)
    H = ar * a * np.sqrt(3.0)
    v = np.zeros((14, 3))
    v[0] = [0, 0, H / 2]
    v[7] = [0, 0, -H / 2]
    for i in range(6):
        ang = 2.0 * np.pi * i / 6.0
        v[1 + i] = [a * np.cos(ang), a * np.sin(ang), H / 2]
        v[8 + i] = [a * np.cos(ang), a * np.sin(ang), -H / 2]
    t = []
    for i in range(6):
        i1, i2, j1, j2 = 1 + i, 1 + (i + 1) % 6, 8 + i, 8 + (i + 1) % 6
        t += [[0, i1, i2], [7, j2, j1], [i1, j1, j2], [i1, j2, i2]]
    return v, np.array(t, dtype=np.int32)


if __name__ == "__main__":
    main()
