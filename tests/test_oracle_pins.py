"""Pins of the FP64 oracle (CPU): golden vectors written by the reference's own
FP64 twin (tests/golden/make_golden.py), closed forms, invariants and library
routines.  A plausible slip in the restatement (dropped term, wrong sign,
transposed operand, wrong ordering) fails at least one of these."""
import math

import numpy as np
import pytest

from conftest import golden, python_to_sorted


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("order,degree", [(4, 3), (7, 5), (13, 7)])
def test_quadrature_exactness(oracle, order, degree):
    pts, w = oracle.quad_rule(order)
    assert abs(w.sum() - 1.0) < 5e-14  # quadrature.h:7-8: weights sum to one (as written: 1 + 2e-14 for 7 pt)
    l1, l2 = pts[:, 0], pts[:, 1]
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = np.sum(w * l1 ** a * l2 ** b)
            assert abs(got - exact) < 1e-12, (order, a, b, got, exact)
    # one degree higher is not integrated exactly (the rule really has this degree)
    a = degree + 1
    exact = 2.0 * math.factorial(a) / math.factorial(a + 2)
    assert abs(np.sum(w * l1 ** a) - exact) > 1e-8


def test_quadrature_fallback_is_seven(oracle):
    p5, w5 = oracle.quad_rule(5)  # quadrature.h:66-68
    p7, w7 = oracle.quad_rule(7)
    assert np.array_equal(p5, p7) and np.array_equal(w5, w7)


# ---------------------------------------------------------------- meshes
@pytest.mark.parametrize("ref", [0, 1, 2, 3])
def test_icosphere_counts_and_radius(oracle, ref):
    v, t = oracle.make_mesh("sphere", ref)
    T = 20 * 4 ** ref
    assert t.shape == (T, 3) and len(v) == T // 2 + 2  # Euler: V - E + F = 2
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-14)
    edges = {tuple(sorted((a, b))) for tri in t for a, b in zip(tri, np.roll(tri, -1))}
    assert len(edges) == 30 * 4 ** ref


@pytest.mark.parametrize("ar,ref", [(1.4286, 0), (1.4286, 2), (0.7, 1), (0.7, 3)])
def test_hex_prism_volume_and_counts(oracle, ar, ref):
    v, t = oracle.make_mesh("hex", ref, ar)
    T = 24 * 4 ** ref
    assert t.shape == (T, 3) and len(v) == T // 2 + 2
    vol = np.sum(np.einsum("ij,ij->i", v[t[:, 0]], np.cross(v[t[:, 1]], v[t[:, 2]]))) / 6.0
    assert abs(vol - 4.0 * np.pi / 3.0) < 1e-12  # unit equal-volume radius, mesh.cpp:103-109
    H = v[:, 2].max() - v[:, 2].min()
    Dx = 2 * np.max(np.linalg.norm(v[:, :2], axis=1)) * np.cos(np.pi / 6)
    assert abs(H / Dx - ar) < 1e-12  # aspect ratio = H / inscribed diameter


# ---------------------------------------------------------------- RWG
@pytest.mark.parametrize("shape,ref,param", [("sphere", 2, 1.0), ("hex", 1, 0.7)])
def test_rwg_invariants(oracle, shape, ref, param):
    p = oracle.Problem.from_shape(shape, ref, param)
    r = p.rwg()
    T = len(p.tris)
    assert r["N"] == 3 * T // 2
    halves = np.bincount(np.concatenate([r["tri_p"], r["tri_m"]]), minlength=T)
    assert np.all(halves == 3)
    assert np.all(r["tri_p"] < r["tri_m"])  # plus = first triangle seen (rwg.cpp:26-31)
    # std::map order: edge keys strictly increasing
    keys = []
    for n in range(r["N"]):
        a = set(p.tris[r["tri_p"][n]]) & set(p.tris[r["tri_m"][n]])
        keys.append(tuple(sorted(a)))
    assert keys == sorted(keys) and len(set(keys)) == r["N"]
    # total charge of every basis function vanishes: (l/A+)A+ - (l/A-)A- = 0
    v = p.verts
    for n in range(0, r["N"], 7):
        a, b = keys[n]
        assert abs(r["length"][n] - np.linalg.norm(v[a] - v[b])) < 1e-15


# ---------------------------------------------------------------- Graglia
def test_graglia_golden(oracle):
    g = golden("graglia.npz")
    for pt, P, V in zip(g["pts"], g["P"], g["V"]):
        assert abs(oracle.potential_integral(pt, *g["tri"]) - P) < 1e-13
        assert np.abs(oracle.vector_potential(pt, *g["tri"]) - V).max() < 1e-9  # FD step 1e-6 amplifies rounding


def _duffy_1_over_R(r, v0, v1, v2, n=80):
    """Brute-force int_T 1/|r-r'| dS' by splitting T at the projection of r and
    integrating each sub-triangle in Duffy coordinates (regular integrand)."""
    nrm = np.cross(v1 - v0, v2 - v0)
    nh = nrm / np.linalg.norm(nrm)
    c = r - np.dot(r - v0, nh) * nh
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x + 1)
    w = 0.5 * w
    total = 0.0
    for a, b in ((v0, v1), (v1, v2), (v2, v0)):
        sgn = np.sign(np.dot(np.cross(a - c, b - c), nh))
        for ui, wu in zip(x, w):
            for vi, wv in zip(x, w):
                q = c + ui * ((a - c) + vi * (b - a))
                jac = ui * np.linalg.norm(np.cross(a - c, b - a))
                total += sgn * wu * wv * jac / np.linalg.norm(r - q)
    return total


@pytest.mark.parametrize("r", [[0.4, 0.3, 0.0], [0.4, 0.3, 0.25], [1.5, -0.7, 0.4], [0.05, 0.9, 0.0]])
def test_potential_integral_brute_force(oracle, r):
    v0, v1, v2 = np.array([0.0, 0, 0]), np.array([1.0, 0.1, 0]), np.array([0.2, 1.1, 0])
    r = np.array(r)
    ref = _duffy_1_over_R(r, v0, v1, v2)
    assert abs(oracle.potential_integral(r, v0, v1, v2) - ref) < 1e-8 * abs(ref)
    # permutation of the vertices does not change P
    assert abs(oracle.potential_integral(r, v1, v2, v0) - ref) < 1e-8 * abs(ref)


def test_potential_integral_far_limit(oracle):
    v0, v1, v2 = np.array([0.0, 0, 0]), np.array([0.1, 0.0, 0]), np.array([0.0, 0.1, 0])
    r = np.array([30.0, 40.0, 120.0])
    A = 0.005
    d = np.linalg.norm(r - (v0 + v1 + v2) / 3.0)  # monopole about the centroid; next term O((h/d)^2)
    assert abs(oracle.potential_integral(r, v0, v1, v2) - A / d) < 1e-6 * A / d


def test_vector_potential_is_fd_of_quadrature_W(oracle):
    """V = r P - grad W with grad W the central difference (eps 1e-6) of the
    Nq-point quadrature of int |r-r'| (graglia.h:90-112); checked against the
    exact gradient of that quadrature sum."""
    v0, v1, v2 = np.array([0.1, 0.0, 0.2]), np.array([1.0, 0.2, 0.1]), np.array([0.3, 0.8, 0.0])
    r = np.array([0.4, 0.3, 0.5])
    pts, w = oracle.quad_rule(7)
    A = 0.5 * np.linalg.norm(np.cross(v1 - v0, v2 - v0))
    q = (1 - pts.sum(1))[:, None] * v0 + pts[:, :1] * v1 + pts[:, 1:] * v2
    d = r - q
    gradW = A * np.sum(w[:, None] * d / np.linalg.norm(d, axis=1)[:, None], axis=0)
    V = r * oracle.potential_integral(r, v0, v1, v2) - gradW
    assert np.abs(oracle.vector_potential(r, v0, v1, v2, 7) - V).max() < 1e-8


# ---------------------------------------------------------------- assembly
def _perm2(keys):
    order = python_to_sorted(keys)
    N = len(order)
    return np.concatenate([order, order + N])


@pytest.mark.parametrize("tag", ["ico0_real", "hex0"])
def test_Z_matches_reference_twin(oracle, tag):
    g = golden(f"{tag}_Z.npz")
    p = oracle.Problem(g["verts"], g["tris"])
    ka, m = float(g["ka"]), complex(g["m"])
    Z = p.assemble_pmchwt(complex(ka), ka * m, 1.0, 1.0 / abs(m))
    P = _perm2(g["keys"])
    Zr = g["Z"][np.ix_(P, P)]
    assert np.abs(Z - Zr).max() < 1e-9 * np.abs(Zr).max()


def test_Z_lossy_matches_up_to_reference_R0_clamp(oracle):
    """With Im k > 0 the Python twin's R^2 expansion (bem_core.py:888-892) leaves
    R^2 = +-2e-16 at identical quadrature points in a few percent of cases,
    switching those pairs from the clamped value i*Re(k)/4pi to the exact limit
    i*k/4pi.  Entries of RWGs sharing no triangle are unaffected and must agree
    exactly; the others must lie within the span of the two limits."""
    g = golden("ico0_lossy_Z.npz")
    p = oracle.Problem(g["verts"], g["tris"])
    ka, m = float(g["ka"]), complex(g["m"])
    args = (complex(ka), ka * m, 1.0, 1.0 / abs(m))
    Za = p.assemble_pmchwt(*args)
    oracle.set_r0_exact(True)
    try:
        Zb = p.assemble_pmchwt(*args)
    finally:
        oracle.set_r0_exact(False)
    P = _perm2(g["keys"])
    Zr = g["Z"][np.ix_(P, P)]
    r = p.rwg()
    N = r["N"]
    tri_sets = [{r["tri_p"][n], r["tri_m"][n]} for n in range(N)]
    share = np.array([[bool(tri_sets[i] & tri_sets[j]) for j in range(N)] for i in range(N)])
    share2 = np.tile(share, (2, 2))
    scale = np.abs(Zr).max()
    assert np.abs(Za - Zr)[~share2].max() < 1e-9 * scale
    span = np.abs(Zb - Za)
    assert span[share2].max() > 1e-7 * scale  # the two limits really differ here
    # diagonal: all pair terms have one sign, so the reference lies between the limits
    d = np.arange(2 * N)
    assert np.all(np.abs(Zr[d, d] - Za[d, d]) <= 1.0001 * span[d, d] + 1e-9 * scale)
    assert np.abs(Za - Zr).max() < 1.5e-3 * scale


def test_Z_sample_ico1(oracle):
    g = golden("ico1_Zsample.npz")
    p = oracle.Problem(g["verts"], g["tris"])
    Z = p.assemble_pmchwt(1.0 + 0j, 1.3116 + 0j, 1.0, 1.0 / 1.3116)
    P = _perm2(g["keys"])
    inv = np.empty_like(P)
    inv[P] = np.arange(len(P))
    z = Z[inv[g["i"]], inv[g["j"]]]
    assert np.abs(z - g["z"]).max() < 1e-9 * np.abs(g["z"]).max()


def test_Z_structure_hex(oracle):
    """Z^T = S Z S (symmetrised L, K: assembly.cu:494-502, pmchwt.cu:36-51), and
    K vanishes for RWG pairs whose four triangles lie in one flat face."""
    p = oracle.Problem.from_shape("hex", 1, 0.7)
    N = p.N
    m = 1.3116 + 0.01j
    Z = p.assemble_pmchwt(2.0 + 0j, 2.0 * m, 1.0, 1.0 / abs(m))
    S = np.concatenate([np.ones(N), -np.ones(N)])
    assert np.abs(Z.T - S[:, None] * Z * S[None, :]).max() == 0.0
    L, K = p.assemble_LK(2.0 + 0j)
    assert np.abs(L - L.T).max() == 0.0 and np.abs(K - K.T).max() == 0.0
    r = p.rwg()
    v, t = p.verts, p.tris
    nrm = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    face = [tuple(np.round(nrm[i], 9)) + (round(float(np.dot(nrm[i], v[t[i, 0]])), 9),) for i in range(len(t))]
    fid = {f: i for i, f in enumerate(sorted(set(face)))}
    tf = np.array([fid[f] for f in face])
    one = (tf[r["tri_p"]] == tf[r["tri_m"]])
    same = one[:, None] & one[None, :] & (tf[r["tri_p"]][:, None] == tf[r["tri_p"]][None, :])
    assert same.sum() > N  # plenty of same-face pairs
    assert np.abs(K[same]).max() < 1e-13 * np.abs(K).max()
    assert np.abs(K[~same]).max() > 1e-6


def test_Z_continuity_in_k(oracle):
    p = oracle.Problem.from_shape("sphere", 0)
    L1, K1 = p.assemble_LK(1.0 + 0j)
    L2, K2 = p.assemble_LK(1.0 + 1e-6 + 0j)
    assert np.abs(L2 - L1).max() < 1e-5 * np.abs(L1).max()
    assert np.abs(K2 - K1).max() < 1e-5 * np.abs(K1).max()


# ---------------------------------------------------------------- RHS / far field
def test_rhs_and_far_field_golden(oracle):
    g = golden("ico1_rhs_ff.npz")
    p = oracle.Problem(g["verts"], g["tris"])
    P = _perm2(g["keys"])
    k = complex(float(g["k"]))
    b = p.rhs_planewave(k, 1.0, g["E0"], g["k_hat"])
    assert np.abs(b - g["b"][P]).max() < 1e-13 * np.abs(b).max()
    order = python_to_sorted(g["keys"])
    Fv = p.farfield_vec(g["cJ"][order], g["cM"][order], k, 1.0, g["r_hats"])
    assert np.abs(Fv - g["Fv"]).max() < 1e-13 * np.abs(Fv).max()


def test_rhs_closed_form_single_triangle_phase(oracle):
    """k_hat normal to a flat face: the phase is constant over the face, so for
    an RWG inside that face b_J = e^{ik z0} E0 . int f = e^{ik z0} E0 . (l/2)(c+ - c-)."""
    p = oracle.Problem.from_shape("hex", 1, 1.0)
    r = p.rwg()
    v, t = p.verts, p.tris
    ztop = v[:, 2].max()
    top = np.all(np.isclose(v[t][:, :, 2], ztop), axis=1)
    n = next(i for i in range(r["N"]) if top[r["tri_p"][i]] and top[r["tri_m"][i]])
    k = 1.3 + 0j
    E0 = np.array([0.6, 0.8, 0.0])
    b = p.rhs_planewave(k, 1.0, E0, [0.0, 0.0, 1.0])
    cp = v[t[r["tri_p"][n]]].mean(0)
    cm = v[t[r["tri_m"][n]]].mean(0)
    # int over T+ of (l/2A)(r - p+) = (l/2)(c+ - p+); minus half -(l/2)(c- - p-)
    integ = 0.5 * r["length"][n] * ((cp - r["free_p"][n]) - (cm - r["free_m"][n]))
    want = np.exp(1j * k * ztop) * np.dot(E0, integ)
    assert abs(b[n] - want) < 1e-13
    H0 = np.cross([0, 0, 1.0], E0)
    assert abs(b[p.N + n] - np.exp(1j * k * ztop) * np.dot(H0, integ)) < 1e-13


# ---------------------------------------------------------------- Mueller / pipeline
def test_mueller_c1_matches_reference_twin(oracle):
    g = golden("c1_mueller.npz")
    p = oracle.Problem.from_shape("sphere", 2)
    info, M, X = p.dense_single(1.0, 1.3116 + 0j, 181)
    assert info == 0
    assert (np.abs(M - g["M"]) / g["M"][0, 0][None, None, :]).max() < 1e-9


@pytest.mark.parametrize("fix,shape,ref,param,ka,orient",
                         [("ico1_orient.npz", "sphere", 1, 1.0, 1.0, (2, 2, 1)),
                          ("hex0_orient.npz", "hex", 0, 1.4286, 1.5, (3, 2, 2))])
def test_orientation_average_matches_reference_twin(oracle, fix, shape, ref, param, ka, orient):
    g = golden(fix)
    p = oracle.Problem.from_shape(shape, ref, param)
    info, M, X = p.dense_orient(ka, 1.3116 + 0j, *orient, ntheta=len(g["theta"]))
    assert info == 0
    assert (np.abs(M - g["M"]) / g["M"][0, 0][None, None, :]).max() < 1e-9


def test_orient_range_partials_sum_to_full(oracle):
    """Sum of --orient-range partial averages = full run (main.cpp:728-734,
    cluster_bem.py:98-128)."""
    p = oracle.Problem.from_shape("sphere", 1)
    _, Mf, _ = p.dense_orient(1.0, 1.3116 + 0j, 3, 2, 1, ntheta=19)
    _, Ma, _ = p.dense_orient(1.0, 1.3116 + 0j, 3, 2, 1, i0=0, i1=2, ntheta=19)
    _, Mb, _ = p.dense_orient(1.0, 1.3116 + 0j, 3, 2, 1, i0=2, i1=6, ntheta=19)
    assert np.abs(Ma + Mb - Mf).max() < 1e-14 * np.abs(Mf).max()


def test_sphere_mueller_symmetries(oracle):
    p = oracle.Problem.from_shape("sphere", 2)
    info, M, _ = p.dense_single(1.0, 1.3116 + 0.01j, 61)
    s = M[0, 0].max()
    for i, j in [(0, 2), (0, 3), (1, 2), (1, 3), (2, 0), (2, 1), (3, 0), (3, 1)]:
        assert np.abs(M[i, j]).max() < 2e-3 * s
    assert np.abs(M[0, 0] - M[1, 1]).max() < 2e-3 * s
    assert np.abs(M[2, 2] - M[3, 3]).max() < 2e-3 * s
    assert np.abs(M[2, 3] + M[3, 2]).max() < 2e-3 * s
    assert np.all(M[0, 0] >= np.abs(M).max(axis=(0, 1)) - 1e-15)


def test_mie_band_c1(oracle):
    from oracle import mie
    g = golden("c1_mueller.npz")
    th = g["theta"]
    Mm = mie.mueller(1.0, 1.3116, th)
    M = g["M"]
    for i, j in [(0, 0), (0, 1), (2, 2), (2, 3)]:
        assert np.abs(M[i, j] - Mm[i, j]).max() < 0.08 * Mm[0, 0].max()


def test_optical_theorem_c1(oracle):
    """C_ext from the forward amplitude equals C_sca = int M11 dOmega for a
    lossless sphere (energy balance, at the discretisation level)."""
    p = oracle.Problem.from_shape("sphere", 2)
    nt = 181
    _, M, X = p.dense_single(1.0, 1.3116 + 0j, nt)
    th = np.linspace(0, np.pi, nt)
    csca = 2 * np.pi * np.trapezoid(M[0, 0] * np.sin(th), th)
    # forward S: S2(0) from the par solution
    Fv = p.farfield_vec(X[:p.N, 0], X[p.N:, 0], 1.0 + 0j, 1.0, [[0.0, 0.0, 1.0]])
    S2 = -1j * 1.0 * Fv[0, 0]  # e_theta at theta=0 is x
    cext = 4 * np.pi / 1.0 ** 2 * S2.real
    assert abs(csca - cext) < 0.02 * cext


# ---------------------------------------------------------------- Mie itself
def test_mie_optical_theorem_and_rayleigh():
    from oracle import mie
    for x, m in [(1.0, 1.3116), (5.0, 1.3116 + 0.01j), (3.0, 1.33 + 0.005j)]:
        S1, S2 = mie.amplitudes(x, m, [0.0])
        qext, qsca = mie.efficiencies(x, m)
        assert abs(4.0 / x ** 2 * S1[0].real - qext) < 1e-12 * qext  # B&H eq. 4.76
        assert abs(S1[0] - S2[0]) < 1e-12 * abs(S1[0])
        th = np.linspace(0, np.pi, 4001)
        M = mie.mueller(x, m, th, k=x)
        csca = 2 * np.pi * np.trapezoid(M[0, 0] * np.sin(th), th)
        assert abs(csca - qsca * np.pi) < 1e-5 * qsca * np.pi
        if complex(m).imag == 0:
            assert abs(qext - qsca) < 1e-12
    x, m = 0.01, 1.5
    qext, qsca = mie.efficiencies(x, m)
    assert abs(qsca - 8.0 / 3.0 * x ** 4 * abs((m * m - 1) / (m * m + 2)) ** 2) < 1e-4 * qsca


# ---------------------------------------------------------------- orientations
@pytest.mark.parametrize("n", [1, 2, 5, 8, 16])
def test_gauss_legendre_vs_numpy(oracle, n):
    x, w = oracle.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.all(np.diff(x) < 0)  # descending, orient.cpp:44
    assert np.abs(x[::-1] - xr).max() < 1e-14 and np.abs(w[::-1] - wr).max() < 1e-14


@pytest.mark.parametrize("args", [(8, 8, 1, 1, 1, False), (4, 3, 2, 2, 6, True), (5, 4, 3, 1, 1, False)])
def test_orientations(oracle, args):
    RT, w = oracle.orientations(*args)
    assert len(w) == args[0] * args[1] * args[2] and np.all(w > 0)
    for R in RT:
        assert np.abs(R.T @ R - np.eye(3)).max() < 1e-14 and abs(np.linalg.det(R) - 1) < 1e-14
    RTu, wu = oracle.orientations(*args, sort=False)
    # sorting only permutes
    keyf = lambda A: tuple(np.round(A.ravel(), 12))
    assert sorted(map(keyf, RT)) == sorted(map(keyf, RTu))


def test_orientation_weights_sum_to_one(oracle):
    for na, nb, ng in [(8, 8, 1), (3, 5, 4)]:
        _, w = oracle.orientations(na, nb, ng)
        assert abs(w.sum() - 1.0) < 1e-13
    _, w = oracle.orientations(4, 3, 2, 2, 6, True)  # symmetric reduction keeps the full-measure weights
    assert abs(w.sum() - 1.0) < 1e-13


# ---------------------------------------------------------------- LU
@pytest.mark.parametrize("n,nrhs", [(5, 1), (64, 3), (200, 7), (333, 2)])
def test_lu_vs_numpy(oracle, n, nrhs):
    rs = np.random.RandomState(42)
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    B = rs.randn(n, nrhs) + 1j * rs.randn(n, nrhs)
    info, X = oracle.lu_solve(Z, B)
    assert info == 0
    Xr = np.linalg.solve(Z, B)
    assert np.abs(X - Xr).max() < 1e-10 * np.abs(Xr).max()
    be = np.abs(Z @ X - B).max() / (np.abs(Z).max() * np.abs(X).max() * n)
    assert be < 1e-15


def test_lu_singular_returns_info_and_leaves_B(oracle):
    rs = np.random.RandomState(1)
    Z = rs.randn(20, 20) + 0j
    Z[7] = Z[3]
    Z[:, 11] = 0
    B = rs.randn(20, 2) + 0j
    info, X = oracle.lu_solve(Z, B)
    assert info > 0
    assert np.array_equal(X, B)


# ---------------------------------------------------------------- JSON
def test_json_writer_format(oracle, tmp_path):
    import json
    th = np.linspace(0, np.pi, 5)
    M = np.arange(16 * 5, dtype=float).reshape(4, 4, 5) * 1.2345e-3
    path = str(tmp_path / "r.json")
    oracle.write_json(path, M, th, 5.0, 1.3116, 0.01, 3, (8, 8, 1), (1.234, 2.0, 3.0, 6.234))
    txt = open(path).read()
    d = json.loads(txt)
    assert d["ka"] == 5.0 and d["ri"] == [1.3116, 0.01] and d["orient"] == [8, 8, 1]
    assert d["timing"] == {"assembly_s": 1.23, "solve_s": 2.0, "farfield_s": 3.0, "total_s": 6.23}
    assert d["theta"][-1] == 180.0
    assert np.allclose(np.array(d["mueller"]), M, rtol=1e-8)
    assert '"ka": 5.000000,' in txt and "1.23450000e-03" in txt


# ---------------------------------------------------------------- near-field corrections (fast operators)
def _csr_dense(N, rp, ci, v):
    D = np.zeros((N, N), np.complex128)
    for i in range(N):
        D[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return D


@pytest.mark.parametrize("tag", ["nearcorr_ico0", "nearcorr_hex0"])
def test_near_corrections_golden(oracle, tag):
    # BemFmmOperator::precompute_corrections (bem_fmm.cu:506-686) against the
    # reference's Python twin (bem_fmm.py:1193-1324) on ref-0 meshes
    g = golden(f"{tag}.npz")
    p = oracle.Problem(g["verts"], g["tris"])
    rp, ci, vals = p.near_corrections(complex(g["k_ext"]), complex(g["k_int"]))
    order = python_to_sorted(g["keys"])
    N = p.N
    for name, v in zip(["L_ext", "L_int"], [vals[0], vals[2]]):
        D = _csr_dense(N, rp, ci, v)
        G = g[name][np.ix_(order, order)]
        # the pattern: RWG pairs with halves on one triangle (bem_fmm.cu:540-555)
        assert np.array_equal(D != 0, G != 0), name
        # the Graglia V (central difference with step 1e-6) limits the agreement
        assert np.abs(D - G).max() <= 1e-9 * np.abs(G).max(), (name, np.abs(D - G).max())
    # every row holds itself and its neighbours through both triangles, sorted
    for i in range(N):
        row = ci[rp[i]:rp[i + 1]]
        assert i in row and np.all(np.diff(row) > 0)


def test_near_corrections_k_vanishes_on_flat_triangles(oracle):
    # f_m(r_i) . ((r_i - r_j) x f_n(r_j)) = 0 for three vectors in one plane:
    # the K correction of a flat triangle is rounding only
    p = oracle.Problem.from_shape("hex", 1, 1.4286)
    rp, ci, vals = p.near_corrections(2.0 + 0j, 2.0 * (1.33 + 0.01j))
    scale = max(np.abs(vals[0]).max(), np.abs(vals[2]).max())
    assert np.abs(vals[1]).max() <= 1e-14 * scale and np.abs(vals[3]).max() <= 1e-14 * scale
