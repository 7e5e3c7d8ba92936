"""GPU parity of the near-field singular corrections of the fast operators
(BemFmmOperator::precompute_corrections, bem_fmm.cu:506-686) through the C ABI
(bem_near_corrections) against the FP64 oracle and the reference twin's golden
vectors.  Everything is FP64 on both sides: the CSR pattern must be identical
and the L values agree to 1e-10 of their largest magnitude (measured ~1e-12:
exp and sincos round differently, and the Graglia V is a central difference
with step 1e-6 of a quadrature sum, which turns rounding of 1e-16 into about
1e-11 of V); the K values are rounding only (their exact value is 0 on a flat
triangle) and stay below 1e-14 of it."""
import numpy as np
import pytest

from conftest import golden, python_to_sorted

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(bem):
    c = bem.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("shape,ref,param,ka,m,quad", [
    ("sphere", 1, 1.0, 1.5, 1.3116 + 0.02j, 7),
    ("hex", 1, 1.4286, 2.0, 1.33 + 0.01j, 7),
    ("sphere", 0, 1.0, 3.0, 1.5 + 0.1j, 4),
    ("hex", 0, 0.7, 1.0, 1.3116 + 0j, 13),
])
def test_near_corrections_vs_oracle(oracle, ctx, shape, ref, param, ka, m, quad):
    print(f"{shape} ref {ref} quad {quad}", flush=True)
    p = oracle.Problem.from_shape(shape, ref, param, quad_order=quad)
    ctx.set_mesh(p.verts, p.tris, quad)
    k_ext, k_int = complex(ka), ka * m
    rp, ci, vals = ctx.near_corrections(k_ext, k_int)
    rpo, cio, valso = p.near_corrections(k_ext, k_int)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    scale = max(np.abs(valso[0]).max(), np.abs(valso[2]).max())
    for j in (0, 2):  # L_ext, L_int
        err = np.abs(vals[j] - valso[j]).max()
        print(f"  L[{j}] max|d| / max|L| = {err / scale:.2e}")
        assert err <= 1e-10 * scale
    for j in (1, 3):  # K_ext, K_int: zero up to rounding
        assert np.abs(vals[j]).max() <= 1e-14 * scale


@pytest.mark.parametrize("tag", ["nearcorr_ico0", "nearcorr_hex0"])
def test_near_corrections_vs_reference_golden(ctx, tag):
    g = golden(f"{tag}.npz")
    ctx.set_mesh(g["verts"], g["tris"], 7)
    rp, ci, vals = ctx.near_corrections(complex(g["k_ext"]), complex(g["k_int"]))
    order = python_to_sorted(g["keys"])
    N = len(rp) - 1
    for name, v in (("L_ext", vals[0]), ("L_int", vals[2])):
        D = np.zeros((N, N), np.complex128)
        for i in range(N):
            D[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
        G = g[name][np.ix_(order, order)]
        assert np.array_equal(D != 0, G != 0)
        assert np.abs(D - G).max() <= 1e-9 * np.abs(G).max()
