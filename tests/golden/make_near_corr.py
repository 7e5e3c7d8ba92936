"""Golden vectors of the near-field singular corrections of the fast solvers.

Run in the build container only (the reference tree is not present on GPU
machines):  python tests/golden/make_near_corr.py

Imports python/bem_fmm.py (with python/bem_core.py) from the read-only
reference checkout and records, for small deterministic meshes, the
correction matrices of BEM_FMM_Operator._precompute_singular_corrections
(bem_fmm.py:1193-1324), the Python twin of BemFmmOperator::precompute_corrections
(src/bem_fmm.cu:506-686): for every pair of RWG functions whose halves share a
triangle, C_L = (smooth same-triangle quadrature + analytical static part) -
(the fast solver's point sum of the full kernel over non-coincident points),
and C_K = -(that point sum's K term), for k_ext and k_int:

  * nearcorr_ico0.npz: icosahedron (ref 0), ka = 1.5, m = 1.3116 + 0.02i;
  * nearcorr_hex0.npz: hexagonal prism H/D = 1.4286 (ref 0), ka = 2,
    m = 1.33 + 0.01i.

The matrices are dense N x N in Python's RWG order; the fixtures store the
Python edge keys so tests map them onto the sorted (std::map) order.  Only data
is written; nothing of the reference is copied.
"""
import contextlib
import io
import os
import sys

import numpy as np

REF = "/root/reference/python"
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, OUT)
from make_golden import _hex_prism_ref0, edge_keys  # noqa: E402


def main():
    sys.path.insert(0, REF)
    import bem_core as bc  # noqa: E402  (reference FP64 twin; fixtures only)
    import bem_fmm as bf  # noqa: E402

    cases = [("nearcorr_ico0", *bc.icosphere(1.0, 0), 1.5, 1.3116 + 0.02j),
             ("nearcorr_hex0", *_hex_prism_ref0(1.4286), 2.0, 1.33 + 0.01j)]
    for tag, verts, tris, ka, m in cases:
        rwg = bc.build_rwg(verts, tris)
        k_ext, k_int = complex(ka), ka * m
        with contextlib.redirect_stdout(io.StringIO()):
            op = bf.BEM_FMM_Operator(rwg, verts, tris, k_ext, k_int, 1.0, 1.0 / abs(m), quad_order=7)
        np.savez_compressed(os.path.join(OUT, f"{tag}.npz"), verts=verts, tris=tris, keys=edge_keys(tris),
                            k_ext=k_ext, k_int=k_int, L_ext=op.corr_L_ext, K_ext=op.corr_K_ext,
                            L_int=op.corr_L_int, K_int=op.corr_K_int)
        print(f"{tag}: N={rwg['N']}, nonzero L entries {int((np.abs(op.corr_L_ext) > 0).sum())}")


if __name__ == "__main__":
    main()
