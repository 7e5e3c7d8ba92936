"""ctypes front end of the FP64 oracle (oracle/bem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as the
thing measured or shipped.  The product path (the HIP library and its host
code) never imports this module.

Complex arrays cross the boundary as numpy complex128 (interleaved re/im).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_N", "orc_Nq"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int
        L.orc_get_rwg.argtypes = [ctypes.c_void_p, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
        L.orc_potential_integral.restype = ctypes.c_double
        L.orc_potential_integral.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_integral_R.restype = ctypes.c_double
        L.orc_integral_R.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_vector_potential.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_assemble_LK.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_assemble_pmchwt.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 6 + [_dp]
        L.orc_rhs_planewave.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        _dp, _dp, _dp]
        L.orc_lu_solve.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_lu_solve.restype = ctypes.c_int
        L.orc_farfield_vec.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp, ctypes.c_int, _dp]
        L.orc_amplitude_to_mueller.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_orientations.argtypes = [ctypes.c_int] * 7 + [_dp, _dp]
        L.orc_orientations.restype = ctypes.c_int
        L.orc_dense_single.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, _dp, _dp]
        L.orc_dense_single.restype = ctypes.c_int
        L.orc_dense_orient.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double] + \
            [ctypes.c_int] * 9 + [_dp, _dp]
        L.orc_dense_orient.restype = ctypes.c_int
        L.orc_last_timings.argtypes = [ctypes.c_void_p, _dp]
        L.orc_quad_rule.argtypes = [ctypes.c_int, _dp]
        L.orc_quad_rule.restype = ctypes.c_int
        L.orc_mesh_counts.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_make_mesh.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _ip]
        L.orc_make_mesh.restype = ctypes.c_int
        L.orc_write_json.argtypes = [ctypes.c_char_p, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double] + [ctypes.c_int] * 4 + [ctypes.c_double] * 4
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_r0_exact.argtypes = [ctypes.c_int]
        L.orc_near_corrections.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_ip, _ip] + [_dp] * 4
        L.orc_near_corrections.restype = ctypes.c_int
        _lib = L
    return _lib


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_r0_exact(on: bool) -> None:
    """Switch the R=0 smooth-kernel limit between the reference's i*Re(k)/4pi
    (default) and the exact i*k/4pi."""
    lib().orc_set_r0_exact(int(bool(on)))


def quad_rule(order: int):
    buf = np.zeros(75)
    n = lib().orc_quad_rule(order, _p(buf))
    return np.stack([buf[:n], buf[25:25 + n]], axis=1), buf[50:50 + n].copy()


def make_mesh(shape: str, ref: int, param: float = None):
    """shape 'sphere' (param = radius, default 1) or 'hex' (param = aspect ratio)."""
    s = 1 if shape == "hex" else 0
    if param is None:
        param = 1.0
    nv, nt = ctypes.c_int(), ctypes.c_int()
    lib().orc_mesh_counts(s, ref, ctypes.byref(nv), ctypes.byref(nt))
    verts = np.zeros((nv.value, 3))
    tris = np.zeros((nt.value, 3), dtype=np.int32)
    got = lib().orc_make_mesh(s, ref, float(param), _p(verts), _pi(tris))
    assert got == nv.value, (got, nv.value)
    return verts, tris


def gauss_legendre(n: int):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def orientations(na, nb, ng, beta_sym=1, gamma_sym=1, mirror=False, sort=True):
    n = na * nb * ng
    RT = np.zeros((n, 3, 3))
    w = np.zeros(n)
    lib().orc_orientations(na, nb, ng, beta_sym, gamma_sym, int(mirror), int(sort), _p(RT), _p(w))
    return RT, w


def potential_integral(r, v0, v1, v2) -> float:
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r, v0, v1, v2)]
    return lib().orc_potential_integral(*[_p(x) for x in a])


def integral_R(r, v0, v1, v2, order=7) -> float:
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r, v0, v1, v2)]
    return lib().orc_integral_R(*[_p(x) for x in a], order)


def vector_potential(r, v0, v1, v2, order=7):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r, v0, v1, v2)]
    out = np.zeros(3)
    lib().orc_vector_potential(*[_p(x) for x in a], order, _p(out))
    return out


def amplitude_to_mueller(S1, S2, S3, S4):
    S = [np.ascontiguousarray(s, dtype=np.complex128) for s in (S1, S2, S3, S4)]
    nt = len(S[0])
    M = np.zeros((4, 4, nt))
    lib().orc_amplitude_to_mueller(*[_p(s) for s in S], nt, _p(M))
    return M


def lu_solve(Z, B):
    """Solve Z X = B (Z row-major n x n, B n x nrhs). Returns (info, X)."""
    Zc = np.array(Z, dtype=np.complex128, order="C", copy=True)
    n = Zc.shape[0]
    B = np.asarray(B, dtype=np.complex128)
    if B.ndim == 1:
        B = B[:, None]
    Bc = np.array(B.T, order="C", copy=True)  # column-major n x nrhs
    info = lib().orc_lu_solve(_p(Zc), n, _p(Bc), Bc.shape[0])
    return info, Bc.T.copy()


class Problem:
    """Mesh + RWG + quadrature held by the oracle."""

    def __init__(self, verts, tris, quad_order: int = 7):
        self.verts = np.ascontiguousarray(verts, dtype=np.float64)
        self.tris = np.ascontiguousarray(tris, dtype=np.int32)
        self._h = lib().orc_create(_p(self.verts), len(self.verts), _pi(self.tris), len(self.tris), quad_order)
        self.N = lib().orc_N(self._h)
        self.Nq = lib().orc_Nq(self._h)

    @classmethod
    def from_shape(cls, shape: str, ref: int, param: float = None, quad_order: int = 7):
        v, t = make_mesh(shape, ref, param)
        return cls(v, t, quad_order)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rwg(self):
        N = self.N
        tp, tm = np.zeros(N, np.int32), np.zeros(N, np.int32)
        fp, fm = np.zeros((N, 3)), np.zeros((N, 3))
        ln, ap, am = np.zeros(N), np.zeros(N), np.zeros(N)
        lib().orc_get_rwg(self._h, _pi(tp), _pi(tm), _p(fp), _p(fm), _p(ln), _p(ap), _p(am))
        return dict(N=N, tri_p=tp, tri_m=tm, free_p=fp, free_m=fm, length=ln, area_p=ap, area_m=am)

    def assemble_LK(self, k: complex):
        L = np.zeros((self.N, self.N), np.complex128)
        K = np.zeros((self.N, self.N), np.complex128)
        lib().orc_assemble_LK(self._h, k.real, k.imag, _p(L), _p(K))
        return L, K

    def assemble_pmchwt(self, k_ext: complex, k_int: complex, eta_ext: float, eta_int: float):
        n = 2 * self.N
        Z = np.zeros((n, n), np.complex128)
        lib().orc_assemble_pmchwt(self._h, k_ext.real, k_ext.imag, k_int.real, k_int.imag, eta_ext, eta_int, _p(Z))
        return Z

    def rhs_planewave(self, k_ext: complex, eta_ext: float, E0, k_hat):
        b = np.zeros(2 * self.N, np.complex128)
        e = np.ascontiguousarray(E0, dtype=np.float64)
        kh = np.ascontiguousarray(k_hat, dtype=np.float64)
        lib().orc_rhs_planewave(self._h, k_ext.real, k_ext.imag, eta_ext, _p(e), _p(kh), _p(b))
        return b

    def farfield_vec(self, cJ, cM, k_ext: complex, eta_ext: float, r_hats):
        cJ = np.ascontiguousarray(cJ, dtype=np.complex128)
        cM = np.ascontiguousarray(cM, dtype=np.complex128)
        rh = np.ascontiguousarray(r_hats, dtype=np.float64).reshape(-1, 3)
        Fv = np.zeros((len(rh), 3), np.complex128)
        lib().orc_farfield_vec(self._h, _p(cJ), _p(cM), k_ext.real, k_ext.imag, eta_ext, _p(rh), len(rh), _p(Fv))
        return Fv

    def dense_single(self, ka: float, m: complex, ntheta: int = 181):
        M = np.zeros((4, 4, ntheta))
        X = np.zeros((2, 2 * self.N), np.complex128)
        info = lib().orc_dense_single(self._h, ka, m.real, m.imag, ntheta, _p(M), _p(X))
        return info, M, X.T.copy()

    def dense_orient(self, ka: float, m: complex, na=8, nb=8, ng=1, beta_sym=1, gamma_sym=1, mirror=False,
                     i0=-1, i1=-1, ntheta: int = 181):
        M = np.zeros((4, 4, ntheta))
        n_all = na * nb * ng
        nr = n_all if not (i0 >= 0 and i1 > i0) else min(i1, n_all) - min(i0, n_all)
        X = np.zeros((2 * nr, 2 * self.N), np.complex128)
        info = lib().orc_dense_orient(self._h, ka, m.real, m.imag, na, nb, ng, beta_sym, gamma_sym, int(mirror),
                                      i0, i1, ntheta, _p(M), _p(X))
        return info, M, X.T.copy()

    def near_corrections(self, k_ext: complex, k_int: complex):
        """Near-field singular corrections of the fast operators
        (bem_fmm.cu:506-686): CSR (row_ptr, col_idx) and the values of
        L_ext, K_ext, L_int, K_int (complex128, nnz each)."""
        k_ext, k_int = complex(k_ext), complex(k_int)
        args = (self._h, k_ext.real, k_ext.imag, k_int.real, k_int.imag)
        rp = np.zeros(self.N + 1, np.int32)
        nnz = lib().orc_near_corrections(*args, _pi(rp), None, None, None, None, None)
        ci = np.zeros(nnz, np.int32)
        vals = [np.zeros(nnz, np.complex128) for _ in range(4)]
        lib().orc_near_corrections(*args, _pi(rp), _pi(ci), *[_p(v) for v in vals])
        return rp, ci, vals

    def last_timings(self):
        out = np.zeros(3)
        lib().orc_last_timings(self._h, _p(out))
        return dict(assembly_s=out[0], solve_s=out[1], farfield_s=out[2])


def write_json(path, M, theta, ka, n_re, n_im, ref, orient=(8, 8, 1), timing=(0, 0, 0, 0)):
    M = np.ascontiguousarray(M, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    return lib().orc_write_json(path.encode(), _p(M), _p(th), len(th), ka, n_re, n_im, ref, *orient, *timing)
