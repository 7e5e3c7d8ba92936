"""Python front end of libbem_amd.so (the MI355X dense PMCHWT library).

Every call goes through the C ABI declared in include/bem_abi.h; there is no
CPU fallback.  If the shared library has not been built, importing this module
raises ``RuntimeError`` -- run ``__graft_entry__.build()`` (or ``make -C
bem-cpp_amd/csrc``) first.

Complex arrays are numpy complex128 (interleaved re/im, the layout the ABI
uses).  Right-hand sides and solutions are passed as (n, nrhs) arrays; the ABI
stores them column-major (one right-hand side per contiguous column), as
solver.h:13-15 of the reference does.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libbem_amd.so")
CLI_PATH = os.path.join(_PKG_ROOT, "bin", "bem_dense")

F64 = 0
F32_3M = 1
ASM_Z64 = 1
ASM_Z32 = 2
ERR_ARG = -1
ERR_STATE = -2
ERR_NOCONV = -5

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


class BemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("residual0", ctypes.c_double),
                ("residual", ctypes.c_double), ("t_solve_ms", ctypes.c_double), ("t_ir_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def _cd(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lib():
    """Loads libbem_amd.so (raises RuntimeError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libbem_amd.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, u = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint
    sig = {
        "bem_create": ([i], vp), "bem_destroy": ([vp], None), "bem_last_error": ([vp], ctypes.c_char_p),
        "bem_timings": ([vp, _dp, i], i), "bem_kernel_stats": ([vp, i, _dp, i], i),
        "bem_autotune": ([vp, i, _dp, i], i), "bem_device_copy": ([vp, d], i),
        "bem_gemm_plan": ([vp, i, i, i, ctypes.c_long], i), "bem_gemm_plan_get": ([vp, i, _dp, i], i),
        "bem_mesh_counts": ([i, i, _ip, _ip], i), "bem_make_mesh": ([i, i, d, _dp, _ip], i),
        "bem_orientations": ([i, i, i, i, i, i, i, _dp, _dp], i),
        "bem_write_json": ([ctypes.c_char_p, _dp, _dp, i, d, d, d, i, i, i, i, _dp], i),
        "bem_write_per_orient": ([ctypes.c_char_p, i, i, i, i, _dp, _dp], i),
        "bem_set_mesh": ([vp, _dp, i, _ip, i, i], i), "bem_num_rwg": ([vp], i),
        "bem_get_rwg": ([vp, _ip, _ip, _dp, _dp, _dp, _dp, _dp], i),
        "bem_assemble": ([vp, d, d, d, d, d, d, u], i), "bem_get_Z": ([vp, _dp, i], i),
        "bem_get_Z_entries": ([vp, _ip, _ip, i, _dp], i),
        "bem_near_corrections": ([vp, d, d, d, d, _ip, _ip, _dp, _dp, _dp, _dp], i),
        "bem_assemble_LK": ([vp, d, d, _dp, _dp], i),
        "bem_factor": ([vp, i], i), "bem_set_Z": ([vp, _dp, i, i], i), "bem_get_pivots": ([vp, _ip], i),
        "bem_get_lu": ([vp, _dp], i), "bem_set_lu": ([vp, _dp, _ip, i], i),
        "bem_solve": ([vp, _dp, i, _dp, d, i, ctypes.POINTER(SolveStats)], i),
        "bem_solve_gmres": ([vp, _dp, i, _dp, d, i, i, i, ctypes.POINTER(SolveStats)], i),
        "bem_block_jacobi": ([vp, d, i, i, _dp], i),
        "bem_block_partition": ([_dp, i, _ip, i, i, d, i, i, _ip, _ip, _ip], i),
        "bem_precond_apply": ([vp, i, _dp, i, _dp], i),
        "bem_rhs_planewave": ([vp, d, d, d, _dp, _dp, i, _dp], i),
        "bem_farfield": ([vp, _dp, _dp, _dp, d, d, d, i, i, i, _dp], i),
        "bem_farfield_noalloc": ([vp, _dp, _dp, _dp, d, d, d, i, i, i, _dp], i),
        "bem_orient_mueller": ([vp, d, d, d, _dp, _dp, i, _dp, i, i, _dp, _dp], i),
        "bem_mueller_grad": ([vp, _dp, _dp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def make_mesh(shape: str, refinements: int, param: float = None):
    """Reference mesh generators (mesh.cpp:11-154) from the library's host code."""
    s = 1 if shape == "hex" else 0
    if param is None:
        param = 1.0
    nv, nt = ctypes.c_int(), ctypes.c_int()
    rc = lib().bem_mesh_counts(s, refinements, ctypes.byref(nv), ctypes.byref(nt))
    if rc:
        raise BemError(rc, "bem_mesh_counts")
    v = np.zeros((nv.value, 3))
    t = np.zeros((nt.value, 3), dtype=np.int32)
    rc = lib().bem_make_mesh(s, refinements, float(param), _p(v), _pi(t))
    if rc:
        raise BemError(rc, "bem_make_mesh")
    return v, t


def orientations(na, nb, ng, beta_sym=1, gamma_sym=1, mirror=False, sort=True):
    n = na * nb * ng
    RT = np.zeros((n, 3, 3))
    w = np.zeros(n)
    got = lib().bem_orientations(na, nb, ng, beta_sym, gamma_sym, int(mirror), int(sort), _p(RT), _p(w))
    if got < 0:
        raise BemError(got, "bem_orientations")
    return RT[:got], w[:got]


def write_json(path, M, theta, ka, n_re, n_im, refinements, orient=(1, 1, 1), timing=(0.0, 0.0, 0.0, 0.0)):
    """The reference's result file (output.cpp:5-64) through the library's
    writer: M is 4 x 4 x ntheta, theta in radians, timing = (assembly, solve,
    far field, total) seconds."""
    M = _cd(M).reshape(16, -1)
    th = _cd(theta)
    if M.shape[1] != th.size:
        raise ValueError("M must be 4 x 4 x len(theta)")
    tm = _cd(timing)
    rc = lib().bem_write_json(os.fsencode(path), _p(M), _p(th), th.size, float(ka), float(n_re), float(n_im),
                              int(refinements), int(orient[0]), int(orient[1]), int(orient[2]), _p(tm))
    if rc:
        raise BemError(rc, f"bem_write_json({path})")


def write_per_orient(path, weights, M_orient, n_alpha, n_beta):
    """The per-orientation Mueller file (main.cpp:640-653): weights (n) and
    M_orient (n x 4 x 4 x ntheta or n x 16 x ntheta), without the 1/|k|^2 of
    the average."""
    w = _cd(weights).ravel()
    Mo = _cd(M_orient).reshape(w.size, 16, -1)
    rc = lib().bem_write_per_orient(os.fsencode(path), w.size, Mo.shape[2], int(n_alpha), int(n_beta), _p(w),
                                    _p(Mo))
    if rc:
        raise BemError(rc, f"bem_write_per_orient({path})")


PRECONDS = {"lu": 0, "block": 1, "none": 2}


def block_partition(verts, tris, quad_order=7, radius_mult=2.0, max_block_rwg=1500, overlap_layers=1):
    """The block-Jacobi / RAS partition of the RWGs (precond.cu:414-605) on the
    host: a list of (own, extended) index arrays (RWG order of Context.rwg)."""
    v = _cd(verts).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    n_ext = ctypes.c_int()
    nb = lib().bem_block_partition(_p(v), len(v), _pi(t), len(t), int(quad_order), float(radius_mult),
                                   int(max_block_rwg), int(overlap_layers), ctypes.byref(n_ext), None, None)
    if nb < 0:
        raise BemError(nb, "bem_block_partition")
    sizes = np.zeros(2 * nb, np.int32)
    rwg = np.zeros(max(n_ext.value, 1), np.int32)
    lib().bem_block_partition(_p(v), len(v), _pi(t), len(t), int(quad_order), float(radius_mult), int(max_block_rwg),
                              int(overlap_layers), ctypes.byref(n_ext), _pi(sizes), _pi(rwg))
    out, o = [], 0
    for b in range(nb):
        own, ext = int(sizes[2 * b]), int(sizes[2 * b + 1])
        out.append((rwg[o:o + own].copy(), rwg[o:o + ext].copy()))
        o += ext
    return out


def auto_precond_params(N, ka):
    """The reference's automatic choice (AutoPrecondParams::compute,
    precond.cu:896-970, without its GPU-memory cap): N <= 500 -> the 2 x 2
    self-interaction blocks, else RAS with one overlap layer and
    block size clamp(sqrt(3e8 / N), 100, 600) (at most 200 for ka > 15 and
    N > 10000).  Returns dict(radius_mult, max_block_rwg, overlap_layers)."""
    if N <= 500:
        return dict(radius_mult=2.0, max_block_rwg=1, overlap_layers=0)
    bs = int(np.sqrt(3e8 / N))
    bs = max(min(bs, 600), 100)
    if ka > 15 and N > 10000:
        bs = min(bs, 200)
    return dict(radius_mult=2.0, max_block_rwg=bs, overlap_layers=1)


class Context:
    """One device context (mesh, Z, factors, right-hand sides, solutions)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        self._h = self._L.bem_create(device)
        if not self._h:
            raise RuntimeError(f"bem_create({device}) failed: no usable GPU")
        self.N = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.bem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what, allow=()):
        if rc < 0 and rc not in allow:
            raise BemError(rc, f"{what}: {self._L.bem_last_error(self._h).decode()}")
        return rc

    # -- mesh -------------------------------------------------------------
    def set_mesh(self, verts, tris, quad_order: int = 7):
        v = _cd(verts)
        t = np.ascontiguousarray(tris, dtype=np.int32)
        self._check(self._L.bem_set_mesh(self._h, _p(v), len(v), _pi(t), len(t), quad_order), "set_mesh")
        self.N = self._L.bem_num_rwg(self._h)
        return self.N

    def rwg(self):
        N = self.N
        tp, tm = np.zeros(N, np.int32), np.zeros(N, np.int32)
        fp, fm = np.zeros((N, 3)), np.zeros((N, 3))
        ln, ap, am = np.zeros(N), np.zeros(N), np.zeros(N)
        self._check(self._L.bem_get_rwg(self._h, _pi(tp), _pi(tm), _p(fp), _p(fm), _p(ln), _p(ap), _p(am)), "get_rwg")
        return dict(N=N, tri_p=tp, tri_m=tm, free_p=fp, free_m=fm, length=ln, area_p=ap, area_m=am)

    # -- assembly ---------------------------------------------------------
    def assemble(self, k_ext: complex, k_int: complex, eta_ext: float, eta_int: float, z64=True, z32=True):
        flags = (ASM_Z64 if z64 else 0) | (ASM_Z32 if z32 else 0)
        k_ext, k_int = complex(k_ext), complex(k_int)
        self._check(self._L.bem_assemble(self._h, k_ext.real, k_ext.imag, k_int.real, k_int.imag, eta_ext, eta_int,
                                         flags), "assemble")

    def get_Z(self):
        n = 2 * self.N
        Z = np.zeros((n, n), np.complex128)
        self._check(self._L.bem_get_Z(self._h, _p(Z), 1), "get_Z")
        return Z

    def get_Z_entries(self, rows, cols):
        """Entries Z[rows[e], cols[e]] of the FP64 matrix (0-based), complex128."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        cl = np.ascontiguousarray(cols, dtype=np.int32)
        if r.shape != cl.shape:
            raise ValueError("rows and cols differ in shape")
        out = np.zeros(r.size, np.complex128)
        self._check(self._L.bem_get_Z_entries(self._h, _pi(r), _pi(cl), r.size, _p(out)), "get_Z_entries")
        return out.reshape(r.shape)

    def near_corrections(self, k_ext: complex, k_int: complex):
        """Near-field singular corrections of the fast operators
        (bem_fmm.cu:506-686) for the mesh: CSR (row_ptr, col_idx) and the values
        of L_ext, K_ext, L_int, K_int (complex128, nnz each)."""
        k_ext, k_int = complex(k_ext), complex(k_int)
        args = (self._h, k_ext.real, k_ext.imag, k_int.real, k_int.imag)
        rp = np.zeros(self.N + 1, np.int32)
        nnz = self._check(self._L.bem_near_corrections(*args, _pi(rp), None, None, None, None, None),
                          "near_corrections")
        ci = np.zeros(nnz, np.int32)
        vals = [np.zeros(nnz, np.complex128) for _ in range(4)]
        self._check(self._L.bem_near_corrections(*args, _pi(rp), _pi(ci), *[_p(v) for v in vals]), "near_corrections")
        return rp, ci, vals

    def assemble_LK(self, k: complex):
        k = complex(k)
        L = np.zeros((self.N, self.N), np.complex128)
        K = np.zeros((self.N, self.N), np.complex128)
        self._check(self._L.bem_assemble_LK(self._h, k.real, k.imag, _p(L), _p(K)), "assemble_LK")
        return L, K

    # -- LU ---------------------------------------------------------------
    def set_Z(self, Z):
        Zc = np.ascontiguousarray(Z, dtype=np.complex128)
        self._check(self._L.bem_set_Z(self._h, _p(Zc), Zc.shape[0], 1), "set_Z")

    def factor(self, precision="fp32"):
        p = F32_3M if precision in ("fp32", F32_3M) else F64
        return self._check(self._L.bem_factor(self._h, p), "factor")

    def get_pivots(self, n):
        ip = np.zeros(n, np.int32)
        self._check(self._L.bem_get_pivots(self._h, _pi(ip)), "get_pivots")
        return ip

    def get_lu(self, n):
        LU = np.zeros((n, n), np.complex128)
        self._check(self._L.bem_get_lu(self._h, _p(LU)), "get_lu")
        return LU

    def set_lu(self, LU, ipiv):
        LUc = np.ascontiguousarray(LU, dtype=np.complex128)
        ip = np.ascontiguousarray(ipiv, dtype=np.int32)
        self._check(self._L.bem_set_lu(self._h, _p(LUc), _pi(ip), LUc.shape[0]), "set_lu")

    def solve(self, B=None, ir_tol=1e-10, ir_max=10, return_X=True, n=None, nrhs=None, allow_noconv=False):
        """Solve Z X = B.  B: (n, nrhs) or (n,) array, or None for the device RHS."""
        st = SolveStats()
        if B is not None:
            B = np.asarray(B, dtype=np.complex128)
            if B.ndim == 1:
                B = B[:, None]
            n, nrhs = B.shape
            Bc = np.ascontiguousarray(B.T)
            bp = _p(Bc)
        else:
            bp = None
            if n is None:
                n = 2 * self.N
        X = np.zeros((nrhs, n), np.complex128) if (return_X and nrhs) else None
        rc = self._L.bem_solve(self._h, bp, nrhs or 0, _p(X) if X is not None else None, ir_tol, ir_max,
                               ctypes.byref(st))
        self._check(rc, "solve", allow=(ERR_NOCONV,) if allow_noconv else ())
        return (X.T.copy() if X is not None else None), st.as_dict()

    def solve_gmres(self, B=None, tol=1e-10, restart=30, max_iter=200, return_X=True, n=None, nrhs=None,
                    allow_noconv=False, precond="lu"):
        """Solve Z X = B by restarted flexible GMRES on the FP64 Z, preconditioned
        by the context's factors (precond="lu"; they may be those of an earlier
        matrix, see bem_solve_gmres), the block-Jacobi preconditioner
        ("block", see block_jacobi) or nothing ("none").  Arguments and results
        as for solve."""
        st = SolveStats()
        if B is not None:
            B = np.asarray(B, dtype=np.complex128)
            if B.ndim == 1:
                B = B[:, None]
            n, nrhs = B.shape
            Bc = np.ascontiguousarray(B.T)
            bp = _p(Bc)
        else:
            bp = None
            if n is None:
                n = 2 * self.N
        X = np.zeros((nrhs, n), np.complex128) if (return_X and nrhs) else None
        rc = self._L.bem_solve_gmres(self._h, bp, nrhs or 0, _p(X) if X is not None else None, tol, restart,
                                     max_iter, PRECONDS[precond], ctypes.byref(st))
        self._check(rc, "solve_gmres", allow=(ERR_NOCONV,) if allow_noconv else ())
        return (X.T.copy() if X is not None else None), st.as_dict()

    def block_jacobi(self, radius_mult=2.0, max_block_rwg=1500, overlap_layers=1):
        """Builds the block-Jacobi / RAS preconditioner from the device's FP64 Z
        (precond.cu:407-895); returns (block count, build ms).  The build time
        grows with the cube of the block size: for large meshes take the
        reference's automatic parameters, auto_precond_params(N, ka)."""
        t = ctypes.c_double()
        nb = self._check(self._L.bem_block_jacobi(self._h, float(radius_mult), int(max_block_rwg),
                                                  int(overlap_layers), ctypes.byref(t)), "block_jacobi")
        return nb, t.value

    def precond_apply(self, R, precond="block"):
        """M^{-1} R for R (n, nrhs) or (n,)."""
        R = np.asarray(R, dtype=np.complex128)
        one = R.ndim == 1
        if one:
            R = R[:, None]
        n, nrhs = R.shape
        Rc = np.ascontiguousarray(R.T)
        Zo = np.zeros((nrhs, n), np.complex128)
        self._check(self._L.bem_precond_apply(self._h, PRECONDS[precond], _p(Rc), nrhs, _p(Zo)), "precond_apply")
        return Zo[0].copy() if one else Zo.T.copy()

    # -- excitation / far field --------------------------------------------
    def rhs_planewave(self, k_ext: complex, eta_ext: float, k_hat, E0, return_B=False):
        kh = _cd(np.atleast_2d(k_hat))
        e0 = _cd(np.atleast_2d(E0))
        nrhs = kh.shape[0]
        B = np.zeros((nrhs, 2 * self.N), np.complex128) if return_B else None
        k_ext = complex(k_ext)
        self._check(self._L.bem_rhs_planewave(self._h, k_ext.real, k_ext.imag, eta_ext, _p(kh), _p(e0), nrhs,
                                              _p(B) if B is not None else None), "rhs_planewave")
        return B.T.copy() if B is not None else None

    def farfield(self, cJ, cM, r_hats, k_ext: complex, eta_ext: float, n_calls: int, n_orient: int,
                 noalloc=False):
        """Far field of n_calls coefficient sets (bem_farfield); noalloc=True
        takes bem_farfield_noalloc (no device allocation, bounded size)."""
        cJ = np.ascontiguousarray(cJ, dtype=np.complex128)
        cM = np.ascontiguousarray(cM, dtype=np.complex128)
        rh = _cd(r_hats).reshape(n_orient, -1, 3)
        ndir = rh.shape[1]
        Fv = np.zeros((n_calls, ndir, 3), np.complex128)
        k_ext = complex(k_ext)
        f = self._L.bem_farfield_noalloc if noalloc else self._L.bem_farfield
        self._check(f(self._h, _p(cJ), _p(cM), _p(rh), k_ext.real, k_ext.imag, eta_ext, n_calls, n_orient, ndir,
                      _p(Fv)), "farfield_noalloc" if noalloc else "farfield")
        return Fv

    def orient_mueller(self, k_ext: complex, eta_ext: float, RT, w, theta, gamma_mirror=False, per_orient=False):
        RTc = _cd(RT).reshape(-1, 3, 3)
        wc = _cd(w)
        th = _cd(theta)
        no, nt = len(wc), len(th)
        M = np.zeros((4, 4, nt))
        Mo = np.zeros((no, 4, 4, nt)) if per_orient else None
        k_ext = complex(k_ext)
        self._check(self._L.bem_orient_mueller(self._h, k_ext.real, k_ext.imag, eta_ext, _p(RTc), _p(wc), no, _p(th),
                                               nt, int(gamma_mirror), _p(M), _p(Mo) if Mo is not None else None),
                    "orient_mueller")
        return (M, Mo) if per_orient else M

    def mueller_grad(self, dL_dM, allow_noconv=False):
        """d L / d (Re m, Im m, ka).  Raises BemError when the adjoint solve
        misses its tolerance, unless allow_noconv (the gradient is then still
        returned)."""
        g = np.zeros(3)
        d = _cd(dL_dM)
        self._check(self._L.bem_mueller_grad(self._h, _p(d), _p(g)), "mueller_grad",
                    allow=(ERR_NOCONV,) if allow_noconv else ())
        return g

    # GEMM configurations (block tile; "k8": K slice of 8 instead of 16)
    GEMM_TILES = ("128x128", "128x64", "64x64", "64x64k8", "32x32")  # FP32
    GEMM_TILES64 = ("128x64", "64x64", "64x32", "64x64k8", "32x32")  # FP64

    def autotune(self, nrhs=128):
        """Times the GEMM tile configurations at this system's shapes and keeps
        the fastest; returns the measured TFLOP/s and the choices."""
        nc = len(self.GEMM_TILES)
        out = np.zeros(4 * nc + 4)
        n = self._L.bem_autotune(self._h, int(nrhs), _p(out), len(out))
        self._check(n, "autotune")
        r = out[:4 * nc].reshape(4, nc)
        return {"fp32_trailing": dict(zip(self.GEMM_TILES, r[0])), "fp32_solve": dict(zip(self.GEMM_TILES, r[1])),
                "fp64_trailing": dict(zip(self.GEMM_TILES64, r[2])), "fp64_solve": dict(zip(self.GEMM_TILES64, r[3])),
                "fp32_choice": (self.GEMM_TILES[int(out[4 * nc])], self.GEMM_TILES[int(out[4 * nc + 1])]),
                "fp64_choice": (self.GEMM_TILES64[int(out[4 * nc + 2])], self.GEMM_TILES64[int(out[4 * nc + 3])])}

    def gemm_plan(self, prec, large, skinny, cap=0):
        """Fixes the GEMM configurations (indices into GEMM_TILES /
        GEMM_TILES64) of one precision and the grid cap of the capped trailing
        updates (0: automatic), replacing the autotuner's choice."""
        p = {"fp32": 0, "fp64": 1}[prec]
        self._check(self._L.bem_gemm_plan(self._h, p, int(large), int(skinny), int(cap)), "gemm_plan")

    # configurations of the FP32 trailing-update kernel (k_cupdate_nt): block
    # tile, "m2": two blocks per CU instead of three, "k32": K slice of 32
    UPDATE_TILES32 = ("64x64", "64x64m2", "128x64", "64x64k32")

    def gemm_plan_info(self, prec):
        """The current GEMM plan of one precision: tile names, grid cap, the
        trailing update's configuration ("update") and the trial factorisation
        time per trailing-update configuration (ms; 0 = not run)."""
        tiles = self.GEMM_TILES if prec == "fp32" else self.GEMM_TILES64
        upd = self.UPDATE_TILES32 if prec == "fp32" else self.GEMM_TILES64
        out = np.zeros(12)
        self._check(self._L.bem_gemm_plan_get(self._h, {"fp32": 0, "fp64": 1}[prec], _p(out), len(out)),
                    "gemm_plan_get")
        return {"large": tiles[int(out[0])], "skinny": tiles[int(out[1])], "panel": tiles[int(out[2])],
                "cap": int(out[3]), "tuned": bool(out[4]), "trialed": bool(out[5]), "update": upd[int(out[11])],
                "trial_factor_ms": {t: float(v) for t, v in zip(upd, out[6:11]) if v > 0}}

    def device_copy(self, nbytes):
        """One plain device-to-device copy kernel of nbytes (a known transfer
        for scaling HBM counters in profiling runs)."""
        self._check(self._L.bem_device_copy(self._h, float(nbytes)), "device_copy")

    KERNEL_CLASSES = ("gemm_fp32_3m", "gemm_fp64", "asm_far", "asm_near", "asm_gather", "tslu", "laswp", "panel",
                      "trsm", "farfield", "rhs", "other", "gemm_fp32_update", "gemm_fp64_update")

    def kernel_stats(self, mode=None):
        """mode 'on' (every class), 'update' (the LU trailing-update GEMMs
        only) and 'off' reset the counters; None returns
        {class: (launches, ms, flops, bytes)}."""
        if mode in ("on", "off", "update"):
            self._L.bem_kernel_stats(self._h, {"on": 1, "off": 0, "update": 3}[mode], None, 0)
            return None
        out = np.zeros(4 * len(self.KERNEL_CLASSES))
        n = self._L.bem_kernel_stats(self._h, 2, _p(out), len(self.KERNEL_CLASSES))
        return {self.KERNEL_CLASSES[i]: tuple(out[4 * i:4 * i + 4]) for i in range(n)}

    def timings(self):
        out = np.zeros(8)
        self._L.bem_timings(self._h, _p(out), 8)
        return dict(assembly_ms=out[0], factor_ms=out[1], solve_ms=out[2], refine_ms=out[3], farfield_ms=out[4],
                    rhs_ms=out[5], adjoint_ms=out[6])


def lab_single_orientation():
    """RT = identity, weight 1: the --single configuration (main.cpp:672-720)."""
    return np.eye(3)[None], np.ones(1)


def planewave_orient_rhs(RT):
    """k_hat = RT z, E_par = RT x, E_perp = RT y per orientation (main.cpp:740-751);
    returns (k_hat, E0) with 2 rows per orientation (par, perp)."""
    RT = np.asarray(RT).reshape(-1, 3, 3)
    kh = np.repeat(RT[:, :, 2], 2, axis=0)
    e0 = np.empty_like(kh)
    e0[0::2] = RT[:, :, 0]
    e0[1::2] = RT[:, :, 1]
    return kh, e0
