"""torch.autograd binding of the dense pipeline.

``MuellerProblem(verts, tris, ...)(params)`` returns the orientation-averaged
Mueller matrix (4 x 4 x ntheta, float64) of the particle for
``params = [Re m, Im m, ka]`` (a float64 tensor), computed on the GPU through
the C ABI: PMCHWT assembly with k_ext = ka, k_int = ka m, eta_ext = 1,
eta_int = 1/|m| (main.cpp:192-196), plane waves of the orientation grid,
FP32 LU on the matrix cores with FP64 refinement (or FP64 LU), far field and
Mueller matrices (main.cpp:758-846).  Its backward pass is the adjoint
gradient of bem_mueller_grad: one adjoint solve with the same factorisation
and a matrix-free contraction of dZ/dp, never a finite difference.

torch is only the autograd plumbing here; every number comes from the HIP
library, which must be built (the import raises otherwise).
"""
import itertools

import numpy as np
import torch

from . import Context, lab_single_orientation, orientations, planewave_orient_rhs

_ids = itertools.count(1)


class MuellerProblem:
    """A mesh, an orientation grid and scattering angles on one device context.

    orient: (n_alpha, n_beta, n_gamma) Gauss-Legendre grid (orient.cpp), or
    None for the single lab-frame orientation of ``--single``.
    """

    def __init__(self, verts, tris, orient=(8, 8, 1), ntheta=181, quad_order=7, precision="fp32", device=0,
                 gamma_mirror=False, ir_tol=1e-12, ir_max=30):
        self.context = Context(device)
        self.context.set_mesh(np.asarray(verts, np.float64), np.asarray(tris, np.int32), quad_order)
        if orient is None:
            self.RT, self.w = lab_single_orientation()
        else:
            self.RT, self.w = orientations(*orient, mirror=gamma_mirror)
        self.k_hat, self.E0 = planewave_orient_rhs(self.RT)
        self.theta = np.linspace(0.0, np.pi, ntheta)
        self.precision = precision
        self.gamma_mirror = gamma_mirror
        self.ir_tol, self.ir_max = ir_tol, ir_max
        self._state = None  # (forward id, ka, m) of the device state

    def mueller(self, ka: float, m: complex) -> np.ndarray:
        """Forward pipeline; leaves its state on the device for the gradient."""
        c = self.context
        c.assemble(complex(ka), ka * m, 1.0, 1.0 / abs(m))
        c.rhs_planewave(ka, 1.0, self.k_hat, self.E0)
        info = c.factor(self.precision)
        if info != 0:
            raise RuntimeError(f"LU: zero pivot at {info}")
        c.solve(None, ir_tol=self.ir_tol, ir_max=self.ir_max, return_X=False, nrhs=len(self.k_hat))
        M = c.orient_mueller(ka, 1.0, self.RT, self.w, self.theta, gamma_mirror=self.gamma_mirror)
        self._state = (next(_ids), float(ka), complex(m))
        return M

    def gradient(self, ka: float, m: complex, dL_dM) -> np.ndarray:
        """d L / d (Re m, Im m, ka) for dL/dM; re-runs the forward pass when
        the device holds another parameter point."""
        if self._state is None or self._state[1:] != (float(ka), complex(m)):
            self.mueller(ka, m)
        return self.context.mueller_grad(np.ascontiguousarray(dL_dM, np.float64))

    def __call__(self, params: torch.Tensor) -> torch.Tensor:
        return DenseMueller.apply(params, self)

    def close(self):
        self.context.close()


class DenseMueller(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, problem):
        p = params.detach().cpu().double().numpy()
        m, ka = complex(p[0], p[1]), float(p[2])
        M = problem.mueller(ka, m)
        ctx.problem = problem
        ctx.point = (ka, m)
        ctx.dtype, ctx.device = params.dtype, params.device
        return torch.from_numpy(M).to(device=params.device, dtype=params.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        ka, m = ctx.point
        g = ctx.problem.gradient(ka, m, grad_out.detach().cpu().double().numpy())
        return torch.from_numpy(g).to(device=ctx.device, dtype=ctx.dtype), None
