"""Differentiable visibility operator and least-squares sky reconstruction.

The forward operator A maps a real sky x (one value per pixel) to complex
visibilities

    (A x)_k = sum_p m_kp x_p exp(-2 pi i b_k . s_p)

(the reference's forward, vissgen3d_noblockage.cu:76-102, with full 3D
phase).  For a real loss L of V = A x, torch hands ``backward`` the gradient
g = dL/dRe V + i dL/dIm V, and

    dL/dx_p = sum_k Re(conj(g_k) A_kp) = Re(A^H g)_p,

which is the masked back-projection with unit weights (the reference's
computeC without density compensation, vissgen3d_blockage.cu:167-192): the
hand-written adjoint kernel ``viss_sky_backproject``.

``lsq_reconstruct`` solves  min_x ||A x - V||^2 + lam ||x||^2  with conjugate
gradients on the normal equations  (Re(A^H A) + lam I) x = Re(A^H V)  (one
forward and one adjoint per iteration).
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import torch

from . import ops


@dataclasses.dataclass
class VisibilityOperator:
    sky: ops.Sky
    uvw: torch.Tensor
    xyz_a: Optional[torch.Tensor] = None
    xyz_b: Optional[torch.Tensor] = None
    mask: bool = True
    R_moon: float = 1737.4e3
    orbit_radius: float = 0.0
    fp64: bool = False

    def __post_init__(self):
        dev = self.sky.device
        self.uvw = self.uvw.to(dev, torch.float64).contiguous()
        if self.mask:
            self.xyz_a = self.xyz_a.to(dev, torch.float64).contiguous()
            self.xyz_b = self.xyz_b.to(dev, torch.float64).contiguous()

    @property
    def shape(self):
        return (self.uvw.shape[1], self.sky.P)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return self.sky.forward(x.detach().to(torch.float64), self.uvw, self.xyz_a, self.xyz_b, mask=self.mask,
                                R_moon=self.R_moon, orbit_radius=self.orbit_radius, fp64=self.fp64)

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        """Re(A^H y)."""
        return self.sky.backproject(self.uvw, y.detach().to(torch.complex128), None, self.xyz_a, self.xyz_b,
                                    mask=self.mask, R_moon=self.R_moon, orbit_radius=self.orbit_radius,
                                    fp64=self.fp64)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return Visibilities.apply(x, self)


class Visibilities(torch.autograd.Function):
    """V = A x with the adjoint kernel as backward."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, op: VisibilityOperator):
        ctx.op = op
        ctx.dtype = x.dtype
        return op.forward(x)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return ctx.op.adjoint(g).to(ctx.dtype), None


def lsq_reconstruct(op: VisibilityOperator, V: torch.Tensor, lam: float = 1e-3, iters: int = 50,
                    x0: Optional[torch.Tensor] = None, callback=None) -> torch.Tensor:
    """Conjugate gradients on (Re(A^H A) + lam I) x = Re(A^H V)."""
    b = op.adjoint(V)
    x = torch.zeros_like(b) if x0 is None else x0.to(b).clone()

    def normal(v):
        return op.adjoint(op.forward(v)) + lam * v

    r = b - normal(x) if x0 is not None else b.clone()
    p = r.clone()
    rs = torch.dot(r, r)
    for it in range(iters):
        Ap = normal(p)
        alpha = rs / torch.dot(p, Ap)
        x += alpha * p
        r -= alpha * Ap
        rs_new = torch.dot(r, r)
        if callback is not None:
            callback(it, x, float(rs_new.sqrt()))
        if float(rs_new) == 0.0:
            break
        p = r + (rs_new / rs) * p
        rs = rs_new
    return x
