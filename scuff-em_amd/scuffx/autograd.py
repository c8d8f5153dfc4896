"""Differentiable solve with respect to a rigid displacement of one
surface, through the adjoint contraction of the BEM matrix derivative
(the reference's GradM path, libscuff.h:500-505, contracted on the GPU).

    x(theta) = M(theta)^{-1} b,   theta = displacement of surface ns

For a real loss L(x), with torch's gradient convention g = dL/dx (complex,
2 dL/d conj(x)):
    dL/dtheta_m = -Re( lambda^H (dM/dtheta_m) x ),   M^H lambda = g.
The right-hand side b is held fixed (it does not move with the surface).
"""
import numpy as np
import torch

from . import HMatrix, HVector


class DisplacedSolve(torch.autograd.Function):
    """DisplacedSolve.apply(theta, G, ns, Omega, b): theta a real 3-vector
    (torch, float64), b a complex128 device tensor of G.TotalBFs; returns
    x = M(theta)^{-1} b."""

    @staticmethod
    def forward(ctx, theta, G, ns, Omega, b):
        d = theta.detach().cpu().numpy().astype(np.float64)
        G.TransformSurface(ns, D=d)
        try:
            M = G.AssembleBEMMatrix(Omega)
            M.LUFactorize()
            x = HVector(b.shape[0], data=b.detach().clone())
            M.LUSolve(x)
        finally:
            G.UnTransform()
        ctx.G, ctx.ns, ctx.Omega, ctx.d, ctx.M = G, ns, Omega, d, M
        ctx.save_for_backward(x.t)
        return x.t.clone()

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        G, ns, Omega, M = ctx.G, ctx.ns, ctx.Omega, ctx.M
        lam = HVector(g.shape[0], data=g.detach().to(torch.complex128).clone())
        M.LUSolve(lam, Trans="C")
        G.TransformSurface(ns, D=ctx.d)
        try:
            s = G.AdjointContraction(ns, Omega, lam, HVector(x.shape[0], data=x))
        finally:
            G.UnTransform()
        grad = torch.tensor(-s.real, dtype=torch.float64, device="cpu")
        return grad, None, None, None, None
