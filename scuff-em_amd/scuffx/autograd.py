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


class AbsorbedPower(torch.autograd.Function):
    """AbsorbedPower.apply(V, G, ns, Omega, IF): the absorbed power of
    G.GetPower (crossed overlaps, OPFT.cc:310-364) for the plane wave IF,
    as a function of the vertex coordinates V (NumVertices x 3, float64
    torch tensor) of surface ns.  Backward, with x = M^{-1} b:
        dP = Re(gamma^T dx) + (dP at fixed x),   dx = M^{-1} (db - dM x),
    so one adjoint solve M^T mu = gamma gives
        dP/dV = explicit + Re(mu^T db/dV) - Re(sum_ij mu_i dM_ij/dV x_j),
    the last term from the vertex-gradient kernels (no dM stored)."""

    @staticmethod
    def forward(ctx, V, G, ns, Omega, IF):
        Vn = V.detach().cpu().numpy().astype(np.float64)
        G.SetVertices(ns, Vn)
        M = G.AssembleBEMMatrix(Omega)
        Mc = M.t.clone()
        b = G.AssembleRHSVector(Omega, IF)
        x = HVector(b.t.shape[0], data=b.t.clone())
        if M.LUFactorize() != 0:
            raise RuntimeError("singular BEM matrix")
        M.LUSolve(x)
        P = G.GetPower(Omega, x, b)["PAbs"]
        ctx.G, ctx.ns, ctx.Omega, ctx.IF, ctx.Vn = G, ns, Omega, IF, Vn
        ctx.M, ctx.Mc, ctx.x = M, Mc, x
        return torch.tensor(P, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gP):
        G, ns, Omega, IF = ctx.G, ctx.ns, ctx.Omega, ctx.IF
        G.SetVertices(ns, ctx.Vn)
        N = ctx.x.t.shape[0]
        gexp, gamma = G.PAbsVertexGradient(ctx.x)
        mu = HVector(N, data=gamma.clone())
        ctx.M.LUSolve(mu, Trans="T")
        gb = G.RHSVertexGradient(Omega, IF, mu)
        gM = G.VertexGradient(Omega, HMatrix(N, N, data=ctx.Mc), u=mu, v=ctx.x)
        g = gexp[ns] + gb[ns].real - gM[ns].real
        return gP.cpu() * torch.from_numpy(g), None, None, None, None


class LogDet(torch.autograd.Function):
    """LogDet.apply(V, G, ns, Omega): Re log det M as a function of the
    vertex coordinates V (NumVertices x 3) of surface ns.  Backward:
    d log det M = Tr(M^{-1} dM) = sum_ij (M^{-1})_ji dM_ij, the inverse
    from the LU factors (LUInvert) as the dense weights of the
    vertex-gradient kernels."""

    @staticmethod
    def forward(ctx, V, G, ns, Omega):
        Vn = V.detach().cpu().numpy().astype(np.float64)
        G.SetVertices(ns, Vn)
        M = G.AssembleBEMMatrix(Omega)
        Mc = M.t.clone()
        if M.LUFactorize() != 0:
            raise RuntimeError("singular BEM matrix")
        ld = M.LogDet()
        ctx.G, ctx.ns, ctx.Omega, ctx.Vn, ctx.M, ctx.Mc = G, ns, Omega, Vn, M, Mc
        return torch.tensor(ld.real, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        G, ns, Omega = ctx.G, ctx.ns, ctx.Omega
        G.SetVertices(ns, ctx.Vn)
        N = ctx.Mc.shape[0]
        Minv = HMatrix(N, N, data=ctx.M.t.clone())
        Minv.ipiv, Minv.factored = ctx.M.ipiv, True
        if Minv.LUInvert() != 0:
            raise RuntimeError("singular BEM matrix")
        gM = G.VertexGradient(Omega, HMatrix(N, N, data=ctx.Mc), C=Minv)
        return g.cpu() * torch.from_numpy(gM[ns].real.copy()), None, None, None
