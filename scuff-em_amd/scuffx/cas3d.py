"""Casimir caller of the BEM hot path (applications/scuff-cas3D).

At each imaginary frequency Xi the caller assembles the diagonal blocks
T_s of every surface and the off-diagonal blocks U_{s,s'} of the BEM
matrix, together with the derivatives dU_{0,s} with respect to a rigid
displacement / rotation of surface 0, stamps them into M, LU-factorizes M
and reads from the factors

  energy integrand  -log|det(M^{-1} M_inf)| / (2 pi)    (M_inf = diag(T_s))
  force / torque    -2 Re Tr(M^{-1} dM) / (2 pi)       (dM = dU^dagger stamped
                                                        into the rows of s >= 1)

(CasimirIntegrand.cc:51-146, :152-184, :298-473).  The imaginary-frequency
integral uses the reference's default 'cliff' rule
(libs/libTriInt/IntegrateCliffFunction.cc:33-320 with the cliff position
0.7/d_min of SumsIntegrals.cc:606-660) or its fixed trapezoid/Simpson
rule (SumsIntegrals.cc:532-587).

With a transformation file the quantities are computed for every entry
(NumTransformations x NumQuantities values per Xi, entry-major): the T
blocks are assembled once per Xi at the untransformed positions, and a U
block is reassembled only when one of its two surfaces has moved under
some entry so far at this Xi (CasimirIntegrand.cc:370-494).

Every block, factorization, determinant and trace runs on the GPU through
the C ABI; this module only orchestrates them and moves device blocks
into place.  As in the reference (CreateSC3Data.cc:100-157), the matrices
of a compact geometry at imaginary frequencies are real: T blocks in packed
symmetric storage, U and dU blocks, M and dM as real matrices, factored by
the real LU (dgetrf semantics).  Real=False keeps everything complex128
(same values: the imaginary parts vanish).  Scope: compact (non-periodic)
geometries at zero temperature.
"""
import math

import numpy as np

from . import HMatrix, SCUFFXError, ReadTransFile, lib, _check, _ptr, _sync_stream

XIMIN_DEFAULT = 0.001          # scuff-cas3D.cc:80
RELTOL_DEFAULT = 1.0e-2        # scuff-cas3D.cc:84
XIMIN_TRAPSIMP = 0.001         # interval of the fixed rule (SumsIntegrals.cc:34-35)
XIMAX_TRAPSIMP = 10.0


class SC3Data:
    """Workspace of the Casimir caller (CreateSC3Data.cc:45-170):
    geometry, requested quantities and the block / matrix storage reused
    at every frequency.

    Quantities are packed in the reference's order: energy, x-, y-,
    z-force, then one torque per axis in TorqueAxes; with a TransFile,
    one such group per transformation in file order."""

    def __init__(self, G, Energy=False, XForce=False, YForce=False, ZForce=False, TorqueAxes=None,
                 XiMin=XIMIN_DEFAULT, AbsTol=0.0, RelTol=RELTOL_DEFAULT, TransFile=None, Real=True):
        self.G = G
        self.Real = bool(Real)
        self.dtype = "float64" if self.Real else "complex128"
        self.GTCs = ReadTransFile(TransFile)          # CreateSC3Data.cc:67-77
        G.CheckGTCList(self.GTCs)
        self.NumTransformations = len(self.GTCs)
        self.NS = G.NumSurfaces
        self.N = G.TotalBFs
        self.N1 = G.BFOffset(0)[1]
        self.TorqueAxes = [] if TorqueAxes is None else [np.asarray(a, dtype=np.float64) for a in TorqueAxes]
        if len(self.TorqueAxes) > 3:
            raise SCUFFXError("at most three torque axes")
        self.Mus = [m for m, on in enumerate((XForce, YForce, ZForce)) if on]
        self.Mus += [3 + t for t in range(len(self.TorqueAxes))]
        self.Energy = bool(Energy)
        if not self.Energy and not self.Mus:
            raise SCUFFXError("no Casimir quantity requested")
        if self.Mus and self.NS < 2:
            raise SCUFFXError("forces and torques need at least two surfaces")
        self.NumQuantities = int(self.Energy) + len(self.Mus)
        self.NTNQ = self.NumTransformations * self.NumQuantities
        self.XiMin, self.AbsTol, self.RelTol = XiMin, AbsTol, RelTol
        self.Mate = G.Mate
        self.M = HMatrix(self.N, self.N, dtype=self.dtype)
        self.dM = HMatrix(self.N, self.N1, dtype=self.dtype) if self.Mus else None
        self.ByXi = []                       # (Xi, EFT) of every evaluation

    def _offsets(self):
        return [self.G.BFOffset(ns) for ns in range(self.NS)]

    def GetCasimirIntegrand(self, Xi, Converged=None):
        """EFT at imaginary frequency Xi for every transformation
        (GetCasimirIntegrand, CasimirIntegrand.cc:298-494, compact
        geometry).  Converged (NTNQ flags, optional): a transformation
        whose quantities have all converged is skipped and reports zeros."""
        G, M = self.G, self.M
        Omega = complex(0.0, Xi)
        NQ = self.NumQuantities
        offs = self._offsets()
        # T blocks, at the untransformed positions; a surface identical to
        # an earlier one (Mate) reuses its block and its LU diagonal
        # (CasimirIntegrand.cc:276-337)
        TBlocks, ld = {}, {}
        logdet_inf = 0.0
        for ns, (o, n) in enumerate(offs):
            nm = self.Mate[ns]
            if nm >= 0:
                TBlocks[ns] = TBlocks[nm]
                if self.Energy:
                    ld[ns] = ld[nm]
                    logdet_inf += ld[ns]
                continue
            T = self._block(ns, ns, Omega, Packed=True)
            TBlocks[ns] = T
            if self.Energy:
                # LU of T for the diagonal of M_inf's factors (:315-337)
                TF = HMatrix(n, n, dtype=self.dtype, data=T.t.clone(), Storage="symmetric" if T.packed else "normal")
                TF.LUFactorize()
                ld[ns] = TF.LogDet().real
                logdet_inf += ld[ns]
        nt_q = len(self.TorqueAxes)
        never_moved = [True] * self.NS
        UBlocks, dU = {}, {}
        EFT = np.zeros(self.NTNQ)
        for nt, GTC in enumerate(self.GTCs):
            if Converged is not None and all(Converged[nt * NQ:(nt + 1) * NQ]):
                continue                                   # CasimirIntegrand.cc:385-418
            G.Transform(GTC)
            for ns, moved in enumerate(G.SurfaceMoved):
                if moved:
                    never_moved[ns] = False
            # U blocks and dU_{0,s}; reused when neither surface has moved
            for ns in range(self.NS):
                for nsp in range(ns + 1, self.NS):
                    if (ns, nsp) in UBlocks and never_moved[ns] and never_moved[nsp]:
                        continue
                    if ns == 0 and self.Mus:
                        r = self._block(0, nsp, Omega, GradM=True, TorqueAxes=self.TorqueAxes if nt_q else None)
                        UBlocks[(ns, nsp)] = r[0]
                        dU[nsp] = list(r[1]) + (list(r[2]) if nt_q else [])
                    else:
                        UBlocks[(ns, nsp)] = self._block(ns, nsp, Omega)
            EFT[nt * NQ:(nt + 1) * NQ] = self._quantities(offs, TBlocks, UBlocks, dU, logdet_inf)
            G.UnTransform()
        self.ByXi.append((Xi, EFT.copy()))
        return EFT

    def _block(self, nsa, nsb, Omega, GradM=False, TorqueAxes=None, Packed=False):
        if self.Real:
            return self.G.AssembleBEMMatrixBlockReal(nsa, nsb, Omega, GradM=GradM, TorqueAxes=TorqueAxes,
                                                     Packed=Packed)
        return self.G.AssembleBEMMatrixBlock(nsa, nsb, Omega, GradM=GradM, TorqueAxes=TorqueAxes)

    def _quantities(self, offs, TBlocks, UBlocks, dU, logdet_inf):
        """stamps the blocks into M, factorizes it and reads the energy,
        force and torque integrands (Factorize, GetLNDetMInvMInf,
        GetTraceMInvdM; CasimirIntegrand.cc:51-184)"""
        M = self.M
        M.t.zero_()
        M.factored = False
        for ns, (o, n) in enumerate(offs):
            T = TBlocks[ns]
            if T.packed:
                # InsertBlock of packed storage: unpacked straight into M
                _sync_stream()
                _check(lib().scuffx_unpack_symmetric(_ptr(T.t), n, _ptr(M.t[o:o + n, o:o + n]), self.N, T.prec))
            else:
                M.t[o:o + n, o:o + n].copy_(T.t)
        for (ns, nsp), U in UBlocks.items():
            (o, n), (op, nq) = offs[ns], offs[nsp]
            M.t[op:op + nq, o:o + n].copy_(U.t)                          # InsertBlock
            M.t[o:o + n, op:op + nq].copy_(U.t.transpose(0, 1).conj())   # InsertBlockAdjoint
        M.LUFactorize()
        out = []
        if self.Energy:
            v = logdet_inf - M.LogDet().real          # GetLNDetMInvMInf, method 1
            out.append(0.0 if not math.isfinite(v) else -v / (2.0 * math.pi))
        for Mu in self.Mus:
            dM = self.dM
            dM.t.zero_()
            for nsp in range(1, self.NS):
                op, nq = offs[nsp]
                D = dU[nsp][Mu]
                dM.t[:, op:op + nq].copy_(D.t.transpose(0, 1).conj())   # InsertBlockAdjoint(dU, off, 0)
            tr = 2.0 * M.TraceSolve(dM).real
            out.append(0.0 if not math.isfinite(tr) else -tr / (2.0 * math.pi))
        return out

    def MinSurfaceDistance(self):
        """smallest vertex-vertex distance between distinct surfaces under
        any transformation (SumsIntegrals.cc:606-631); 1.0 if it is zero"""
        d = math.inf
        for GTC in self.GTCs:
            self.G.Transform(GTC)
            V = [self.G.Vertices(ns) for ns in range(self.NS)]
            for a in range(self.NS):
                for b in range(a + 1, self.NS):
                    diff = V[a][:, None, :] - V[b][None, :, :]
                    d = min(d, float(np.sqrt((diff * diff).sum(-1)).min()))
            self.G.UnTransform()
        return 1.0 if d == 0.0 or not math.isfinite(d) else d

    def GetXiIntegral_Cliff(self):
        """integral over Xi from XiMin to infinity with the cliff rule
        (SumsIntegrals.cc:606-660); returns (EFT, Error)"""
        XiCliff = 0.7 / self.MinSurfaceDistance()
        return IntegrateCliffFunction(lambda xi, skip: self.GetCasimirIntegrand(xi, skip), self.NTNQ,
                                      self.XiMin, 0.0, XiCliff, self.AbsTol, self.RelTol)

    def GetXiIntegral_TrapSimp(self, NumIntervals=50):
        """XiMin * f(XiMin) plus NumIntervals Simpson panels of width
        (10 - 0.001) / NumIntervals from XiMin on, with trapezoid-difference
        errors (SumsIntegrals.cc:532-587)"""
        fL = self.GetCasimirIntegrand(self.XiMin)
        I = fL * self.XiMin
        E = np.zeros_like(I)
        Delta = (XIMAX_TRAPSIMP - XIMIN_TRAPSIMP) / NumIntervals
        Xi = self.XiMin
        for _ in range(NumIntervals):
            fM = self.GetCasimirIntegrand(Xi + 0.5 * Delta)
            fR = self.GetCasimirIntegrand(Xi + Delta)
            Xi += Delta
            simp = (fL + 4.0 * fM + fR) * Delta / 6.0
            trap = (fL + 2.0 * fM + fR) * Delta / 4.0
            I = I + simp
            E = E + np.abs(simp - trap)
            fL = fR
        return I, E


def IntegrateCliffFunction(f, nFun, xMin, xMax, xCliff, AbsTol, RelTol, MaxEvals=100000):
    """Integral of a vector function that is flat up to a 'cliff' and then
    decays (libs/libTriInt/IntegrateCliffFunction.cc:98-320): march from
    xMin in steps Delta, each step estimated by Simpson (kept) against the
    two-panel trapezoid (error); a step whose worst component misses both
    tolerances is halved and retried, one that beats RelTol doubles the
    next.  A component converges once its local mean |f| has stayed below
    RelTol times its running mean for two consecutive steps (after the
    first three); it is then no longer accumulated.  xMax <= xMin means
    infinity.  f(x, skip) returns nFun values; skip flags converged ones.
    Returns (Integral, Error)."""
    if xCliff <= xMin:
        if xMax > xMin:
            xCliff = 0.5 * (xMax - xMin)
        elif xMin > 0.0:
            xCliff = 1.5 * xMin
        else:
            xCliff = 1.0
    if xMax <= xMin:
        xMax = 1.234e89
    conv = np.zeros(nFun, dtype=bool)
    conv_iters = np.zeros(nFun, dtype=np.int64)
    I = np.zeros(nFun)
    E = np.zeros(nFun)
    AbsI = np.zeros(nFun)
    fb = [None, None, None]
    full = [False, False, False]
    min_conv_iters, min_iters = 2, 3
    Delta = 0.5 * xCliff
    iters = evals = 0
    x = xMin
    all_conv = False
    while not all_conv and x < xMax:
        if evals > MaxEvals:
            raise SCUFFXError("cliff integration did not converge within %d evaluations" % MaxEvals)
        if x + Delta > xMax:
            Delta = xMax - x
        for i, xi in enumerate((x, x + 0.5 * Delta, x + Delta)):
            if not full[i]:
                fb[i] = np.asarray(f(xi, conv.copy()), dtype=np.float64)
                evals += 1
        fL, fM, fR = fb
        simp = Delta * (fL + 4.0 * fM + fR) / 6.0
        trap = Delta * (fL + 2.0 * fM + fR) / 4.0
        dI = simp
        dE = np.abs(simp - trap)
        dAbs = Delta * (np.abs(fL) + 4.0 * np.abs(fM) + np.abs(fR)) / 6.0
        act = ~conv
        max_abs = 0.0
        max_rel = 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            for nf in np.nonzero(act)[0]:
                max_abs = float(np.fmax(max_abs, dE[nf]))
                max_rel = float(np.fmax(max_rel, dE[nf] / abs(dI[nf])))   # fmax drops 0/0
        if Delta > 1.0e-8 and max_rel > 2.0 * RelTol and max_abs > AbsTol:
            Delta *= 0.5
            fb[2] = fb[1]
            full = [True, False, True]
            continue
        x += Delta
        all_conv = True
        for nf in np.nonzero(act)[0]:
            I[nf] += dI[nf]
            E[nf] += dE[nf]
            AbsI[nf] += dAbs[nf]
            if iters < min_iters:
                all_conv = False
                continue
            full_slope = AbsI[nf] / (x - xMin)
            local_slope = dAbs[nf] / Delta
            conv_iters[nf] = conv_iters[nf] + 1 if local_slope <= RelTol * full_slope else 0
            if conv_iters[nf] == min_conv_iters:
                conv[nf] = True
            else:
                all_conv = False
        iters += 1
        fb[0] = fb[2]
        full = [True, False, False]
        if max_rel < RelTol:
            Delta *= 2.0
    return I, E
