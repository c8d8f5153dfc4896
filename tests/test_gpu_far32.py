"""The mixed form's single-precision far panel pairs (SetFarPrecision(32)):
the matrix against the double-precision assembly, and the mixed solve's
powers against the plain complex128 chain and the Mie series, for the
eps = 4+0.1i level-3 sphere (N=3840) at omega 0.5, 1, 2.  The stated
errors are what FP32 far pairs cost: ~1e-7 relative per far entry."""
import numpy as np
import pytest
import torch

import scuffx

pytestmark = pytest.mark.gpu

PW = scuffx.PlaneWave([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])


@pytest.mark.parametrize("omega", [0.5, 1.0, 2.0])
def test_far32_matrix_and_power(sx, O, omega):
    V, T = scuffx.icosphere(3)
    G = scuffx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_4+0.1i")
    N = G.TotalBFs
    M64 = G.AssembleBEMMatrix(omega)
    b = G.AssembleRHSVector(omega, PW)
    x64 = scuffx.HVector(N, data=b.t.clone())
    M64c = M64.t.clone()
    M64.LUFactorize()
    M64.LUSolve(x64)
    P64 = G.GetPower(omega, x64, b)
    G.SetFarPrecision(32)
    M32 = G.AssembleBEMMatrix(omega)
    G.SetFarPrecision(64)
    dM = float(torch.linalg.norm(M32.t - M64c) / torch.linalg.norm(M64c))
    x32, iters, rel = scuffx.solve_mixed(M32, b, tol=1e-12, maxiter=30)
    P32 = G.GetPower(omega, x32, b)
    perr = max(abs(P32[k] - P64[k]) / abs(P64[k]) for k in ("PAbs", "PExt", "PScat"))
    qs, qe = O.mie(omega, np.sqrt(4 + 0.1j))
    f = np.pi / (2 * O.ZVAC)
    mie = {"PScat": qs * f, "PAbs": (qe - qs) * f}
    merr = max(abs(P32[k] - mie[k]) / abs(mie[k]) for k in mie)
    merr64 = max(abs(P64[k] - mie[k]) / abs(mie[k]) for k in mie)
    print(f"\n  omega={omega}: matrix rel Fro {dM:.2e}, power vs FP64 {perr:.2e} ({iters} refinement steps), "
          f"vs Mie {merr:.2e} (FP64 chain {merr64:.2e})")
    assert 1e-9 < dM <= 1e-6
    assert perr <= 1e-5
    assert abs(merr - merr64) <= 1e-5
