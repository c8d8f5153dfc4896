"""GPU parity of the callers either side of the system path: the
plane-wave RHS (AssembleRHSVector.cc:190-427) and the power routine
(OPFT.cc:259-364), the whole assemble -> LU -> solve -> power chain
against the oracle (1e-10) and the Mie series, and the mixed-precision
solve against the FP64 one (1e-8 relative power error)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def geoms(sx, O, level, material, kind, eps=1.0):
    V, T = O.icosphere(level)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, material)
    Go = O.Geometry()
    Go.add_surface(V, T, kind=kind, eps=eps)
    return G, Go


PW = ([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])


@pytest.mark.parametrize("material,kind,eps", [("PEC", 0, 1.0), ("CONST_EPS_4+0.1i", 1, 4 + 0.1j)])
@pytest.mark.parametrize("omega", [0.5, 1.0])
def test_rhs_matches_oracle(sx, O, material, kind, eps, omega):
    G, Go = geoms(sx, O, 1, material, kind, eps)
    E0 = [0.3 + 0.1j, 1.0, -0.2j]
    nh = np.array([1.0, 2.0, 2.0]) / 3.0
    r = G.AssembleRHSVector(omega, sx.PlaneWave(E0, nh)).numpy()
    ro = Go.rhs_planewave(omega, E0, nh)
    assert np.linalg.norm(r - ro) <= 1e-13 * np.linalg.norm(ro)


@pytest.mark.parametrize("material,kind,eps", [("PEC", 0, 1.0), ("CONST_EPS_4+0.1i", 1, 4 + 0.1j)])
def test_power_matches_oracle(sx, O, material, kind, eps):
    import torch
    G, Go = geoms(sx, O, 1, material, kind, eps)
    n = G.TotalBFs
    rng = np.random.default_rng(3)
    KN = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    R = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    P = G.GetPower(1.0, sx.HVector(n, data=torch.from_numpy(KN).cuda()),
                   sx.HVector(n, data=torch.from_numpy(R).cuda()))
    Po = Go.power(1.0, KN, R)
    for key in ("PAbs", "PExt"):
        assert abs(P[key] - Po[key]) <= 1e-12 * max(abs(Po[key]), 1e-300) + 1e-300, key


def solve_chain(sx, G, omega):
    M = G.AssembleBEMMatrix(omega)
    KN = G.AssembleRHSVector(omega, sx.PlaneWave(*PW))
    RHS = G.AssembleRHSVector(omega, sx.PlaneWave(*PW))
    assert M.LUFactorize() == 0
    M.LUSolve(KN)
    return G.GetPower(omega, KN, RHS), KN


def test_chain_pec_l2_config1(sx, O):
    """PEC sphere, icosphere level 2, N=480, omega=1: scattered power
    against the oracle chain and the exact Mie value 8.488640e-03
    (unitTests/unit-test-PFT.cc:81-93)."""
    G, Go = geoms(sx, O, 2, "PEC", 0)
    P, KN = solve_chain(sx, G, 1.0)
    Mo = Go.assemble(1.0)
    ro = Go.rhs_planewave(1.0)
    LU, piv, _ = O.zgetrf(Mo)
    KNo = O.zgetrs(LU, piv, ro)[:, 0]
    Po = Go.power(1.0, KNo, ro)
    print(f"  P={P} oracle={Po}", flush=True)
    assert abs(P["PScat"] - Po["PScat"]) <= 1e-10 * abs(Po["PScat"])
    assert abs(P["PExt"] - Po["PExt"]) <= 1e-10 * abs(Po["PExt"])
    # mesh error of the level-2 icosphere (320 flat panels): -3.8%
    assert abs(P["PScat"] - 8.488640e-03) <= 0.05 * 8.488640e-03


def test_chain_dielectric_l2_vs_mie(sx, O):
    G, Go = geoms(sx, O, 2, "CONST_EPS_4+0.1i", 1, 4 + 0.1j)
    P, _ = solve_chain(sx, G, 1.0)
    qs, qe = O.mie(1.0, np.sqrt(4 + 0.1j))
    f = np.pi / (2 * O.ZVAC)
    # mesh error of the level-2 icosphere: PExt -6.3%, PAbs -5.0% (the
    # chain itself is pinned to 1e-5 by the LossySphere_327 regression)
    assert abs(P["PExt"] - qe * f) <= 0.08 * qe * f
    assert abs(P["PAbs"] - (qe - qs) * f) <= 0.07 * (qe - qs) * f


@pytest.mark.parametrize("level,omega", [(2, 1.0), (3, 1.0)])
def test_mixed_precision_power(sx, O, level, omega):
    """complex64 LU + complex128 refinement against the complex128 LU:
    power within 1e-8."""
    G, _ = geoms(sx, O, level, "CONST_EPS_12+0.01i", 1, 12 + 0.01j)
    M = G.AssembleBEMMatrix(omega)
    RHS = G.AssembleRHSVector(omega, sx.PlaneWave(*PW))
    x, iters, rel = sx.solve_mixed(M, RHS, tol=1e-13, maxiter=30)
    Pm = G.GetPower(omega, x, RHS)
    KN = G.AssembleRHSVector(omega, sx.PlaneWave(*PW))
    M.LUFactorize()
    M.LUSolve(KN)
    Pd = G.GetPower(omega, KN, RHS)
    print(f"  level={level} iters={iters} rel={rel:.2e} Pmixed={Pm} Pfp64={Pd}", flush=True)
    for key in ("PAbs", "PExt", "PScat"):
        assert abs(Pm[key] - Pd[key]) <= 1e-8 * abs(Pd[key]), key
