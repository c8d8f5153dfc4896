"""The frequency sweep entry point (scuffx_batch_sweep): the scattering
caller's loop over frequencies (scuff-scatter.cc:373-567) on a Drude
sphere, eps(w) = 1 - 100 / (w (w + 0.1 i)), over log-spaced frequencies
in [0.05, 5] as in the batch configuration.

Pins: every frequency equals the step-by-step chain (AssembleBEMMatrix,
AssembleRHSVector, LUFactorize, LUSolve, GetPower) to the bit in complex128,
the mixed-precision sweep equals it to 1e-10 in power, and one frequency
equals the oracle's chain to 1e-10."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PW = ([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])
DRUDE = "DRUDE_10_0.1"


def chain(sx, G, w):
    M = G.AssembleBEMMatrix(w)
    KN = G.AssembleRHSVector(w, sx.PlaneWave(*PW))
    RHS = G.AssembleRHSVector(w, sx.PlaneWave(*PW))
    assert M.LUFactorize() == 0
    M.LUSolve(KN)
    P = G.GetPower(w, KN, RHS)
    return np.array([P["PAbs"], P["PExt"], P["PScat"]])


def test_sweep_matches_chain(sx, O):
    print("sweep: Drude sphere L1, 8 frequencies", flush=True)
    V, T = O.icosphere(1)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, DRUDE)
    om = 0.05 * 100.0 ** (np.arange(8) / 7.0)
    t0 = time.perf_counter()
    P64 = sx.BatchSweep(G, om, sx.PlaneWave(*PW), prec=64)
    t1 = time.perf_counter()
    P32 = sx.BatchSweep(G, om, sx.PlaneWave(*PW), prec=32)
    print(f"  complex128 sweep {1e3 * (t1 - t0):.1f} ms, mixed {1e3 * (time.perf_counter() - t1):.1f} ms", flush=True)
    for k, w in enumerate(om):
        ref = chain(sx, G, w)
        assert np.array_equal(P64[k], ref), (w, P64[k], ref)
        assert np.all(np.abs(P32[k] - ref) <= 1e-10 * np.abs(ref).max()), (w, P32[k], ref)
    # the oracle's chain at one frequency
    Go = O.Geometry()
    Go.add_surface(V, T, kind=1, eps=1 - 100.0 / (om[4] * (om[4] + 0.1j)))
    n = Go.num_bfs(0)
    Mo = Go.assemble(om[4])
    R = Go.rhs_planewave(om[4], PW[0], PW[1])
    KN = np.linalg.solve(Mo, R)
    Po = Go.power(om[4], KN, R)
    assert n == G.TotalBFs
    for i, key in enumerate(("PAbs", "PExt")):
        assert abs(P64[4][i] - Po[key]) <= 1e-10 * abs(Po[key]), key


def test_sweep_arguments(sx, O):
    V, T = O.icosphere(0)
    G = sx.RWGGeometry()
    G.AddMesh(V, T)
    with pytest.raises(sx.SCUFFXError):
        sx.BatchSweep(G, [1.0], sx.PlaneWave(*PW), prec=16)
    assert sx.BatchSweep(G, [], sx.PlaneWave(*PW)).shape == (0, 3)


def test_batched_sweep_config3(sx, O):
    """The batch of 64 instances: 32 log-spaced frequencies in [0.05, 5]
    on two Drude spheres of uneven order (level 2, N=960; level 3,
    N=3840), swept by one captured graph per call (batched complex64 LU,
    complex128 refinement).  Every instance's power equals the
    step-by-step complex128 chain (AssembleBEMMatrix, LUFactorize, LUSolve,
    GetPower) within 1e-10 of its largest component."""
    om = 0.05 * 100.0 ** (np.arange(32) / 31.0)
    Gs = []
    for L in (2, 3):
        V, T = O.icosphere(L)
        G = sx.RWGGeometry()
        G.AddMesh(V, T, DRUDE)
        Gs.append(G)
    assert [G.TotalBFs for G in Gs] == [960, 3840]
    print("batched sweep: 2 Drude spheres x 32 frequencies", flush=True)
    t0 = time.perf_counter()
    P, rel, info = sx.SweepBatched(Gs, om, sx.PlaneWave(*PW))
    print(f"  {1e3 * (time.perf_counter() - t0):.1f} ms, solved again {info[0]}, LU block width {info[1]}, "
          f"max residual {rel.max():.2e}", flush=True)
    worst = 0.0
    for g, G in enumerate(Gs):
        for k, w in enumerate(om):
            ref = chain(sx, G, w)
            d = np.abs(P[g, k] - ref).max() / np.abs(ref).max()
            worst = max(worst, d)
            assert d <= 1e-10, (g, w, P[g, k], ref)
    print(f"  worst power difference vs the step-by-step chain {worst:.2e}", flush=True)
    assert rel.max() <= 1e-11


def test_batched_sweep_rejects_short_wavelength(sx, O):
    """a frequency whose near pairs need the short-wavelength kernels is
    reported, not captured"""
    V, T = O.icosphere(1)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_100")
    with pytest.raises(sx.SCUFFXError):
        sx.SweepBatched([G], [1.0], sx.PlaneWave(*PW))
