"""GPU parity of the short-wavelength branches of GetPanelPanelInteractions
(PanelPanelInteractions.cc:426-491): near pairs whose larger panel has
|k| R > pi take the 78-point product rule (no common vertex) or the
Taylor-Duffy method (common edge / vertex) instead of desingularization.
Full matrices against the oracle, which restates every branch and is
pinned by the reference's PPI goldens (unit-test-PPIs.cc, incl. the
common-vertex short-wavelength case)."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def sphere(sx, O, level, mat, kind, eps, shift=None):
    V, T = O.icosphere(level)
    G = sx.RWGGeometry()
    Go = O.Geometry()
    for d in ([None] if shift is None else [None, shift]):
        Vi = V if d is None else V + np.asarray(d)
        G.AddMesh(Vi, T, mat)
        Go.add_surface(Vi, T, kind=kind, eps=eps)
    return G, Go


def max_kr(O, level, k):
    V, T = O.icosphere(level)
    c = V[T].mean(axis=1)
    r = np.linalg.norm(V[T] - c[:, None, :], axis=2).max()
    return abs(k) * r


@pytest.mark.parametrize("mat,kind,eps,omega", [
    ("CONST_EPS_100", 1, 100.0, 1.0),        # interior |k| R ~ 3.6, exterior desingularized
    ("CONST_EPS_10", 1, 10.0, 4.0j),         # imaginary frequency, interior |k| R ~ 4.5
    ("PEC", 0, 1.0, 10.0),                   # exterior |k| R ~ 3.6
])
def test_short_wavelength_matrix_matches_oracle(sx, O, mat, kind, eps, omega):
    print(f"sw {mat} omega={omega}", flush=True)
    kin = abs(omega) * np.sqrt(abs(eps)) if kind else abs(omega)
    assert np.pi < max_kr(O, 1, kin) < 20.0
    G, Go = sphere(sx, O, 1, mat, kind, eps)
    M = G.AssembleBEMMatrix(omega).numpy()
    Mo = Go.assemble(omega)
    e = rel_fro(M, Mo)
    print(f"  rel Frobenius {e:.2e}", flush=True)
    assert e <= 1e-12


def test_short_wavelength_two_surfaces(sx, O):
    """near pairs without common vertices across a small gap, both blocks"""
    print("sw two spheres", flush=True)
    G, Go = sphere(sx, O, 1, "CONST_EPS_60", 1, 60.0, shift=(0.0, 0.0, 2.3))
    M = G.AssembleBEMMatrix(0.8).numpy()
    Mo = Go.assemble(0.8)
    e = rel_fro(M, Mo)
    print(f"  rel Frobenius {e:.2e}", flush=True)
    assert e <= 1e-12
    o1, n1 = G.BFOffset(1)
    B = G.AssembleBEMMatrixBlock(1, 1, 0.8).numpy()
    assert np.array_equal(B, M[o1:o1 + n1, o1:o1 + n1])


def test_very_short_wavelength_is_rejected(sx, O):
    """|k| R > 20 on common-edge pairs: the reference's high-k gradient
    kernel reads outside its factorial table (TaylorDuffy.cc:777,812)"""
    G, _ = sphere(sx, O, 1, "PEC", 0, 1.0)
    assert max_kr(O, 1, 60.0) > 20.0
    with pytest.raises(sx.SCUFFXError):
        G.AssembleBEMMatrix(60.0)
