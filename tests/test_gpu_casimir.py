"""GPU parity of the Casimir caller (applications/scuff-cas3D): diagonal T
blocks, U blocks with their displacement / rotation derivatives, the LU of
the stamped matrix, log det and Tr(M^{-1} dM) (CasimirIntegrand.cc:51-184,
:298-473), and the imaginary-frequency integral (SumsIntegrals.cc:606-660).

Pins:
  * integrands at one Xi against the oracle's full matrix and GradM/dMdT
    blocks, with numpy's slogdet / solve as the linear algebra (1e-10);
  * the z-force between two R = 1 spheres of Sphere_327.msh at centre
    distances 3.0 and 4.9 (PEC) and 3.0 (eps = 10, which reaches the
    short-wavelength panel-pair branches inside the spheres), against the
    reference's own published outputs PECSpheres_327.out /
    E10Spheres_327.out (doc/docs/tests/CasimirSpheres), within the two
    quadratures' combined error estimates."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, read_msh

pytestmark = pytest.mark.gpu

MESH = os.path.join(GOLDEN, "Sphere_327.msh")


def spheres_geo(tmp_path, sep, material=None):
    mat = f"\tMATERIAL {material}\n" if material else ""
    geo = tmp_path / f"spheres_{sep}.scuffgeo"
    geo.write_text(f"OBJECT UpperSphere\n\tMESHFILE {MESH}\n\tDISPLACED 0 0 {sep}\n{mat}ENDOBJECT\n\n"
                   f"OBJECT LowerSphere\n\tMESHFILE {MESH}\n{mat}ENDOBJECT\n")
    return str(geo)


def oracle_pair(O, sep, kind, eps):
    V, T = read_msh(MESH)
    Go = O.Geometry()
    Go.add_surface(V + np.array([0.0, 0.0, sep]), T, kind=kind, eps=eps)
    Go.add_surface(V, T, kind=kind, eps=eps)
    Go.set_mate(1, 0)          # one mesh file, one material: the reference's Mate
    return Go


def golden(name):
    rows = [ln.split() for ln in open(os.path.join(GOLDEN, name)) if ln.strip() and not ln.startswith("#")]
    return {r[0]: (float(r[1]), float(r[2])) for r in rows}


@pytest.mark.parametrize("kind,material,eps", [(0, None, 1.0), (1, "CONST_EPS_10", 10.0)])
def test_integrand_matches_oracle(sx, O, tmp_path, kind, material, eps):
    from scuffx.cas3d import SC3Data
    sep, xi = 3.0, 0.7
    print(f"casimir integrand kind={kind} xi={xi}", flush=True)
    G = sx.RWGGeometry(spheres_geo(tmp_path, sep, material))
    axes = [np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])]
    S = SC3Data(G, Energy=True, XForce=True, ZForce=True, TorqueAxes=axes)
    EFT = S.GetCasimirIntegrand(xi)

    Go = oracle_pair(O, sep, kind, eps)
    n0 = Go.num_bfs(0)
    M = Go.assemble(1j * xi)
    _, GB, TB = Go.assemble_block_grad(1j * xi, 0, 1, gammas=[sx.CreateGammaMatrix(a) for a in axes])
    ld = [np.linalg.slogdet(M[o:o + n, o:o + n])[1] for o, n in ((0, n0), (n0, M.shape[0] - n0))]
    ldm = np.linalg.slogdet(M)[1]
    energy = -(sum(ld) - ldm) / (2 * np.pi)
    # the energy is a difference of log determinants of size ~N: its
    # rounding error scales with their magnitude, not with the difference
    assert abs(EFT[0] - energy) <= 1e-14 * (abs(ldm) + sum(abs(x) for x in ld)) / (2 * np.pi)

    def force(dU):
        dM = np.zeros((M.shape[0], n0), dtype=complex)
        dM[n0:, :] = dU.conj().T
        X = np.linalg.solve(M, dM)
        return -2.0 * np.trace(X[:n0, :]).real / (2 * np.pi)

    want = np.array([force(GB[0]), force(GB[2]), force(TB[0]), force(TB[1])])
    scale = np.abs(want).max()
    assert np.all(np.abs(EFT[1:] - want) <= 1e-10 * scale), (EFT, want)


@pytest.mark.parametrize("name,material,sep", [
    ("PECSpheres_327.out", None, "3.0"),
    ("PECSpheres_327.out", None, "4.9"),
    ("E10Spheres_327.out", "CONST_EPS_10", "3.0"),        # reaches |k| R > pi inside the spheres
])
def test_zforce_matches_reference_output(sx, tmp_path, name, material, sep):
    from scuffx.cas3d import SC3Data
    print(f"casimir zforce {name} sep={sep}", flush=True)
    G = sx.RWGGeometry(spheres_geo(tmp_path, sep, material))
    S = SC3Data(G, ZForce=True)
    I, E = S.GetXiIntegral_Cliff()
    ref, ref_err = golden(name)[sep]
    print(f"  F={I[0]:.7e} (+- {E[0]:.1e})  reference {ref:.7e} (+- {ref_err:.1e})  "
          f"evaluations {len(S.ByXi)}", flush=True)
    # the two Xi integrals (same integrand, different quadrature runs)
    # differ by no more than their combined error estimates
    assert abs(I[0] - ref) <= E[0] + 3.0 * ref_err
