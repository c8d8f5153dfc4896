"""The oracle against the reference's own regression for the scattering
path: tests/Mie/reference/LossySphere_327.EMTPFT (scuff-scatter on
Sphere_327.msh, MATERIAL CONST_EPS_10_SIGMA_1, plane wave along +z with
polarization (0.70710678, 0.70710678i, 0), powers by energy/momentum
transfer; tests/Mie/TestMie.sh:77-88).  The file prints 7 digits; the
whole chain -- mesh, edges, assembly, RHS, solve and the EMT power -- is
held to 1e-5 relative (the reference's own check, Checklist.EMTPFT, allows
1e-2).  Five of the fifteen frequencies here; all fifteen run on the GPU."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

E0 = np.array([0.70710678, 0.70710678j, 0.0])
TOL = 1e-5


def reference_rows():
    return np.loadtxt(os.path.join(GOLDEN, "LossySphere_327.EMTPFT"), usecols=(0, 2, 3))


@pytest.fixture(scope="module")
def lossy(O):
    G = O.Geometry()
    G.add_msh(os.path.join(GOLDEN, "Sphere_327.msh"), kind=3, eps=10.0, wp=1.0)   # CONST_EPS_10_SIGMA_1
    return G


def test_reference_file_shape():
    ref = reference_rows()
    assert ref.shape == (15, 3)
    assert np.allclose(ref[:, 0], 0.2 * np.arange(1, 16))


@pytest.mark.parametrize("row", [0, 4, 8, 11, 14])
def test_oracle_emt_power_matches_reference(lossy, row):
    w, pabs_ref, pscat_ref = reference_rows()[row]
    M = lossy.assemble(w)
    R = lossy.rhs_planewave(w, E0=E0, nhat=(0, 0, 1))
    KN = np.linalg.solve(M, R)
    P = lossy.power_emt(w, KN, R, lossy.assemble_region(w, 0))
    assert abs(P["PAbs"] / pabs_ref - 1) <= TOL, (w, P)
    assert abs(P["PScat"] / pscat_ref - 1) <= TOL, (w, P)


def test_oracle_region_matrices_sum_to_the_bem_matrix(lossy):
    """exterior + interior region matrices = the full matrix"""
    w = 0.8
    M = lossy.assemble(w)
    Me, Mi = lossy.assemble_region(w, 0), lossy.assemble_region(w, 1)
    assert np.linalg.norm(Me + Mi - M) <= 1e-14 * np.linalg.norm(M)
    assert np.linalg.norm(Mi) > 0.1 * np.linalg.norm(M)
