"""GPU parity of RWGGeometry::AssembleBEMMatrix (AssembleBEMMatrix.cc:
684-773) against the oracle: full matrices at sizes the oracle assembles
in seconds, sampled rows of the 3840-unknown dielectric sphere, the
frequencies of unitTests/unit-test-BEMMatrix.cc (0.01, 1, 0.1i), several
surfaces with PEC and dielectric blocks, and size-independent properties
(symmetry, FIPPI table size) at every size.

Tolerance: relative Frobenius error 1e-12 over the whole matrix (or over
the sampled rows)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_fro

pytestmark = pytest.mark.gpu

TOL = 1e-12


def both(sx, O, V, T, material, kind, eps=1.0):
    G = sx.RWGGeometry()
    G.AddMesh(V, T, material)
    Go = O.Geometry()
    Go.add_surface(V, T, kind=kind, eps=eps)
    return G, Go


def check_full(G, Go, omega, tol=TOL):
    M = G.AssembleBEMMatrix(omega).numpy()
    Mo = Go.assemble(omega)
    err = rel_fro(M, Mo)
    print(f"  N={M.shape[0]} omega={omega} rel_fro={err:.3e}", flush=True)
    assert err <= tol, err
    return M, Mo


def test_pec_icosphere_l1(sx, O):
    V, T = O.icosphere(1)
    G, Go = both(sx, O, V, T, "PEC", 0)
    M, Mo = check_full(G, Go, 1.0)
    assert np.array_equal(M, M.T)
    # the same FIPPI table: one record per distinct canonical key
    assert G.AssemblyStats()["nfippi"] == Go.fippi_count()


def test_dielectric_icosphere_l1(sx, O):
    V, T = O.icosphere(1)
    G, Go = both(sx, O, V, T, "CONST_EPS_4+0.1i", 1, 4 + 0.1j)
    M, Mo = check_full(G, Go, 1.0)
    assert np.array_equal(M, M.T)


@pytest.mark.parametrize("omega", [0.01, 1.0, 0.1j])
def test_pec_reference_mesh_unit_test_frequencies(sx, O, omega):
    """SSphere_255.msh at the frequencies of unit-test-BEMMatrix.cc:136-157."""
    path = os.path.join(GOLDEN, "SSphere_255.msh")
    G = sx.RWGGeometry()
    G.AddMeshFile(path, "PEC")
    Go = O.Geometry()
    Go.add_msh(path)
    check_full(G, Go, omega)


@pytest.mark.parametrize("omega", [0.01, 1.0, 0.1j])
def test_dielectric_reference_mesh(sx, O, omega):
    path = os.path.join(GOLDEN, "SSphere_255.msh")
    G = sx.RWGGeometry()
    G.AddMeshFile(path, "CONST_EPS_11.7+0.1i")
    Go = O.Geometry()
    Go.add_msh(path, kind=1, eps=11.7 + 0.1j)
    check_full(G, Go, omega)


def test_two_surfaces_pec_and_dielectric(sx, O, tmp_path):
    """Two objects, one PEC and one dielectric (mixed 1x2 blocks and the
    exterior-region coupling of SurfaceSurfaceInteractions.cc:200-327)."""
    Va, Ta = O.icosphere(1, 1.0)
    Vb = Va + np.array([0.0, 0.0, 2.5])
    G = sx.RWGGeometry()
    G.AddMesh(Va, Ta, "PEC")
    G.AddMesh(Vb, Ta, "CONST_EPS_4+0.1i")
    Go = O.Geometry()
    Go.add_surface(Va, Ta, kind=0)
    Go.add_surface(Vb, Ta, kind=1, eps=4 + 0.1j)
    M, Mo = check_full(G, Go, 0.7)
    assert M.shape == (120 + 240, 120 + 240)
    assert np.array_equal(M, M.T)


def test_touching_surfaces_near_pairs_across_blocks(sx, O):
    """Two spheres 0.05 apart: near pairs (and FIPPI records) between
    different surfaces."""
    Va, Ta = O.icosphere(1, 1.0)
    Vb = Va + np.array([2.05, 0.0, 0.0])
    G = sx.RWGGeometry()
    G.AddMesh(Va, Ta, "CONST_EPS_2+0.5i")
    G.AddMesh(Vb, Ta, "CONST_EPS_2+0.5i")
    Go = O.Geometry()
    Go.add_surface(Va, Ta, kind=1, eps=2 + 0.5j)
    Go.add_surface(Vb, Ta, kind=1, eps=2 + 0.5j)
    check_full(G, Go, 1.0)
    assert G.AssemblyStats()["nfippi"] == Go.fippi_count()


def test_pec_icosphere_l2_config1(sx, O):
    """PEC sphere, icosphere level 2, EFIE N=480, omega=1."""
    V, T = O.icosphere(2)
    G, Go = both(sx, O, V, T, "PEC", 0)
    check_full(G, Go, 1.0)


@pytest.mark.parametrize("omega", [0.5, 1.0, 2.0])
def test_dielectric_l3_sampled_rows(sx, O, omega):
    """Dielectric sphere eps=4+0.1i, level 3, PMCHWT N=3840: the GPU
    matrix against oracle rows of 12 edges spread over the mesh."""
    V, T = O.icosphere(3)
    G, Go = both(sx, O, V, T, "CONST_EPS_4+0.1i", 1, 4 + 0.1j)
    M = G.AssembleBEMMatrix(omega).numpy()
    N = M.shape[0]
    assert N == 3840
    assert np.array_equal(M, M.T)
    Mo = np.zeros((N, N), dtype=np.complex128, order="F")
    for e0 in (0, 613, 1400, 1916):
        Go.assemble_rows(omega, e0, e0 + 3, Mo)
    errs = []
    for e0 in (0, 613, 1400, 1916):
        r0, r1 = 2 * e0, 2 * (e0 + 3)
        mask = np.zeros((r1 - r0, N), dtype=bool)
        for i in range(r1 - r0):
            # the oracle fills the upper triangle of its rows
            mask[i, r0 + i:] = True
        A = np.where(mask, M[r0:r1], 0)
        B = np.where(mask, Mo[r0:r1], 0)
        errs.append(rel_fro(A, B))
    print(f"  omega={omega} sampled rel_fro={max(errs):.3e}", flush=True)
    assert max(errs) <= TOL


def test_wrong_size_matrix_is_replaced(sx, O):
    V, T = O.icosphere(0)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "PEC")
    M0 = sx.HMatrix(5, 5)
    M = G.AssembleBEMMatrix(1.0, M0)
    assert M is not M0 and M.NR == 30
    M2 = G.AssembleBEMMatrix(1.0, M)
    assert M2 is M


def test_mate_blocks_copied(sx, O, tmp_path):
    """two identical dielectric spheres from one mesh file: the second
    diagonal block is the first's, copied (AssembleBEMMatrix.cc:733-740);
    the whole matrix matches the oracle, which computes both blocks"""
    from conftest import GOLDEN, read_msh
    mesh = os.path.join(GOLDEN, "SSphere_255.msh")
    geo = tmp_path / "pair.scuffgeo"
    geo.write_text(f"OBJECT A\n  MESHFILE {mesh}\n  MATERIAL CONST_EPS_4+0.1i\nENDOBJECT\n"
                   f"OBJECT B\n  MESHFILE {mesh}\n  DISPLACED 0.1 0 2.7\n  MATERIAL CONST_EPS_4+0.1i\nENDOBJECT\n")
    print("mate blocks", flush=True)
    G = sx.RWGGeometry(str(geo))
    assert G.Mate == [-1, 0]
    M = G.AssembleBEMMatrix(0.8).numpy()
    V, T = read_msh(mesh)
    Go = O.Geometry()
    Go.add_surface(V, T, kind=1, eps=4 + 0.1j)
    Go.add_surface(V + np.array([0.1, 0.0, 2.7]), T, kind=1, eps=4 + 0.1j)
    Go.set_mate(1, 0)
    Mo = Go.assemble(0.8)
    e = rel_fro(M, Mo)
    print(f"  rel Frobenius {e:.2e}", flush=True)
    assert e <= 1e-12
    o1, n1 = G.BFOffset(1)
    assert np.array_equal(M[o1:o1 + n1, o1:o1 + n1], M[:n1, :n1])
