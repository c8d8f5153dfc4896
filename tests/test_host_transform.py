"""Host-side pins of the geometrical transformations (GTransformation.cc,
RWGGeometry::Transform / UnTransform, RWGSurface.cc:588-667) and of the
.trans reader (GTransformation.cc:373-580), without a GPU.

Pins: rotations against the closed-form Rodrigues matrix; the order in
which DISP / ROT compose (each later step acts on the result of the
earlier ones); panel centroids / normals re-derived from the moved
vertices; UnTransform back to the original coordinates up to rounding;
the reference's own Spheres.trans fixture (doc/docs/tests/CasimirSpheres)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def rodrigues(axis, deg):
    u = np.asarray(axis, float) / np.linalg.norm(axis)
    t = np.radians(deg)
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return np.cos(t) * np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * np.outer(u, u)


def two_spheres(sx, O, level=1):
    V, T = O.icosphere(level)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, Label="Upper")
    G.AddMesh(V + np.array([0.0, 0.0, -3.0]), T, "CONST_EPS_4", Label="Lower")
    return G, V


def write(tmp_path, text, name="t.trans"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reference_spheres_trans(sx):
    G = sx.ReadTransFile(os.path.join(GOLDEN, "Spheres.trans"))
    tags = [g.Tag for g in G]
    assert len(tags) == 21 and tags[0] == "3.0" and tags[4] == "4.9" and tags[-1] == "100"


def test_no_file_is_one_identity(sx, O):
    T = sx.ReadTransFile(None)
    assert [g.Tag for g in T] == ["DEFAULT"]
    G, V = two_spheres(sx, O, 0)
    G.Transform(T[0])
    assert G.SurfaceMoved == [False, False]
    assert np.array_equal(G.Vertices(0), V)


def test_rotation_then_displacement_one_line(sx, O, tmp_path):
    T = sx.ReadTransFile(write(tmp_path, "TRANS rot OBJECT Upper ROT 30 ABOUT 0 1 0 DISP 0 0 2.5 "
                                         "OBJECT lower DISPLACED 1 0 0\n"))
    G, V = two_spheres(sx, O)
    V1 = G.Vertices(1).copy()
    G.Transform(T[0])
    assert G.SurfaceMoved == [True, True]
    want = V @ rodrigues([0, 1, 0], 30).T + np.array([0, 0, 2.5])
    assert np.abs(G.Vertices(0) - want).max() <= 1e-15 * 4
    assert np.array_equal(G.Vertices(1), V1 + np.array([1.0, 0, 0]))
    # panels follow their vertices: centroid, rotated normal, same area
    P0 = G.Panels(0)
    Vt = G.Vertices(0)
    _, Tr = O.icosphere(1)
    assert np.abs(P0[:, 0:3] - Vt[Tr].mean(axis=1)).max() <= 1e-15 * 4
    G.UnTransform()
    assert np.abs(G.Vertices(0) - V).max() <= 1e-15 * 8
    assert np.abs(G.Vertices(1) - V1).max() <= 1e-15 * 8
    P1 = G.Panels(0)
    assert np.abs(P1[:, 3:6] - (P0[:, 3:6] @ rodrigues([0, 1, 0], 30))).max() <= 1e-14


def test_section_composes_in_order(sx, O, tmp_path):
    """DISPLACED then ROTATED: the rotation acts on the displaced surface"""
    T = sx.ReadTransFile(write(tmp_path, "# comment\n\nTRANSFORMATION sec\n OBJECT UPPER\n"
                                         " DISPLACED 0 0 1\n ROTATED 90 ABOUT 0 0 1\nENDTRANSFORMATION\n"
                                         "TRANS d SURFACE Upper DISP 0 0 3.0\n"))
    assert [g.Tag for g in T] == ["sec", "d"]
    G, V = two_spheres(sx, O)
    G.Transform(T[0])
    assert G.SurfaceMoved == [True, False]
    R = rodrigues([0, 0, 1], 90)
    want = (V + np.array([0, 0, 1.0])) @ R.T
    assert np.abs(G.Vertices(0) - want).max() <= 1e-15 * 8
    G.UnTransform()
    G.Transform(T[1])
    assert G.SurfaceMoved == [True, False]
    assert np.abs(G.Vertices(0) - (V + np.array([0, 0, 3.0]))).max() <= 1e-15 * 8


@pytest.mark.parametrize("text,msg", [
    ("FOO bar\n", "unknown token"),
    ("TRANSFORMATION t\n OBJECT Upper\n DISPLACED 0 0 1 2\nENDTRANSFORMATION\n", "junk at end of line"),
    ("TRANSFORMATION t\n OBJECT\nENDTRANSFORMATION\n", "requires one argument"),
    ("TRANSFORMATION t\n OBJECT Upper\n DISPLACED 0 0 1\n", "end of file"),
    ("TRANS t OBJECT Upper DISP 0 0\n", "too few values"),
    ("TRANS t OBJECT Upper ROT 30 AROUND 0 0 1\n", "invalid syntax"),
    ("TRANS t OBJECT Upper DISP 0 x 1\n", "bad value"),
    ("TRANS t OBJECT\n", "empty OBJECT"),
    ("TRANS t OBJECT Upper SPIN 3\n", "unknown keyword"),
])
def test_trans_file_errors(sx, tmp_path, text, msg):
    with pytest.raises(sx.SCUFFXError, match=msg):
        sx.ReadTransFile(write(tmp_path, text))


def test_labels_and_checks(sx, O, tmp_path):
    G, _ = two_spheres(sx, O, 0)
    assert (G.Label(0), G.Label(1)) == ("Upper", "Lower")
    assert G.GetSurfaceByLabel("lower") == 1 and G.GetSurfaceByLabel("Middle") == -1
    V, T = O.icosphere(0)
    G.AddMesh(V + 5.0, T)
    assert G.Label(2) == "ByHand"
    bad = sx.ReadTransFile(write(tmp_path, "TRANS a OBJECT Middle DISP 0 0 1\n"))
    with pytest.raises(sx.SCUFFXError, match="unknown OBJECT/SURFACE Middle"):
        G.CheckGTCList(bad)
    v0 = G.Vertices(0).copy()
    with pytest.raises(sx.SCUFFXError):
        G.Transform(bad[0])
    assert np.array_equal(G.Vertices(0), v0)
    nolabel = sx.ReadTransFile(write(tmp_path, "TRANS a DISP 0 0 1\n", "n.trans"))
    with pytest.raises(sx.SCUFFXError, match="no OBJECT/SURFACE"):
        G.CheckGTCList(nolabel)


def test_scuffgeo_labels(sx, tmp_path):
    mesh = os.path.join(GOLDEN, "Sphere_327.msh")
    geo = tmp_path / "g.scuffgeo"
    geo.write_text(f"OBJECT UpperSphere\n\tMESHFILE {mesh}\n\tDISPLACED 0 0 3\nENDOBJECT\n"
                   f"OBJECT\n\tMESHFILE {mesh}\nENDOBJECT\n")
    G = sx.RWGGeometry(str(geo))
    assert G.Label(0) == "UpperSphere"
    assert G.Label(1) == mesh
    T = sx.ReadTransFile(os.path.join(GOLDEN, "Spheres.trans"))
    G.CheckGTCList(T)
    v = G.Vertices(0).copy()
    G.Transform(T[4])                      # tag 4.9: UpperSphere DISP 0 0 1.9
    assert np.array_equal(G.Vertices(0), v + np.array([0.0, 0.0, 1.9]))
    G.UnTransform()
    assert np.abs(G.Vertices(0) - v).max() <= 1e-15 * 8
