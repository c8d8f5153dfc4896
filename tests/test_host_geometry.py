"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point of include/scuffx.h, and the product's mesh topology and
edge indexing agree bit for bit with the oracle (InitEdgeList.cc:47-371,
RWGSurface.cc:55-97) on icospheres and on the reference's own test meshes.
"""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, read_msh

HEADER = os.path.join(ROOT, "include", "scuffx.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(scuffx_\w+)\s*\(", text, re.M)))


def test_abi_exports_every_declared_symbol(sx):
    import subprocess
    syms = declared_symbols()
    assert len(syms) >= 20
    L = sx.lib()
    for s in syms:
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", sx._LIBPATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (scuffx_\w+)", out))
    assert set(syms) == exported


def test_abi_version_and_errors(sx):
    L = sx.lib()
    assert L.scuffx_version() >= 100
    # argument errors are reported without touching the device
    assert L.scuffx_geom_total_bfs(None) == -1
    assert b"null" in L.scuffx_last_error()
    assert L.scuffx_lu_factor(None, 4, 4, None, 64) == -1
    assert L.scuffx_lu_solve(None, 4, 4, None, b"N", 1, None, 4, 64) == -1
    with pytest.raises(sx.SCUFFXError):
        G = sx.RWGGeometry()
        G.AddMesh(np.zeros((3, 3)), np.array([[0, 1, 2]]), "NOT_A_MATERIAL")


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_icosphere_closed_forms(sx, O, level):
    V, T = sx.icosphere(level)
    assert V.shape == (10 * 4 ** level + 2, 3) and T.shape == (20 * 4 ** level, 3)
    assert np.allclose(np.linalg.norm(V, axis=1), 1.0, atol=1e-15)
    Vo, To = O.icosphere(level)
    assert np.array_equal(V, Vo) and np.array_equal(T, To)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "PEC")
    nV, nP, nE = G.Counts(0)
    assert nE == 30 * 4 ** level and nV - nE + nP == 2
    assert G.TotalBFs == nE
    # outward normals
    P = G.Panels(0)
    assert np.all(np.einsum("ij,ij->i", P[:, 0:3], P[:, 3:6]) > 0)


def _compare_topology(G, Go, ns=0):
    e, le = G.Edges(ns)
    eo, leo = Go.edges(ns)
    assert np.array_equal(e, eo)
    assert np.array_equal(le, leo)
    assert np.array_equal(G.Panels(ns), Go.panels(ns))


@pytest.mark.parametrize("level", [1, 2, 3])
def test_edges_bit_exact_icosphere(sx, O, level):
    V, T = O.icosphere(level)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_4+0.1i")
    Go = O.Geometry()
    Go.add_surface(V, T, kind=1, eps=4 + 0.1j)
    assert G.TotalBFs == Go.N == 2 * 30 * 4 ** level
    _compare_topology(G, Go)


@pytest.mark.parametrize("mesh,nedges", [("UnitTestSphere_R0P75_414.msh", 414), ("SSphere_255.msh", 357)])
def test_edges_bit_exact_reference_meshes(sx, O, mesh, nedges):
    path = os.path.join(GOLDEN, mesh)
    G = sx.RWGGeometry()
    G.AddMeshFile(path, "PEC")
    Go = O.Geometry()
    Go.add_msh(path)
    nV, nP, nE = G.Counts(0)
    assert nE == nedges and nV - nE + nP == 2
    _compare_topology(G, Go)


def test_gmsh_roundtrip(sx, O, tmp_path):
    V, T = O.icosphere(2, 1.0)
    fn = str(tmp_path / "s.msh")
    O.write_gmsh(fn, V, T)
    G = sx.RWGGeometry()
    G.AddMeshFile(fn, "PEC")
    Go = O.Geometry()
    Go.add_surface(V, T, kind=0)
    _compare_topology(G, Go)


def test_scuffgeo_objects_and_displacement(sx, O, tmp_path):
    mesh = os.path.join(GOLDEN, "SSphere_255.msh")
    geo = tmp_path / "two.scuffgeo"
    geo.write_text(
        "# two spheres\n"
        f"OBJECT Lower\n  MESHFILE {mesh}\nENDOBJECT\n\n"
        f"OBJECT Upper\n  MESHFILE {mesh}\n  DISPLACED 0 0 3\n  MATERIAL CONST_EPS_4+0.1i\nENDOBJECT\n")
    G = sx.RWGGeometry(str(geo))
    assert G.NumSurfaces == 2
    assert G.TotalBFs == 357 + 2 * 357
    # the oracle sees the displaced vertices directly
    Go = O.Geometry()
    Go.add_msh(mesh)
    P0 = Go.panels(0)
    P1 = G.Panels(1)
    assert np.array_equal(P1[:, 3:], P0[:, 3:]) or np.allclose(P1[:, 3:], P0[:, 3:], rtol=0, atol=1e-15)
    assert np.allclose(P1[:, 0:3] - P0[:, 0:3], [0, 0, 3], rtol=0, atol=1e-14)


def test_scuffgeo_rotation_composes(sx, tmp_path):
    mesh = os.path.join(GOLDEN, "SSphere_255.msh")
    geo = tmp_path / "rot.scuffgeo"
    geo.write_text(f"OBJECT A\n MESHFILE {mesh}\n ROTATED 90 ABOUT 0 0 1\n DISPLACED 1 0 0\nENDOBJECT\n")
    G = sx.RWGGeometry(str(geo))
    G0 = sx.RWGGeometry()
    G0.AddMeshFile(mesh)
    c0 = G0.Panels(0)[:, 0:3]
    c1 = G.Panels(0)[:, 0:3]
    want = np.stack([-c0[:, 1] + 1, c0[:, 0], c0[:, 2]], axis=1)
    assert np.allclose(c1, want, atol=1e-14)


def test_scuffgeo_errors(sx, tmp_path):
    bad = tmp_path / "bad.scuffgeo"
    bad.write_text("OBJECT A\n MESHFILE nothere.msh\nENDOBJECT\n")
    with pytest.raises(sx.SCUFFXError):
        sx.RWGGeometry(str(bad))
    bad.write_text("OBJECT A\n MESHFILE x.msh\n DISPLACED 1 2\nENDOBJECT\n")
    with pytest.raises(sx.SCUFFXError, match="DISPLACED"):
        sx.RWGGeometry(str(bad))


@pytest.mark.parametrize("name,ok", [("PEC", True), ("VACUUM", True), ("CONST_EPS_12+0.01i", True),
                                     ("CONST_EPS_2.5_MU_1.5", True), ("DRUDE_10_0.1", True),
                                     ("CONST_EPS_2.5_SIGMA_1e3", True), ("GOLD", False),
                                     ("CONST_EPS_x", False)])
def test_material_names(sx, name, ok):
    V, T = sx.icosphere(0)
    G = sx.RWGGeometry()
    if ok:
        G.AddMesh(V, T, name)
        assert G.TotalBFs == (30 if name == "PEC" else 60)
    else:
        with pytest.raises(sx.SCUFFXError):
            G.AddMesh(V, T, name)


def test_mates(sx, tmp_path):
    """Mate[] (RWGGeometry.cc:733-770): same mesh file and material ->
    the first earlier such surface; another material, PEC vs dielectric
    or raw arrays -> -1"""
    mesh = os.path.join(GOLDEN, "SSphere_255.msh")
    geo = tmp_path / "mates.scuffgeo"
    geo.write_text(
        f"OBJECT A\n  MESHFILE {mesh}\n  MATERIAL CONST_EPS_4\nENDOBJECT\n"
        f"OBJECT B\n  MESHFILE {mesh}\n  DISPLACED 0 0 3\n  MATERIAL CONST_EPS_4\nENDOBJECT\n"
        f"OBJECT C\n  MESHFILE {mesh}\n  DISPLACED 0 0 6\n  MATERIAL CONST_EPS_5\nENDOBJECT\n"
        f"OBJECT D\n  MESHFILE {mesh}\n  DISPLACED 0 0 9\nENDOBJECT\n"
        f"OBJECT E\n  MESHFILE {mesh}\n  ROTATED 30 ABOUT 0 0 1\n  DISPLACED 0 0 12\nENDOBJECT\n"
        f"OBJECT F\n  MESHFILE {mesh}\n  DISPLACED 0 0 15\n  MATERIAL CONST_EPS_4\nENDOBJECT\n")
    G = sx.RWGGeometry(str(geo))
    assert G.Mate == [-1, 0, -1, -1, 3, 0]
    G2 = sx.RWGGeometry()
    G2.AddMeshFile(mesh, "PEC")
    G2.AddMeshFile(mesh, "PEC")
    assert G2.Mate == [-1, 0]


@pytest.mark.parametrize("mesh", ["Sphere_327.msh", "UnitTestSphere_R0P75_414.msh", "reversed-icosphere"])
def test_panel_normals_point_into_exterior(sx, O, mesh):
    """Panel normals of a dielectric surface point into its exterior
    region whatever the vertex order of the mesh's triangles
    (RWGGeometry.cc:653-690; the crossed-overlap absorbed power depends on
    it).  A PEC surface keeps the right-hand-rule normals: no non-PEC
    surface bounds its exterior.  Product and oracle agree bit for bit."""
    if mesh == "reversed-icosphere":
        V, T = O.icosphere(2)
        T = T[:, ::-1].copy()
        path = None
    else:
        path = os.path.join(GOLDEN, mesh)
    for mat, kind in (("CONST_EPS_4", 1), ("PEC", 0)):
        G = sx.RWGGeometry()
        Go = O.Geometry()
        if path:
            G.AddMeshFile(path, mat)
            Go.add_msh(path, kind=kind, eps=4.0)
        else:
            G.AddMesh(V, T, mat)
            Go.add_surface(V, T, kind=kind, eps=4.0)
        P = G.Panels(0)
        assert np.array_equal(P, Go.panels(0))
        c = P[:, 0:3] - P[:, 0:3].mean(axis=0)
        outward = np.einsum("ij,ij->i", c, P[:, 3:6]) > 0
        if kind == 1:
            assert np.all(outward)
        else:
            # right-hand rule from the vertex order
            Vg, Tg = (V, T) if path is None else read_msh(path)
            n = np.cross(Vg[Tg[:, 1]] - Vg[Tg[:, 0]], Vg[Tg[:, 2]] - Vg[Tg[:, 0]])
            assert np.all(np.einsum("ij,ij->i", n, P[:, 3:6]) > 0)


def test_scuffgeo_user_defined_material(sx, tmp_path):
    """MATERIAL ... ENDMATERIAL sections (UserDefinedMaterials.cc:145-320):
    constants, Eps(w) and Mu(w) expressions of w in rad/s; an OBJECT refers
    to the material by name.  A Drude block equals the closed form, and
    equals the built-in DRUDE material at the same plasma frequency."""
    mesh = os.path.join(GOLDEN, "Sphere_327.msh")
    geo = tmp_path / "gold.scuffgeo"
    geo.write_text("MATERIAL Gold\n"
                   "   wp = 1.37e16;    # plasma frequency, rad/s\n"
                   "   gamma = 5.32e13;\n"
                   "   Eps(w) = 1 - wp^2 / (w*(w + i*gamma));\n"
                   "ENDMATERIAL\n\n"
                   "MATERIAL Lossy\n  s = 2.5\n  Eps(w) = 3 + i*s*2.99792458e14/w\n  Mu(w) = 1.5\nENDMATERIAL\n\n"
                   f"OBJECT A\n\tMESHFILE {mesh}\n\tMATERIAL Gold\nENDOBJECT\n"
                   f"OBJECT B\n\tMESHFILE {mesh}\n\tDISPLACED 0 0 3\n\tMATERIAL Lossy\nENDOBJECT\n")
    G = sx.RWGGeometry(str(geo))
    c = 2.99792458e14
    for om in (0.1, 1.0, 2.5, 1.0j):
        w = om * c
        e, m = G.GetEpsMu(om, Region=1)
        ref = 1 - 1.37e16 ** 2 / (w * (w + 1j * 5.32e13))
        assert abs(e - ref) <= 1e-13 * abs(ref) and m == 1
        e2, m2 = G.GetEpsMu(om, Region=2)
        assert abs(e2 - (3 + 2.5j / om)) <= 1e-14 * abs(e2) and m2 == 1.5
    assert G.GetEpsMu(1.0, Region=0) == (1, 1)
    # the built-in Drude material in units of c / 1 um
    geo2 = tmp_path / "drude.scuffgeo"
    geo2.write_text(f"OBJECT A\n\tMESHFILE {mesh}\n\tMATERIAL DRUDE_{1.37e16 / c!r}_{5.32e13 / c!r}\nENDOBJECT\n")
    G2 = sx.RWGGeometry(str(geo2))
    for om in (0.1, 1.0):
        e, _ = G.GetEpsMu(om, Region=1)
        e2, _ = G2.GetEpsMu(om, Region=1)
        assert abs(e - e2) <= 1e-12 * abs(e)
    for bad in ("MATERIAL X\n  Eps(w) = 1 + foo\nENDMATERIAL\n", "MATERIAL X\n  a = 1 2\nENDMATERIAL\n",
                "MATERIAL X\n  Eps(w) = (1 + w\nENDMATERIAL\n", "MATERIAL X\n  Eps(w) = 1\n"):
        geo3 = tmp_path / "bad.scuffgeo"
        geo3.write_text(bad + f"OBJECT A\n\tMESHFILE {mesh}\n\tMATERIAL X\nENDOBJECT\n")
        with pytest.raises(sx.SCUFFXError):
            sx.RWGGeometry(str(geo3))


def test_set_vertices_rederives_panels_and_edges(sx, O):
    """SetVertices keeps the topology and re-derives panel and edge data:
    a uniform scaling by 2 doubles lengths and radii and quadruples areas;
    a vertex moved and moved back restores everything bit for bit; the
    result equals a fresh surface built from the new vertices (the oracle's
    geometry on them, bit for bit); errors leave the mesh untouched."""
    V, T = O.icosphere(2)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_4")
    P0, E0 = G.Panels(0), G.Edges(0)
    G.SetVertices(0, 2.0 * V)
    P1, E1 = G.Panels(0), G.Edges(0)
    assert np.array_equal(E1[0], E0[0])                    # topology
    assert np.allclose(E1[1], 2.0 * E0[1], rtol=1e-15)     # lengths
    assert np.allclose(P1[:, 7], 4.0 * P0[:, 7], rtol=1e-14)
    assert np.allclose(P1[:, 3:6], P0[:, 3:6], atol=1e-15)  # normals keep their orientation
    W = V.copy()
    W[5] += np.array([0.01, -0.02, 0.03])
    G.SetVertices(0, W)
    Go = O.Geometry()
    Go.add_surface(W, T, kind=1, eps=4.0)
    assert np.array_equal(G.Panels(0), Go.panels(0))
    assert np.array_equal(G.Vertices(0), W)
    G.SetVertices(0, V)
    assert np.array_equal(G.Panels(0), P0) and np.array_equal(G.Edges(0)[1], E0[1])
    with pytest.raises(sx.SCUFFXError):
        G.SetVertices(0, V[:-1])
    D = V.copy()
    D[T[0, 1]] = D[T[0, 0]]            # a degenerate panel
    with pytest.raises(sx.SCUFFXError):
        G.SetVertices(0, D)
    assert np.array_equal(G.Panels(0), P0)
    G.TransformSurface(0, D=[0.0, 0.0, 1.0])
    with pytest.raises(sx.SCUFFXError):
        G.SetVertices(0, V)
    G.UnTransform()
    G.SetVertices(0, V)
