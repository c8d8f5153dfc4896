"""Pin of the oracle's displacement gradient of a BEM block (GradM of
AssembleBEMMatrixBlock, SurfaceSurfaceInteractions.cc:195-327,
EdgeEdgeInteractions.cc:183-198, PanelPanelInteractions.cc:132-199)
against central finite differences of the oracle's own block under a
rigid displacement of the row surface.  Near pairs are present (gap 0.5
on level-0 spheres), so both the 6-point and the 78-point gradient rules
are exercised.

Tolerance 1e-4: for near pairs the reference takes the gradient from the
78-point rule without desingularization (PanelPanelInteractions.cc:504-510)
while the matrix entries themselves come from desingularization plus the
FIPPI terms, so the two differ by the 78-point rule's error (~1e-5 here,
independent of the step h).  A wrong sign or component gives O(1)."""
import numpy as np
import pytest


def two_spheres(O, d, kind, eps):
    V, T = O.icosphere(0)
    G = O.Geometry()
    G.add_surface(V, T, kind=kind, eps=eps)
    G.add_surface(V + np.asarray(d), T, kind=kind, eps=eps)
    return G


@pytest.mark.parametrize("kind,eps,omega", [(0, 1.0, 0.9), (1, 4 + 0.1j, 0.7), (1, 2.0, 0.5j)])
def test_gradm_matches_central_differences(O, kind, eps, omega):
    d0 = np.array([0.3, -0.2, 2.5])
    B, GB, _ = two_spheres(O, d0, kind, eps).assemble_block_grad(omega, 0, 1)
    h = 1e-4
    for m in range(3):
        e = np.zeros(3)
        e[m] = h
        # moving the row surface by +h is moving the column surface by -h
        Bp = two_spheres(O, d0 - e, kind, eps).assemble_block_grad(omega, 0, 1)[0]
        Bm = two_spheres(O, d0 + e, kind, eps).assemble_block_grad(omega, 0, 1)[0]
        fd = (Bp - Bm) / (2 * h)
        assert np.linalg.norm(GB[m] - fd) <= 1e-4 * np.linalg.norm(fd), m


def gamma(axis):
    """rotation generator G with G x = u x x for the unit axis u (the
    matrix CreateGammaMatrix builds, EdgeEdgeInteractions.cc:241-270)"""
    u = np.asarray(axis, dtype=float)
    u = u / np.linalg.norm(u)
    return np.array([[0.0, -u[2], u[1]], [u[2], 0.0, -u[0]], [-u[1], u[0], 0.0]])


def rotate(V, axis, th):
    u = np.asarray(axis, dtype=float)
    u = u / np.linalg.norm(u)
    K = gamma(u)
    R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    return V @ R.T


@pytest.mark.parametrize("kind,eps,omega", [(0, 1.0, 0.8), (1, 4 + 0.1j, 0.6)])
def test_dmdtheta_matches_central_differences(O, kind, eps, omega):
    """Angular derivative of the block under a rotation of the row surface
    about an axis through the origin (PanelPanelInteractions.cc:182-199)."""
    V, T = O.icosphere(0)
    V = V + np.array([0.4, 0.1, -0.2])
    W = V + np.array([0.3, -0.2, 2.5])
    axes = [(0, 0, 1), (1, 1, 0), (0.3, -0.5, 0.8)]

    def geom(Va):
        G = O.Geometry()
        G.add_surface(Va, T, kind=kind, eps=eps)
        G.add_surface(W, T, kind=kind, eps=eps)
        return G
    B, GB, TB = geom(V).assemble_block_grad(omega, 0, 1, gammas=[gamma(a) for a in axes])
    h = 1e-4
    for t, a in enumerate(axes):
        Bp = geom(rotate(V, a, h)).assemble_block_grad(omega, 0, 1)[0]
        Bm = geom(rotate(V, a, -h)).assemble_block_grad(omega, 0, 1)[0]
        fd = (Bp - Bm) / (2 * h)
        assert np.linalg.norm(TB[t] - fd) <= 1e-4 * np.linalg.norm(fd), a
