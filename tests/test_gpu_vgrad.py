"""Vertex-position derivatives (scuffx_vertex_gradient and its callers)
against central finite differences of the GPU chain itself, and the
summed vertex gradient of a rigidly moved surface against the GradM path
(AssembleBEMMatrixBlock with GradM, libscuff.h:500-505), which the oracle
pins by its own central differences (tests/test_oracle_gradm.py).

Finite differences use h = 1e-5 (radius 1 spheres).  The kernels
differentiate far pairs exactly (the assembly's degree-4 rule), separated
near pairs with the 78-point rule on each panel quarter and touching /
self pairs by central differences of their Taylor-Duffy integrals, so the
agreement with finite differences of the FIPPI-based assembly is set by
the quadrature accuracy of the near terms.  The assembly integrates its
near-pair records to the reference's tolerances (relative 1e-6 for
separated pairs, 1e-5 for common-edge/vertex pairs: QIFIPPICubature.cc:41-42
and its Taylor-Duffy counterpart), while the kernels differentiate
integrals converged to 1e-10; the two derivatives therefore agree to about
1e-5 of the largest component, and the tolerances below are set there.

The icosphere's vertices carry exact coordinate ties (symmetric vertices
share coordinates).  At a tie the canonical vertex order of a near pair's
FIPPI record (AssessPanelPair.cc:342-455) switches under a 1e-5 move, and
the finite difference of the assembled matrix then carries the record's
integration error divided by h.  Every vertex is therefore displaced by a
fixed random amount (1-2% of the panel size) before differencing; the
sphere keeps its topology and vertex count."""
import time

import numpy as np
import pytest
import torch

import scuffx
from scuffx.autograd import AbsorbedPower, LogDet

pytestmark = pytest.mark.gpu

PW = scuffx.PlaneWave([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])
H = 1.0e-5


def sphere(level, mat, center=None, jitter=None):
    V, T = scuffx.icosphere(level, center=center)
    V = np.array(V, dtype=np.float64).reshape(-1, 3)
    if jitter is None:
        jitter = 0.02 / 2 ** level
    V = V + np.random.default_rng(11 + level).uniform(-jitter, jitter, V.shape)
    G = scuffx.RWGGeometry()
    G.AddMesh(V, T, mat)
    return G, V


def rnd(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return torch.tensor(z, dtype=torch.complex128, device="cuda")


def uMv(G, Omega, V, u, v, ns=0):
    G.SetVertices(ns, V)
    M = G.AssembleBEMMatrix(Omega)
    return complex(torch.dot(u, M.t.T @ v))


def fd_all(f, V0, coords):
    out = {}
    for (iv, c) in coords:
        Vp, Vm = V0.copy(), V0.copy()
        Vp[iv, c] += H
        Vm[iv, c] -= H
        out[(iv, c)] = (f(Vp) - f(Vm)) / (2 * H)
    return out


@pytest.mark.parametrize("mat,Omega", [("CONST_EPS_4+0.1i", 1.0), ("PEC", 1.0), ("PEC", 0.7j)])
def test_matrix_gradient_vs_fd_l1(mat, Omega):
    G, V0 = sphere(1, mat)
    N = G.TotalBFs
    u, v = rnd(N, 1), rnd(N, 2)
    M = G.AssembleBEMMatrix(Omega)
    t0 = time.perf_counter()
    g = G.VertexGradient(Omega, M, u=u, v=v)[0]
    t1 = time.perf_counter()
    coords = [(iv, c) for iv in range(V0.shape[0]) for c in range(3)]
    fd = fd_all(lambda V: uMv(G, Omega, V, u, v), V0, coords)
    G.SetVertices(0, V0)
    err = max(abs(g[iv, c] - fd[(iv, c)]) for (iv, c) in coords)
    scale = max(abs(fd[k]) for k in fd)
    print(f"\n  {mat} omega={Omega}: N={N}, gradient {1e3 * (t1 - t0):.1f} ms, "
          f"max |adjoint - FD| / max |FD| = {err / scale:.2e}")
    assert err <= 1e-5 * scale


def test_dense_weights_equal_rank_one():
    """C = u v^T as a dense matrix gives the rank-one result"""
    G, V0 = sphere(1, "CONST_EPS_4+0.1i")
    N = G.TotalBFs
    u, v = rnd(N, 3), rnd(N, 4)
    M = G.AssembleBEMMatrix(1.0)
    g1 = G.VertexGradient(1.0, M, u=u, v=v)[0]
    Cd = scuffx.HMatrix(N, N, data=torch.outer(v, u).contiguous())   # t[c, r] = u_r v_c
    g2 = G.VertexGradient(1.0, M, C=Cd)[0]
    assert np.abs(g1 - g2).max() <= 1e-12 * np.abs(g1).max()


def test_rhs_gradient_vs_fd():
    G, V0 = sphere(1, "CONST_EPS_4+0.1i")
    N = G.TotalBFs
    mu = rnd(N, 5)
    g = G.RHSVertexGradient(1.0, PW, mu)[0]

    def f(V):
        G.SetVertices(0, V)
        return complex(torch.dot(mu, G.AssembleRHSVector(1.0, PW).t))

    coords = [(iv, c) for iv in range(V0.shape[0]) for c in range(3)]
    fd = fd_all(f, V0, coords)
    G.SetVertices(0, V0)
    err = max(abs(g[iv, c] - fd[(iv, c)]) for (iv, c) in coords)
    assert err <= 1e-7 * max(abs(x) for x in fd.values())


def test_pabs_explicit_and_gamma():
    G, V0 = sphere(1, "CONST_EPS_4+0.1i")
    N = G.TotalBFs
    M = G.AssembleBEMMatrix(1.0)
    b = G.AssembleRHSVector(1.0, PW)
    x = scuffx.HVector(N, data=b.t.clone())
    M.LUFactorize()
    M.LUSolve(x)
    gexp, gamma = G.PAbsVertexGradient(x)

    def P(V, xx):
        G.SetVertices(0, V)
        return G.GetPower(1.0, xx, b)["PAbs"]

    coords = [(iv, c) for iv in range(V0.shape[0]) for c in range(3)]
    fd = fd_all(lambda V: P(V, x), V0, coords)
    err = max(abs(gexp[0][iv, c] - fd[(iv, c)]) for (iv, c) in coords)
    assert err <= 1e-7 * max(abs(v) for v in fd.values())
    # PAbs is quadratic in KN: the central difference along d is exact
    G.SetVertices(0, V0)
    d = rnd(N, 6) * 1e-3 * float(x.t.abs().max())
    xp = scuffx.HVector(N, data=x.t + d)
    xm = scuffx.HVector(N, data=x.t - d)
    lhs = (P(V0, xp) - P(V0, xm)) / 2.0
    rhs = float(torch.dot(gamma, d).real)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def _pabs(G, V, Omega):
    return float(AbsorbedPower.apply(torch.tensor(V), G, 0, Omega, PW))


def test_absorbed_power_gradient_l1_every_vertex():
    G, V0 = sphere(1, "CONST_EPS_4+0.1i")
    Vt = torch.tensor(V0, requires_grad=True)
    P = AbsorbedPower.apply(Vt, G, 0, 1.0, PW)
    P.backward()
    g = Vt.grad.numpy()
    coords = [(iv, c) for iv in range(V0.shape[0]) for c in range(3)]
    fd = fd_all(lambda V: _pabs(G, V, 1.0), V0, coords)
    err = max(abs(g[iv, c] - fd[(iv, c)]) for (iv, c) in coords)
    scale = max(abs(v) for v in fd.values())
    print(f"\n  L1 PAbs={float(P):.6e}: max |adjoint - FD| / max |FD| = {err / scale:.2e}")
    assert err <= 1e-4 * scale


def test_absorbed_power_gradient_l3_all_642_vertices():
    """eps = 4+0.1i sphere, level 3 (642 vertices, N=3840), omega=1: the
    gradient over all 642 vertices against central differences along two
    random directions of all 1926 coordinates, and coordinate by
    coordinate at 6 vertices"""
    G, V0 = sphere(3, "CONST_EPS_4+0.1i")
    assert V0.shape[0] == 642
    Vt = torch.tensor(V0, requires_grad=True)
    t0 = time.perf_counter()
    P = AbsorbedPower.apply(Vt, G, 0, 1.0, PW)
    t1 = time.perf_counter()
    P.backward()
    t2 = time.perf_counter()
    g = Vt.grad.numpy()
    print(f"\n  L3 PAbs={float(P):.6e}: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s")
    rng = np.random.default_rng(7)
    for k in range(2):
        d = rng.standard_normal(V0.shape)
        d /= np.linalg.norm(d)
        fd = (_pabs(G, V0 + H * d, 1.0) - _pabs(G, V0 - H * d, 1.0)) / (2 * H)
        ad = float((g * d).sum())
        print(f"  direction {k}: adjoint {ad:.10e} FD {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e} "
              f"(|gradient| {np.linalg.norm(g):.3e})")
        # a random direction mostly cancels: its derivative is ~1e-2 of the
        # gradient's norm, so the per-component agreement shows ~1e2 larger
        assert abs(ad - fd) <= 2e-3 * abs(fd)
    scale = np.abs(g).max()
    for iv in rng.choice(642, 6, replace=False):
        for c in range(3):
            Vp, Vm = V0.copy(), V0.copy()
            Vp[iv, c] += H
            Vm[iv, c] -= H
            fd = (_pabs(G, Vp, 1.0) - _pabs(G, Vm, 1.0)) / (2 * H)
            assert abs(g[iv, c] - fd) <= 1e-4 * scale, (iv, c, g[iv, c], fd)


def test_logdet_vertex_gradient_two_pec_spheres():
    """two PEC spheres (level 3 each, EFIE N=3840) 3 um apart along z at
    the imaginary frequency xi = 1: the z-components of the log det
    gradient summed over the second sphere's vertices equal
    Re Tr(M^{-1} dM/dz) from the GradM blocks; a random vertex direction
    against central differences"""
    V1, T1 = scuffx.icosphere(3)
    V2, T2 = scuffx.icosphere(3, center=[0.0, 0.0, 3.0])
    G = scuffx.RWGGeometry()
    G.AddMesh(V1, T1, "PEC")
    G.AddMesh(V2, T2, "PEC")
    Omega = 1.0j
    N = G.TotalBFs
    V20 = np.array(V2, dtype=np.float64).reshape(-1, 3)
    Vt = torch.tensor(V20, requires_grad=True)
    L = LogDet.apply(Vt, G, 1, Omega)
    L.backward()
    g = Vt.grad.numpy()
    # GradM path: block (sphere 2 rows, sphere 1 columns) and its z-derivative
    M = G.AssembleBEMMatrix(Omega)
    M.LUFactorize()
    _, grads = G.AssembleBEMMatrixBlock(1, 0, Omega, GradM=True)
    o1, n1 = G.BFOffset(0)
    o2, n2 = G.BFOffset(1)
    dM = scuffx.HMatrix(N, N)
    dB = grads[2].t.T            # (n2 rows, n1 columns)
    dM.t[o1:o1 + n1, o2:o2 + n2] = dB.T    # t[c, r]: columns of sphere 1, rows of sphere 2
    dM.t[o2:o2 + n2, o1:o1 + n1] = dB      # the transposed block
    tr = M.TraceSolve(dM).real
    sz = g[:, 2].sum()
    # the sum also carries the sphere's own terms, which cancel under a
    # translation only to the accuracy of each vertex's derivative
    mass = np.abs(g[:, 2]).sum()
    print(f"\n  d logdet/dz: vertex sum {sz:.12e}, GradM trace {tr:.12e}, rel {abs(sz - tr) / abs(tr):.2e}, "
          f"sum |g_z| {mass:.3e}")
    assert abs(sz - tr) <= 1e-6 * mass + 1e-8 * abs(tr)

    def f(V):
        return float(LogDet.apply(torch.tensor(V), G, 1, Omega))

    d = np.random.default_rng(9).standard_normal(V20.shape)
    d /= np.linalg.norm(d)
    fd = (f(V20 + H * d) - f(V20 - H * d)) / (2 * H)
    ad = float((g * d).sum())
    print(f"  random direction: adjoint {ad:.10e} FD {fd:.10e}")
    assert abs(ad - fd) <= 1e-4 * abs(fd)
