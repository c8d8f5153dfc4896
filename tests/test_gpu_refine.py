"""Refinement of complex64 LU solutions through the C ABI
(scuffx_refine, scuffx_gmres_ir): the status they return, the best
iterate they keep, and GMRES-IR where classical refinement stalls.

Matrices: A = U diag(s) V^H with Haar-random unitary U, V (fixed seeds)
and singular values log-spaced from 1 to 1/kappa.  The residual relative to
|b| cannot fall below about kappa * 1e-16, so tolerances sit above that.
With kappa = 1e7 classical refinement contracts by only ~0.7 per step and
misses 1e-9 within 30 steps; with kappa = 3e7 its first correction raises
the residual (kappa * 6e-8 > 1) and is undone.  In both cases the library
continues with GMRES-IR (20 and ~300 steps in a numpy model of the same
algorithm).  Answers are checked against numpy's complex128 solve
(forward error <= kappa * 1e-14) and the residual against the tolerance."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def haar(n, rng):
    Z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2)
    Q, R = np.linalg.qr(Z)
    return Q * (np.diag(R) / np.abs(np.diag(R)))


def conditioned(n, kappa, seed):
    rng = np.random.default_rng(seed)
    s = np.logspace(0, -np.log10(kappa), n)
    A = (haar(n, rng) * s) @ haar(n, rng).conj().T
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b


def run(sx, A, b, fn, tol, maxiter):
    import torch
    n = A.shape[0]
    M = sx.HMatrix.from_numpy(A)
    LU32 = torch.empty((n, n), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    bt = torch.from_numpy(b).cuda()
    x = torch.zeros(n, dtype=torch.complex128, device="cuda")
    L = sx.lib()
    sx._sync_stream()
    assert L.scuffx_lu32_from_c128(sx._ptr(M.t), n, n, sx._ptr(LU32), n, sx._ptr(ipiv)) == 0
    it, rel = C.c_int(), C.c_double()
    rc = getattr(L, fn)(sx._ptr(M.t), n, n, sx._ptr(LU32), n, sx._ptr(ipiv), sx._ptr(bt), sx._ptr(x),
                        tol, maxiter, C.byref(it), C.byref(rel))
    return rc, x.cpu().numpy(), it.value, rel.value


def resid(A, x, b):
    return np.max(np.abs((b - A @ x).real) + np.abs((b - A @ x).imag)) / np.max(np.abs(b.real) + np.abs(b.imag))


@pytest.mark.parametrize("fn", ["scuffx_refine", "scuffx_gmres_ir"])
def test_well_conditioned_converges(sx, fn):
    A, b = conditioned(700, 1e3, 11)
    rc, x, it, rel = run(sx, A, b, fn, 1e-12, 40)
    xr = np.linalg.solve(A, b)
    print(f"  {fn} rc={rc} iters={it} relres={rel:.2e}", flush=True)
    assert rc == 0
    assert rel <= 1e-12 and resid(A, x, b) <= 2e-12
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


@pytest.mark.parametrize("kappa,tol,maxiter", [(1e7, 1e-9, 30), (3e7, 1e-8, 400)])
def test_slow_or_stalled_refinement_continues_with_gmres(sx, kappa, tol, maxiter, monkeypatch):
    """At condition 3e7 the complex64 factors barely precondition: the
    factorization's updates use four real products per complex product
    there (SCUFFX_GEMM_3M=0); the three-product updates, whose normwise
    error is a few times larger, stall at that condition."""
    if kappa > 2e7:
        monkeypatch.setenv("SCUFFX_GEMM_3M", "0")
    A, b = conditioned(640, kappa, 5)
    rc, x, it, rel = run(sx, A, b, "scuffx_refine", tol, maxiter)
    xr = np.linalg.solve(A, b)
    fe = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(f"  kappa={kappa:.0e} rc={rc} iters={it} relres={rel:.2e} forward error {fe:.2e}", flush=True)
    assert rc == 0
    assert rel <= tol and resid(A, x, b) <= 2 * tol
    assert fe <= kappa * 1e-14


def test_not_converged_status_and_best_iterate(sx):
    """maxiter = 0: no correction is allowed, so the status is
    SCUFFX_NOT_CONVERGED, x is the complex64 solution and relres its
    residual; the Python front end raises on it."""
    A, b = conditioned(400, 1e6, 7)
    rc, x, it, rel = run(sx, A, b, "scuffx_refine", 1e-14, 0)
    assert rc == sx.NOT_CONVERGED
    assert it == 0 and np.isfinite(rel) and rel > 1e-14
    assert abs(resid(A, x, b) - rel) <= 1e-3 * rel
    # an unreachable tolerance: the best iterate is returned, never a worse one
    rc2, x2, it2, rel2 = run(sx, A, b, "scuffx_refine", 1e-30, 25)
    assert rc2 == sx.NOT_CONVERGED and rel2 <= rel
    # near the floor (~kappa * 1e-16) host and device residuals differ by
    # rounding of the same order
    assert resid(A, x2, b) <= max(2 * rel2, 1e-9)
    M = sx.HMatrix.from_numpy(A)
    import torch
    bb = sx.HVector(400, data=torch.from_numpy(b).cuda())
    with pytest.raises(sx.SCUFFXError):
        sx.solve_mixed(M, bb, tol=1e-30, maxiter=3)


def test_gmres_ir_bem_power(sx, O):
    """GMRES-IR on the dielectric sphere (level 3, N=3840): powers equal the
    complex128 LU's within 1e-10."""
    V, T = O.icosphere(3)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_12+0.01i")
    pw = sx.PlaneWave([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    M = G.AssembleBEMMatrix(1.0)
    RHS = G.AssembleRHSVector(1.0, pw)
    x, it, rel = sx.solve_mixed(M, RHS, tol=1e-12, maxiter=60, method="gmres")
    Pg = G.GetPower(1.0, x, RHS)
    KN = G.AssembleRHSVector(1.0, pw)
    M.LUFactorize()
    M.LUSolve(KN)
    Pd = G.GetPower(1.0, KN, RHS)
    print(f"  gmres iters={it} relres={rel:.2e} {Pg} {Pd}", flush=True)
    for k in Pd:
        assert abs(Pg[k] - Pd[k]) <= 1e-10 * abs(Pd[k]), k
