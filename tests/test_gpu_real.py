"""Real matrices of imaginary frequencies (LHM_REAL, AllocateBEMMatrix(
PureImagFreq=true), AssembleBEMMatrix.cc:778-783), packed symmetric storage
(LHM_SYMMETRIC, HMatrix.cc:325-342) and LUInvert (LBWrappers.cc:251-335).

* the real LU (dgetrf semantics: pivot = first largest |a|) against
  SciPy's LAPACK dgetrf: identical pivots, factors within 1e-12; solves
  'N', 'T', 'C' within 1e-12 of numpy's;
* LUInvert (complex128, complex64, real) against numpy.linalg.inv;
* at omega = i xi the complex assembly has exactly zero imaginary parts, so
  the real matrix equals its real part bit for bit (full and packed
  storage, whole matrix and blocks with derivatives), and the oracle's
  matrix within 1e-12;
* packed storage: the factors of the unpacked copy give the same solution
  and log det as the full matrix;
* the Casimir integrands computed in real arithmetic equal the complex ones
  within 1e-12 relative."""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import rel_fro

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [100, 700, 2600])
def test_real_lu_matches_dgetrf(sx, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    M = sx.HMatrix.from_numpy(A)
    assert M.prec == 8 and M.RealComplex == "LHM_REAL"
    assert M.LUFactorize() == 0
    lu, piv = sl.lu_factor(A)
    ipiv = M.ipiv.cpu().numpy() - 1
    assert np.array_equal(ipiv, piv)
    F = M.numpy()
    assert rel_fro(F, lu) <= 1e-12
    b = rng.standard_normal((n, 3))
    for trans in ("N", "T", "C"):
        X = sx.HMatrix.from_numpy(b)
        M.LUSolve(X, Trans=trans)
        ref = np.linalg.solve(A if trans == "N" else A.T, b)
        assert rel_fro(X.numpy(), ref) <= 1e-10, trans
    la = M.LogDet()
    sign, ld = np.linalg.slogdet(A)
    assert abs(la.real - ld) <= 1e-10 * abs(ld)
    # the phase sums arg(U_ii) and pi per interchange (up to ~n terms)
    assert abs(np.exp(1j * la.imag) - sign) <= 1e-14 * n


def test_real_lu_singular_info(sx):
    A = np.random.default_rng(3).standard_normal((200, 200))
    A[:, 57] = 0.0
    M = sx.HMatrix.from_numpy(A)
    assert M.LUFactorize() == 58


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-10), ("complex64", 2e-3), ("float64", 1e-10)])
def test_lu_invert(sx, dtype, tol):
    n = 640
    rng = np.random.default_rng(7)
    A = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if dtype != "float64" else 0)
    M = sx.HMatrix.from_numpy(A, dtype=dtype)
    assert M.LUFactorize() == 0
    assert M.LUInvert() == 0
    assert rel_fro(M.numpy(), np.linalg.inv(A)) <= tol
    # a zero diagonal entry of U is reported and the matrix left as it was
    Z = np.triu(rng.standard_normal((50, 50)))
    Z[10, 10] = 0.0
    Mz = sx.HMatrix.from_numpy(Z.astype(np.complex128) if dtype != "float64" else Z, dtype=dtype)
    info = Mz.LUFactorize()
    assert info == 11
    before = Mz.numpy()
    assert Mz.LUInvert() == 11
    assert np.array_equal(Mz.numpy(), before)


@pytest.mark.parametrize("material,kind,eps", [("PEC", 0, 1.0), ("CONST_EPS_4", 1, 4.0)])
def test_real_assembly_is_real_part(sx, O, material, kind, eps):
    """omega = 0.7i on a level-2 sphere with a material real on the
    imaginary axis (the case the reference's real matrices assume): exact
    zero imaginary parts, the real matrix (full and packed) equal to the
    real part bit for bit, and the oracle's matrix within 1e-12."""
    V, T = O.icosphere(2)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, material)
    w = 0.7j
    Mc = G.AssembleBEMMatrix(w).numpy()
    assert np.all(Mc.imag == 0.0)
    Mr = G.AllocateBEMMatrix(PureImagFreq=True)
    assert Mr.prec == 8
    Mr = G.AssembleBEMMatrix(w, Mr)
    R = Mr.numpy()
    assert np.array_equal(R, Mc.real)
    Mp = G.AssembleBEMMatrix(w, G.AllocateBEMMatrix(PureImagFreq=True, Packed=True))
    assert Mp.packed and Mp.t.numel() == R.shape[0] * (R.shape[0] + 1) // 2
    assert np.array_equal(Mp.numpy(), R)
    Go = O.Geometry()
    Go.add_surface(V, T, kind=kind, eps=eps)
    Mo = Go.assemble(w)
    assert rel_fro(R, Mo.real) <= 1e-12
    # complex packed storage of a real frequency: the upper triangle of the full matrix
    Mcp = G.AssembleBEMMatrix(0.7, G.AllocateBEMMatrix(Packed=True))
    assert np.array_equal(Mcp.numpy(), G.AssembleBEMMatrix(0.7).numpy())


def test_packed_lu_solve_logdet(sx, O):
    V, T = O.icosphere(2)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_4")
    w = 1.3j
    Mf = G.AssembleBEMMatrix(w, G.AllocateBEMMatrix(PureImagFreq=True))
    Mp = G.AssembleBEMMatrix(w, G.AllocateBEMMatrix(PureImagFreq=True, Packed=True))
    A = Mf.numpy()
    assert Mf.LUFactorize() == 0 and Mp.LUFactorize() == 0
    assert abs(Mf.LogDet() - Mp.LogDet()) == 0.0
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    import torch
    x1 = sx.HVector(len(b), data=torch.from_numpy(b.copy()).cuda())
    x2 = sx.HVector(len(b), data=torch.from_numpy(b.copy()).cuda())
    Mf.LUSolve(x1)
    Mp.LUSolve(x2)
    assert np.array_equal(x1.numpy(), x2.numpy())
    assert np.linalg.norm(A @ x1.numpy() - b) <= 1e-10 * np.linalg.norm(b)
    assert Mp.LUInvert() == 0
    assert rel_fro(Mp.numpy(), np.linalg.inv(A)) <= 1e-9


def test_real_blocks_with_derivatives(sx, O):
    V, T = O.icosphere(1)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "PEC")
    G.AddMesh(V + np.array([0.0, 0.0, 3.0]), T, "CONST_EPS_10")
    w = 0.4j
    axes = [[0.0, 0.0, 1.0]]
    Bc, Gc, Tc = G.AssembleBEMMatrixBlock(0, 1, w, GradM=True, TorqueAxes=axes)
    Br, Gr, Tr = G.AssembleBEMMatrixBlockReal(0, 1, w, GradM=True, TorqueAxes=axes)
    assert Br.prec == 8
    for c, r in [(Bc, Br)] + list(zip(Gc, Gr)) + list(zip(Tc, Tr)):
        cn = c.numpy()
        assert np.all(cn.imag == 0.0)
        assert np.array_equal(r.numpy(), cn.real)
    Dc = G.AssembleBEMMatrixBlock(1, 1, w).numpy()
    Dp = G.AssembleBEMMatrixBlockReal(1, 1, w, Packed=True)
    assert Dp.packed and np.array_equal(Dp.numpy(), Dc.real)


def test_casimir_real_equals_complex(sx, O):
    from scuffx.cas3d import SC3Data
    V, T = O.icosphere(1)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "PEC")
    G.AddMesh(V + np.array([0.0, 0.0, 2.6]), T, "PEC")
    r = SC3Data(G, Energy=True, ZForce=True, TorqueAxes=[[0.0, 0.0, 1.0]])
    c = SC3Data(G, Energy=True, ZForce=True, TorqueAxes=[[0.0, 0.0, 1.0]], Real=False)
    assert r.M.prec == 8 and c.M.prec == 64
    for xi in (0.05, 0.8, 3.0):
        er, ec = r.GetCasimirIntegrand(xi), c.GetCasimirIntegrand(xi)
        print(f"  xi={xi} real {er} complex {ec}", flush=True)
        assert np.all(np.abs(er - ec) <= 1e-12 * np.maximum(np.abs(ec), 1e-300) + 1e-300)
