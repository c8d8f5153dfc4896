"""Pins of the oracle's Mie series and dense LU.

* Mie: the power table of unitTests/unit-test-PFT.cc:81-120 (PEC, SiO2 and
  gold spheres of radius 1 at omega = 0.01, 0.1, 1; absorbed and scattered
  power for a unit plane wave, P = Q pi R^2 / (2 Z0)), printed to 7 digits.
* LU: the LAPACK contract the reference calls (zgetrf_/zgetrs_,
  libs/libhmat/LBWrappers.cc:151-245): same pivots and factors as the
  LAPACK zgetrf that scipy links, and solves with Trans N/T/C.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

PFT = {
    "PEC": [(0.01, None, 0.0, 1.389884e-10), (0.1, None, 0.0, 1.393182e-06), (1.0, None, 0.0, 8.488640e-03)],
    "SiO2": [(0.01, 3.634170e+00 + 3.794466e-03j, 5.981414e-08, 2.430533e-11),
             (0.1, 3.778579e+00 + 4.786666e-02j, 7.245380e-06, 2.580820e-07),
             (1.0, 1.429487e+00 + 3.647089e-02j, 1.774032e-04, 1.351696e-04)],
    "Gold": [(0.01, -6.610472e+04 + 1.172275e+06j, 1.373257e-05, 1.297492e-10),
             (0.1, -5.031480e+04 + 8.922668e+04j, 3.741193e-05, 1.352849e-06),
             (1.0, -2.020863e+03 + 3.585437e+02j, 8.270806e-05, 8.860794e-03)],
}


@pytest.mark.parametrize("material", sorted(PFT))
def test_mie_pft_table(O, material):
    f = np.pi / (2 * O.ZVAC)
    for omega, eps, pabs, pscat in PFT[material]:
        qs, qe = O.mie(omega, None if eps is None else np.sqrt(eps))
        assert abs(qs * f - pscat) <= 1e-6 * pscat, (material, omega)
        if pabs:
            assert abs((qe - qs) * f - pabs) <= 1e-6 * pabs, (material, omega)
        else:
            assert abs(qe - qs) < 1e-12


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


@pytest.mark.parametrize("n", [1, 2, 7, 64, 130])
def test_zgetrf_matches_lapack(O, n):
    A = _rand(n, 20261202 + n)
    LU, ipiv, info = O.zgetrf(A)
    lu2, piv2, info2 = lapack.zgetrf(A)
    assert info == info2 == 0
    assert np.array_equal(ipiv, piv2 + 1)          # oracle pivots are 1-based
    assert np.max(np.abs(LU - lu2)) <= 1e-12 * np.max(np.abs(lu2))


def test_zgetrf_singular_info(O):
    A = _rand(6, 3)
    A[:, 3] = 0
    _, _, info = O.zgetrf(A)
    _, _, info2 = lapack.zgetrf(A)
    assert info == info2 > 0


def test_zgetrf_pivot_rule_uses_abs1(O):
    """zgetrf picks the pivot by |Re|+|Im| (izamax), not by modulus."""
    # |0.7+0.7i|: |Re|+|Im| = 1.4 beats 1.2, while the modulus 0.99 would not
    A = np.array([[0.7 + 0.7j, 1.0], [1.2 + 0.0j, 2.0]])
    _, ipiv, _ = O.zgetrf(A)
    assert ipiv[0] == 1
    B = np.array([[1.2 + 0.0j, 1.0], [0.7 + 0.7j, 2.0]])
    _, ipiv, _ = O.zgetrf(B)
    assert ipiv[0] == 2
    assert np.array_equal(lapack.zgetrf(A)[1] + 1, O.zgetrf(A)[1])
    assert np.array_equal(lapack.zgetrf(B)[1] + 1, O.zgetrf(B)[1])


@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_zgetrs(O, trans):
    n = 40
    A = _rand(n, 11)
    B = _rand(n, 12)[:, :3]
    LU, ipiv, _ = O.zgetrf(A)
    X = O.zgetrs(LU, ipiv, B, trans)
    op = {"N": A, "T": A.T, "C": A.conj().T}[trans]
    assert np.linalg.norm(op @ X - B) <= 1e-12 * np.linalg.norm(B) * np.linalg.cond(A)


def test_oracle_mate_copies_diagonal_block(O):
    """Mate (AssembleBEMMatrix.cc:728-740): the declared duplicate's
    diagonal block is the first surface's; everything else unchanged"""
    V, T = O.icosphere(0)
    d = np.array([0.0, 0.0, 3.0])
    A = O.Geometry(); A.add_surface(V, T, kind=1, eps=4.0); A.add_surface(V + d, T, kind=1, eps=4.0)
    B = O.Geometry(); B.add_surface(V, T, kind=1, eps=4.0); B.add_surface(V + d, T, kind=1, eps=4.0)
    B.set_mate(1, 0)
    Ma, Mb = A.assemble(0.7), B.assemble(0.7)
    n = B.num_bfs(0)
    assert np.array_equal(Mb[n:, n:], Mb[:n, :n])
    assert np.array_equal(Mb[:n, :], Ma[:n, :]) and np.array_equal(Mb[:, :n], Ma[:, :n])
    # the computed block agrees with the copy to the FIPPI cubature tolerance
    assert np.linalg.norm(Ma[n:, n:] - Mb[n:, n:]) <= 1e-6 * np.linalg.norm(Ma[n:, n:])


@pytest.mark.parametrize("kind", [0, 1])
def test_oracle_compact_rows_equal_full_matrix_rows(O, kind):
    """The compact row assembly (used where the full matrix does not fit on
    the host) gives the full assembly's upper-triangle rows bit for bit."""
    V, T = O.icosphere(1)
    G = O.Geometry()
    G.add_surface(V, T, kind=kind, eps=4 + 0.1j)
    M = G.assemble(0.7)
    nb = 1 if kind == 0 else 2
    for e0 in (0, 17, 27):
        R = G.rows(0.7, e0, e0 + 3)
        r0 = nb * e0
        for i in range(R.shape[0]):
            r = r0 + i
            assert np.array_equal(R[i, r:], M[r, r:])
