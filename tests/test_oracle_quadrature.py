"""Pins of the oracle's quadrature layer against mathematics.

* triangle rules (libs/libTriInt/TCR.cc:83-91, :556-576): exact for all
  monomials u^a v^b up to their stated degree on the standard triangle,
  whose moments are a! b! / (a+b+2)!;
* hcubature (libs/libSGJC/hcubature.c:897+) and pcubature
  (libs/libSGJC/pcubature.c) on integrals with closed forms, including
  vector integrands and an endpoint singularity;
* Taylor-Duffy (libs/libscuff/TaylorDuffy.cc:277-367) on the worked
  example of doc/docs/reference/SingularIntegrals.md:440-493.
"""
from math import factorial

import numpy as np
import pytest


@pytest.mark.parametrize("order,degree,npts", [(4, 4, 6), (20, 20, 78)])
def test_tcr_polynomial_exactness(O, order, degree, npts):
    r = O.tcr(order)
    assert r.shape == (npts, 3)
    assert np.all(r[:, 0] >= 0) and np.all(r[:, 1] >= 0) and np.all(r[:, 0] + r[:, 1] <= 1 + 1e-15)
    for d in range(degree + 1):
        for a in range(d + 1):
            b = d - a
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            q = np.sum(r[:, 2] * r[:, 0] ** a * r[:, 1] ** b)
            # TCR4 constants carry 13 digits
            assert abs(q - exact) <= 2e-11 * exact, (order, a, b)
    # one degree higher is no longer exact: a wrong table would not show this
    worst = 0.0
    for a in range(degree + 2):
        b = degree + 1 - a
        exact = factorial(a) * factorial(b) / factorial(a + b + 2)
        worst = max(worst, abs(np.sum(r[:, 2] * r[:, 0] ** a * r[:, 1] ** b) - exact) / exact)
    assert worst > 1e-9


def test_hcubature_2d_exponential(O):
    v, e, n = O.hcubature(lambda x: [np.exp(x[0] + x[1]), x[0] * x[1]], 2, [0, 0], [1, 1], reltol=1e-10)
    assert abs(v[0] - (np.e - 1) ** 2) < 1e-9
    assert abs(v[1] - 0.25) < 1e-14
    assert n > 17


def test_hcubature_3d_vector(O):
    # 3-D rule (Genz-Malik 7/5, 33 points per region)
    f = lambda x: [np.cos(x[0]) * np.cos(x[1]) * np.cos(x[2]), x[0] ** 2 + x[1] * x[2]]
    v, e, n = O.hcubature(f, 2, [0, 0, 0], [1, 2, 3], reltol=1e-11)
    assert abs(v[0] - np.sin(1) * np.sin(2) * np.sin(3)) < 1e-10
    assert abs(v[1] - (1 / 3 * 6 + 1 * 2 * 4.5)) < 1e-10
    assert n % 33 == 0


def test_hcubature_1d_singular(O):
    # Gauss-Kronrod 15 with bisection: int_0^1 x^(-1/2) = 2
    v, e, n = O.hcubature(lambda x: [1 / np.sqrt(x[0])], 1, [0], [1], abstol=1e-9, reltol=1e-9, maxeval=200000)
    assert abs(v[0] - 2.0) < 1e-7
    assert n % 15 == 0


def test_hcubature_maxeval_stops(O):
    v, e, n = O.hcubature(lambda x: [1 / np.sqrt(x[0] + x[1])], 1, [0, 0], [1, 1], reltol=1e-15, maxeval=200)
    assert n <= 200 + 17 * 2


def test_pcubature_polynomial_and_smooth(O):
    v, e, n = O.pcubature(lambda x: [np.cos(x[0]) * x[1] ** 2], 1, [0, 0], [1, 1], reltol=1e-12)
    assert abs(v[0] - np.sin(1) / 3) < 1e-13
    # Clenshaw-Curtis of order 2^m+1 integrates low polynomials exactly at once
    v, e, n = O.pcubature(lambda x: [x[0] ** 3 - 2 * x[0] + 1], 1, [-1], [2], reltol=1e-12)
    assert abs(v[0] - (15 / 4 - 3 + 3)) < 1e-13
    assert n <= 9


def test_pcubature_3d(O):
    v, e, n = O.pcubature(lambda x: [np.exp(-(x[0] ** 2 + x[1] ** 2 + x[2] ** 2))], 1, [0] * 3, [1] * 3, reltol=1e-12)
    from scipy.special import erf
    assert abs(v[0] - (np.sqrt(np.pi) / 2 * erf(1)) ** 3) < 1e-12


def test_taylor_duffy_doc_example(O):
    """doc/docs/forDevelopers/SingularIntegrals.md (TestTaylorDuffy
    listing and sample output): common-edge pair, P in {UNITY, PMCHWC},
    kernels r^-1 and r^-3 (KIndex RP), Q=V3, Q'=V3', RelTol 1e-10 and
    MaxEval 25 give 33 integrand samples and the printed values."""
    V1, V2, V3 = [0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.05, 0.1, 0.0]
    V3P = [0.07, -0.08, 0.03]
    Vs = np.array([V1, V2, V3, V2, V3P])
    Q = np.array([V3, V3P])
    res, ncalls = O.taylor_duffy(2, [0, 3], [0, 0], [-1.0, -3.0], Vs, Q, maxeval=25)
    assert ncalls == 33
    assert f"{res[0].real:.8e}" == "3.56771698e-01"
    assert f"{res[1].real:.8e}" == "1.57003236e-03"
