"""Host-side pins of the Casimir caller's imaginary-frequency integrator
(IntegrateCliffFunction, libs/libTriInt/IntegrateCliffFunction.cc:98-320)
against closed-form integrals of Fermi-type 'cliff' functions
    f(x) = A / (1 + exp((x - c) / w)),
    int_a^inf f = A w log(1 + exp((c - a) / w)),
and of the argument checks of the Casimir workspace."""
import math

import numpy as np
import pytest


def cliff(A, c, w):
    return lambda x: A / (1.0 + math.exp(min((x - c) / w, 700.0)))


def exact(A, c, w, a):
    return A * w * math.log1p(math.exp((c - a) / w))


@pytest.mark.parametrize("reltol", [1e-2, 1e-4, 1e-6])
def test_cliff_integral_closed_form(reltol):
    from scuffx.cas3d import IntegrateCliffFunction
    comps = [(-7.9e-3, 1.2, 0.4), (2.5, 0.05, 0.01), (1e-9, 3.0, 1.0)]
    fs = [cliff(*p) for p in comps]
    calls = []

    def f(x, skip):
        calls.append(x)
        return [g(x) for g in fs]

    I, E = IntegrateCliffFunction(f, 3, 0.001, 0.0, 0.7, 0.0, reltol)
    for k, p in enumerate(comps):
        ref = exact(*p, 0.001)
        err = abs(I[k] - ref) / abs(ref)
        assert err <= 10 * reltol, (k, err)
        assert E[k] >= 0
    assert all(x >= 0.001 for x in calls)


def test_cliff_zero_component_and_finite_range():
    from scuffx.cas3d import IntegrateCliffFunction
    g = cliff(1.0, 2.0, 0.3)
    I, E = IntegrateCliffFunction(lambda x, s: [g(x), 0.0], 2, 0.0, 0.0, 1.0, 0.0, 1e-5)
    assert I[1] == 0.0 and E[1] == 0.0
    assert abs(I[0] - exact(1.0, 2.0, 0.3, 0.0)) <= 1e-4
    # finite upper limit: a constant integrates exactly (Simpson)
    I, _ = IntegrateCliffFunction(lambda x, s: [2.0], 1, 1.0, 4.0, 0.0, 0.0, 1e-3)
    assert abs(I[0] - 6.0) <= 1e-12


def test_casimir_workspace_argument_checks(sx, O):
    from scuffx.cas3d import SC3Data
    V, T = O.icosphere(0)
    G = sx.RWGGeometry()
    G.AddMesh(V, T)
    with pytest.raises(sx.SCUFFXError):
        SC3Data(G)                           # nothing requested
    with pytest.raises(sx.SCUFFXError):
        SC3Data(G, ZForce=True)              # forces need two surfaces
    G.AddMesh(V + np.array([0.0, 0.0, 3.0]), T)
    with pytest.raises(sx.SCUFFXError):
        SC3Data(G, TorqueAxes=[(0, 0, 1)] * 4)
