"""The adjoint contraction of the BEM matrix derivative (scuffx_adjoint)
and the differentiable solve built on it (scuffx.autograd.DisplacedSolve).

Pins: the contraction equals the dense GradM / dMdT blocks of the
reference's path (AssembleBEMMatrixBlock with GradM, itself pinned by
central differences in the oracle) contracted in numpy, over both the
(a, b) blocks and their transposes (1e-12); and the autograd gradient of
two real losses of the solution equals central finite differences of the
solve under the displacement (1e-6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def two_spheres(sx, O, mat="CONST_EPS_4+0.1i"):
    V, T = O.icosphere(1)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, mat, Label="a")
    # far apart enough that no cross pair is near (no adaptive FIPPI
    # cubature in the derivative blocks, whose noise would swamp the
    # finite differences)
    G.AddMesh(V + np.array([0.3, 0.0, 3.2]), T, mat, Label="b")
    G.AddMesh(V + np.array([0.0, 3.3, -0.4]), T, mat, Label="c")
    return G


def test_contraction_matches_dense_blocks(sx, O):
    import torch
    print("adjoint contraction vs dense GradM", flush=True)
    G = two_spheres(sx, O)
    N = G.TotalBFs
    rng = np.random.default_rng(11)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    X = sx.HVector(N, data=torch.from_numpy(x).cuda())
    Y = sx.HVector(N, data=torch.from_numpy(y).cuda())
    axes = [np.array([0.0, 0.0, 1.0]), np.array([1.0, 1.0, 0.0])]
    omega = 0.9
    for ns in range(3):
        s = G.AdjointContraction(ns, omega, X, Y, TorqueAxes=axes)
        want = np.zeros(5, dtype=complex)
        oa, na = G.BFOffset(ns)
        for b in range(3):
            if b == ns:
                continue
            ob, nb = G.BFOffset(b)
            _, GB, TB = G.AssembleBEMMatrixBlock(ns, b, omega, GradM=True, TorqueAxes=axes)
            for m, D in enumerate(list(GB) + list(TB)):
                Dm = D.numpy()
                want[m] += x[oa:oa + na].conj() @ Dm @ y[ob:ob + nb] + x[ob:ob + nb].conj() @ Dm.T @ y[oa:oa + na]
        e = np.abs(s - want).max() / np.abs(want).max()
        print(f"  surface {ns}: {e:.2e}", flush=True)
        assert e <= 1e-12


@pytest.mark.parametrize("loss", ["norm", "linear"])
def test_autograd_matches_finite_differences(sx, O, loss):
    import torch
    from scuffx.autograd import DisplacedSolve
    print(f"autograd displaced solve, loss {loss}", flush=True)
    G = two_spheres(sx, O)
    N = G.TotalBFs
    omega = 0.9
    b = G.AssembleRHSVector(omega, sx.PlaneWave([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])).t
    c = torch.from_numpy(np.random.default_rng(5).standard_normal(N) + 0j).cuda()

    def L(xv):
        return (xv.abs() ** 2).sum() if loss == "norm" else (c.conj() * xv).sum().real

    theta = torch.tensor([0.02, -0.01, 0.03], dtype=torch.float64, requires_grad=True)
    x = DisplacedSolve.apply(theta, G, 1, omega, b)
    L(x).backward()
    g = theta.grad.numpy()
    h = 1e-4
    fd = np.zeros(3)
    for m in range(3):
        tp = theta.detach().clone(); tp[m] += h
        tm = theta.detach().clone(); tm[m] -= h
        fd[m] = (L(DisplacedSolve.apply(tp, G, 1, omega, b)).item() -
                 L(DisplacedSolve.apply(tm, G, 1, omega, b)).item()) / (2 * h)
    e = np.abs(g - fd).max() / np.abs(fd).max()
    print(f"  grad {g} fd {fd} rel {e:.1e}", flush=True)
    assert e <= 1e-6
