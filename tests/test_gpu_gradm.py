"""GPU parity of RWGGeometry::AssembleBEMMatrixBlock with GradM
(libscuff.h:500-505): the block of two distinct surfaces and its three
displacement derivatives against the oracle, relative Frobenius 1e-12,
in both block orientations, for PEC, dielectric and mixed pairs, with
and without near panel pairs between the surfaces."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def pair(sx, O, level, mats, d):
    V, T = O.icosphere(level)
    G = sx.RWGGeometry()
    Go = O.Geometry()
    for i, (name, kind, eps) in enumerate(mats):
        Vi = V + (np.asarray(d) if i else 0.0)
        G.AddMesh(Vi, T, name)
        Go.add_surface(Vi, T, kind=kind, eps=eps)
    return G, Go


PEC = ("PEC", 0, 1.0)
DIEL = ("CONST_EPS_4+0.1i", 1, 4 + 0.1j)


@pytest.mark.parametrize("level,mats,d,omega", [
    (1, (PEC, PEC), (0.0, 0.0, 3.0), 1.0),
    (1, (DIEL, DIEL), (0.2, 0.1, 2.6), 0.7),      # near pairs across the gap
    (1, (PEC, DIEL), (0.0, 2.8, 0.0), 0.9),
    (0, (DIEL, PEC), (0.3, -0.2, 2.5), 0.5j),
])
@pytest.mark.parametrize("order", [(0, 1), (1, 0)])
def test_block_and_gradm_match_oracle(sx, O, level, mats, d, omega, order):
    nsa, nsb = order
    print(f"gradm level={level} {mats[0][0]}/{mats[1][0]} order={order} omega={omega}", flush=True)
    G, Go = pair(sx, O, level, mats, d)
    B, GB = G.AssembleBEMMatrixBlock(nsa, nsb, omega, GradM=True)
    Bo, GBo, _ = Go.assemble_block_grad(omega, nsa, nsb)
    eb = rel_fro(B.numpy(), Bo)
    eg = [rel_fro(GB[m].numpy(), GBo[m]) for m in range(3)]
    print(f"  block {eb:.2e} grad {max(eg):.2e}", flush=True)
    assert max(eg) <= 1e-12
    if nsa < nsb:
        assert eb <= 1e-12
    else:
        # the (1,0) block is the transpose of the (0,1) block, as the
        # lower blocks of the full matrix are (AssembleBEMMatrix.cc:753-757);
        # a direct (1,0) evaluation differs from that only by the adaptive
        # FIPPI tolerance of the near pairs taken in the other order
        Bt, _, _ = Go.assemble_block_grad(omega, nsb, nsa)
        assert rel_fro(B.numpy(), Bt.T) <= 1e-12
        assert eb <= 1e-8


def test_block_matches_full_matrix(sx, O):
    """The block equals the corresponding part of AssembleBEMMatrix."""
    G, _ = pair(sx, O, 1, (DIEL, PEC), (0.0, 0.0, 2.7))
    M = G.AssembleBEMMatrix(1.0).numpy()
    o0, n0 = G.BFOffset(0)
    o1, n1 = G.BFOffset(1)
    B01 = G.AssembleBEMMatrixBlock(0, 1, 1.0).numpy()
    B10 = G.AssembleBEMMatrixBlock(1, 0, 1.0).numpy()
    assert np.array_equal(B01, M[o0:o0 + n0, o1:o1 + n1])
    assert np.array_equal(B10, M[o1:o1 + n1, o0:o0 + n0])
    # diagonal (T) blocks, self terms and near pairs included
    for ns, (o, n) in enumerate([(o0, n0), (o1, n1)]):
        T = G.AssembleBEMMatrixBlock(ns, ns, 1.0).numpy()
        assert np.array_equal(T, M[o:o + n, o:o + n])


def test_block_rejects_same_surface_gradient(sx, O):
    G, _ = pair(sx, O, 0, (PEC, PEC), (0.0, 0.0, 3.0))
    with pytest.raises(sx.SCUFFXError):
        G.AssembleBEMMatrixBlock(1, 1, 1.0, GradM=True)


@pytest.mark.parametrize("mats", [(PEC, DIEL), (DIEL, DIEL)])
def test_torque_derivatives_match_oracle(sx, O, mats):
    """dMdT: derivatives with respect to rotations of the row surface about
    three axes through the origin."""
    axes = [(0, 0, 1), (1, 1, 0), (0.3, -0.5, 0.8)]
    G, Go = pair(sx, O, 1, mats, (0.2, 0.1, 2.6))
    B, GB, TB = G.AssembleBEMMatrixBlock(0, 1, 0.8, GradM=True, TorqueAxes=axes)
    Bo, GBo, TBo = Go.assemble_block_grad(0.8, 0, 1, gammas=[sx.CreateGammaMatrix(a) for a in axes])
    assert rel_fro(B.numpy(), Bo) <= 1e-12
    for m in range(3):
        assert rel_fro(GB[m].numpy(), GBo[m]) <= 1e-12
    for t in range(3):
        e = rel_fro(TB[t].numpy(), TBo[t])
        print(f"  torque axis {axes[t]}: {e:.2e}", flush=True)
        assert e <= 1e-12
