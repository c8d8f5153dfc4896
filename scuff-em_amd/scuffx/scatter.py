"""BEM matrices of a geometry under the entries of a transformation file,
as the scattering caller builds them (applications/scuff-scatter/
scuff-scatter.cc:345-451, :565).

With several entries the diagonal (T) blocks are assembled once per
frequency at the untransformed positions (a surface identical to an
earlier one reuses its block, RWGGeometry Mate), and for each entry the
geometry is transformed, the off-diagonal (U) blocks are assembled and
stamped with their transposes below the diagonal.  A single entry is
applied before one full assembly.

Two places where this deliberately differs from scuff-scatter:
  * the reference stamps the conjugate transpose of each U block below the
    diagonal (InsertBlockAdjoint, scuff-scatter.cc:440, HMatrix.cc:638-653);
    at a real frequency that is not the matrix its own full assembly
    produces (which fills the lower triangle with the plain transpose,
    AssembleBEMMatrix.cc:753-758).  Here the lower blocks are transposes,
    so every entry's matrix equals the full assembly at the transformed
    positions up to the T blocks' rounding.  At imaginary frequencies the
    blocks are real and the two agree.
  * with exactly one entry the reference assembles the full matrix before
    applying it (scuff-scatter.cc:401-402 precede :428), so the matrix
    belongs to the untransformed geometry while the incident field and
    power see the transformed one; here the entry is applied first.

Every block is assembled on the GPU through the C ABI; this module moves
device blocks into place.
"""
from . import HMatrix


def TransformedBEMMatrices(G, Omega, GTCs, M=None):
    """Yields (GTC, M) for every entry of GTCs with G transformed by that
    entry while the consumer holds it; G is untransformed before the next
    entry and when the generator is closed."""
    G.CheckGTCList(GTCs)
    N = G.TotalBFs
    NS = G.NumSurfaces
    if M is None or M.NR != N or M.NC != N:
        M = HMatrix(N, N)
    if len(GTCs) == 1:
        G.Transform(GTCs[0])
        try:
            G.AssembleBEMMatrix(Omega, M)
            yield GTCs[0], M
        finally:
            G.UnTransform()
        return
    mate = G.Mate
    offs = [G.BFOffset(ns) for ns in range(NS)]
    TBlocks = {}
    for ns in range(NS):
        TBlocks[ns] = TBlocks[mate[ns]] if mate[ns] >= 0 else G.AssembleBEMMatrixBlock(ns, ns, Omega)
    for GTC in GTCs:
        G.Transform(GTC)
        try:
            M.t.zero_()
            M.factored = False
            for ns, (o, n) in enumerate(offs):
                M.t[o:o + n, o:o + n].copy_(TBlocks[ns].t)
                for nsp in range(ns + 1, NS):
                    op, nq = offs[nsp]
                    U = G.AssembleBEMMatrixBlock(ns, nsp, Omega)
                    M.t[op:op + nq, o:o + n].copy_(U.t)                    # rows ns, columns nsp
                    M.t[o:o + n, op:op + nq].copy_(U.t.transpose(0, 1))    # rows nsp, columns ns
            yield GTC, M
        finally:
            G.UnTransform()
