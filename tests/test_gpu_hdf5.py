"""HMatrix / HVector HDF5 export from GPU memory, as the scattering caller
writes M, RHS and KN (scuff-scatter.cc:456-457, :497-500), read back through
the HDF5 layout (HDF5IO.cc) bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_export_scatter_datasets(sx, O, tmp_path):
    import ctypes
    try:
        ctypes.CDLL("/opt/conda/lib/libhdf5.so")
    except OSError:
        pytest.skip("HDF5 libraries not installed")
    print("hdf5 export of M, RHS, KN", flush=True)
    V, T = O.icosphere(1)
    G = sx.RWGGeometry()
    G.AddMesh(V, T)
    M = G.AssembleBEMMatrix(1.0)
    Mh = M.numpy()
    RHS = G.AssembleRHSVector(1.0, sx.PlaneWave([1, 0, 0], [0, 0, 1]))
    path = str(tmp_path / "scatter.hdf5")
    ctx = sx.OpenHDF5Context(path)
    M.ExportToHDF5(ctx, "M_1")
    RHS.ExportToHDF5(ctx, "RHS_1")
    M.LUFactorize()
    KN = sx.HVector(RHS.N, data=RHS.t.clone())
    M.LUSolve(KN)
    KN.ExportToHDF5(ctx, "KN_1")
    sx.CloseHDF5Context(ctx)
    assert np.array_equal(sx.ImportFromHDF5(path, "M_1"), Mh)
    assert np.array_equal(sx.HMatrix.FromHDF5(path, "M_1").numpy(), Mh)
    assert np.array_equal(sx.HVector.FromHDF5(path, "RHS_1").numpy(), RHS.numpy())
    assert np.array_equal(sx.ImportFromHDF5(path, "KN_1", Vector=True), KN.numpy())
