"""HDF5 export / import in the reference's layout (libs/libhmat/HDF5IO.cc),
without a GPU.

The layout is pinned independently of this project's reader: the files are
opened with the HDF5 library's own lite API through ctypes, and the
dataset shapes (HDF5 row-major: {NC, NR} / {NC, 2 NR} for matrices,
{N} / {2 N} for vectors), the raw interleaved column-major data and the
Storage_Type / RealComplex attributes are checked against
HDF5IO.cc:171-236 and :436-475.  Then the round trip through the product's
reader, and its error messages."""
import ctypes as C
import os

import numpy as np
import pytest


def h5lib():
    for d in ("", os.environ.get("SCUFFX_HDF5_DIR", ""), "/opt/conda/lib"):
        try:
            C.CDLL(os.path.join(d, "libhdf5.so") if d else "libhdf5.so", mode=C.RTLD_GLOBAL)
            return C.CDLL(os.path.join(d, "libhdf5_hl.so") if d else "libhdf5_hl.so", mode=C.RTLD_GLOBAL), \
                C.CDLL(os.path.join(d, "libhdf5.so") if d else "libhdf5.so")
        except OSError:
            continue
    pytest.skip("HDF5 libraries not installed")


def raw(path, name):
    """dims, RealComplex, Storage_Type (or None) and raw doubles of a dataset"""
    hl, h5 = h5lib()
    h5.H5open()
    h5.H5Fopen.restype = C.c_int64
    h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    f = h5.H5Fopen(path.encode(), 0, 0)
    assert f >= 0
    rank = C.c_int()
    hl.H5LTget_dataset_ndims.argtypes = [C.c_int64, C.c_char_p, C.c_void_p]
    assert hl.H5LTget_dataset_ndims(f, name.encode(), C.byref(rank)) >= 0
    dims = (C.c_ulonglong * 2)()
    cls, sz = C.c_int(), C.c_size_t()
    hl.H5LTget_dataset_info.argtypes = [C.c_int64, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p]
    assert hl.H5LTget_dataset_info(f, name.encode(), dims, C.byref(cls), C.byref(sz)) >= 0
    shape = tuple(int(dims[i]) for i in range(rank.value))
    data = np.zeros(int(np.prod(shape)))
    hl.H5LTread_dataset_double.argtypes = [C.c_int64, C.c_char_p, C.c_void_p]
    assert hl.H5LTread_dataset_double(f, name.encode(), data.ctypes.data_as(C.c_void_p)) >= 0
    attrs = {}
    hl.H5LTget_attribute_int.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_void_p]
    for a in ("RealComplex", "Storage_Type"):
        v = C.c_int(-7)
        attrs[a] = v.value if hl.H5LTget_attribute_int(f, name.encode(), a.encode(), C.byref(v)) >= 0 else None
    h5.H5Fclose.argtypes = [C.c_int64]
    h5.H5Fclose(f)
    return shape, attrs, data


def test_layout_and_round_trip(sx, tmp_path):
    h5lib()
    rng = np.random.default_rng(20261202)
    Mc = rng.standard_normal((5, 3)) + 1j * rng.standard_normal((5, 3))
    Mr = rng.standard_normal((4, 6))
    vc = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    vr = rng.standard_normal(2)
    path = str(tmp_path / "m.hdf5")
    ctx = sx.OpenHDF5Context(path)
    sx.ExportToHDF5(ctx, "M_1_A", Mc)
    sx.ExportToHDF5(ctx, "Mr", Mr)
    sx.ExportToHDF5(ctx, "KN", vc)
    sx.ExportToHDF5(ctx, "vr", vr)
    sx.CloseHDF5Context(ctx)

    shape, attrs, data = raw(path, "M_1_A")
    assert shape == (3, 10) and attrs == {"RealComplex": 1, "Storage_Type": 0}
    # column-major, interleaved re/im: row c of the dataset is column c
    assert np.array_equal(data.reshape(3, 10)[:, 0::2], Mc.real.T)
    assert np.array_equal(data.reshape(3, 10)[:, 1::2], Mc.imag.T)
    shape, attrs, data = raw(path, "Mr")
    assert shape == (6, 4) and attrs == {"RealComplex": 0, "Storage_Type": 0}
    assert np.array_equal(data.reshape(6, 4), Mr.T)
    shape, attrs, data = raw(path, "KN")
    assert shape == (14,) and attrs["RealComplex"] == 1 and attrs["Storage_Type"] is None
    assert np.array_equal(data[0::2], vc.real) and np.array_equal(data[1::2], vc.imag)
    shape, attrs, _ = raw(path, "vr")
    assert shape == (2,) and attrs["RealComplex"] == 0

    assert np.array_equal(sx.ImportFromHDF5(path, "M_1_A"), Mc)
    assert np.array_equal(sx.ImportFromHDF5(path, "Mr"), Mr)
    assert np.array_equal(sx.ImportFromHDF5(path, "KN", Vector=True), vc)
    assert np.array_equal(sx.ImportFromHDF5(path, "vr", Vector=True), vr)
    # a file of its own
    one = str(tmp_path / "one.hdf5")
    sx.ExportToHDF5(one, "RHS", vc)
    assert np.array_equal(sx.ImportFromHDF5(one, "RHS", Vector=True), vc)


def test_import_errors(sx, tmp_path):
    h5lib()
    path = str(tmp_path / "e.hdf5")
    sx.ExportToHDF5(path, "v", np.arange(3.0))
    with pytest.raises(sx.SCUFFXError, match="does not contain a dataset named M"):
        sx.ImportFromHDF5(path, "M")
    with pytest.raises(sx.SCUFFXError, match="is not a matrix"):
        sx.ImportFromHDF5(path, "v")
    with pytest.raises(sx.SCUFFXError, match="could not open file"):
        sx.ImportFromHDF5(str(tmp_path / "missing.hdf5"), "v")
