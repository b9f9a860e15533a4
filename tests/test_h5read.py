"""The test-side HDF5 reader against files written with the HDF5 C library directly (the
same calls the driver makes): names, classes, sizes, extents and values come back."""
import ctypes as C

import numpy as np
import pytest

import h5read


def _write(path, datasets):
    lib = h5read.library()
    hid = C.c_int64
    lib.H5Fcreate.restype, lib.H5Fcreate.argtypes = hid, [C.c_char_p, C.c_uint, hid, hid]
    lib.H5Screate_simple.restype = hid
    lib.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.c_void_p]
    lib.H5Dcreate2.restype, lib.H5Dcreate2.argtypes = hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]
    lib.H5Dwrite.restype, lib.H5Dwrite.argtypes = C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]
    f = lib.H5Fcreate(str(path).encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    for name, a in datasets.items():
        dims = (C.c_uint64 * a.ndim)(*a.shape)
        sp = lib.H5Screate_simple(a.ndim, dims, None)
        t = C.c_int64.in_dll(lib, "H5T_NATIVE_INT_g" if a.dtype == np.int32 else "H5T_NATIVE_DOUBLE_g").value
        ds = lib.H5Dcreate2(f, name.encode(), t, sp, 0, 0, 0)
        assert lib.H5Dwrite(ds, t, 0, 0, 0, np.ascontiguousarray(a).ctypes.data) >= 0
        lib.H5Dclose(ds)
        lib.H5Sclose(sp)
    lib.H5Fclose(f)


def test_reader_round_trip(tmp_path):
    if h5read.library() is None:
        pytest.skip("no HDF5 C library on this host")
    rng = np.random.default_rng(0)
    data = {"/id": rng.permutation(7).astype(np.int32), "/x": rng.uniform(size=7),
            "/T": rng.uniform(size=(5, 3)), "/f": rng.uniform(size=(2, 3, 9))}
    _write(tmp_path / "a.h5", data)
    got = h5read.read_root(tmp_path / "a.h5")
    assert sorted(got) == sorted(k.lstrip("/") for k in data)
    for k, a in data.items():
        arr, (cls, size, le, signed) = got[k.lstrip("/")]
        assert arr.dtype == a.dtype and arr.shape == a.shape and np.array_equal(arr, a)
        assert le and size == a.dtype.itemsize
        assert (cls, signed) == ((h5read.H5T_INTEGER, True) if a.dtype == np.int32 else (h5read.H5T_FLOAT, False))
