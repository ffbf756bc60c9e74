"""HDF5 dumps of app/mhdjet (mhdjet.cpp:25-89): written by the library's
serial writer and read back through the HDF5 C API, on the CPU.  The file
layout is fixed by the reference: one dataset "/3DArray" of native doubles,
dimensions (N_x, N_y, N_z), element [i][j][k] = node (i, j, k)."""
import ctypes as C

import numpy as np
import pytest


def _h5_available():
    try:
        C.CDLL("libhdf5.so")
        return True
    except OSError:
        try:
            C.CDLL("/opt/conda/lib/libhdf5.so")
            return True
        except OSError:
            return False


pytestmark = pytest.mark.skipif(not _h5_available(), reason="no HDF5 C library in this environment")


def test_hdf5_round_trip_and_layout(tmp_path):
    import novaclone_amd as nv
    rng = np.random.default_rng(1)
    a = rng.standard_normal((6, 5, 4))
    f = tmp_path / "Ux0.000000.h5"
    nv.hdf5_write_host(f, a)
    assert f.read_bytes()[:8] == b"\x89HDF\r\n\x1a\n"  # the HDF5 signature
    back = nv.hdf5_read_host(f)
    assert back.shape == (6, 5, 4) and np.array_equal(back, a)


def test_hdf5_walled_extent_is_the_cell_count(tmp_path):
    # a walled direction stores N + 1 nodes; the reference's dataset holds N
    import novaclone_amd as nv
    a = np.arange(6 * 5 * 4, dtype=np.float64).reshape(6, 5, 4)
    f = tmp_path / "p_0.000000.h5"
    nv.hdf5_write_host(f, a, dims=(6, 4, 4))
    back = nv.hdf5_read_host(f)
    assert back.shape == (6, 4, 4) and np.array_equal(back, a[:, :4, :])


def test_hdf5_file_read_by_the_c_api_directly(tmp_path):
    # an independent reader: the HDF5 C library through ctypes (H5Fopen,
    # H5Dopen2, H5Dread), so the file is a real HDF5 file, not merely
    # readable by the library that wrote it
    import novaclone_amd as nv
    try:
        h5 = C.CDLL("libhdf5.so")
    except OSError:
        h5 = C.CDLL("/opt/conda/lib/libhdf5.so")
    a = np.random.default_rng(2).standard_normal((4, 3, 8))
    f = tmp_path / "Bx_0.020000.h5"
    nv.hdf5_write_host(f, a)
    h5.H5open()
    h5.H5Fopen.restype = C.c_int64
    h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    h5.H5Dopen2.restype = C.c_int64
    h5.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    h5.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    h5.H5Dclose.argtypes = [C.c_int64]
    h5.H5Fclose.argtypes = [C.c_int64]
    native = C.c_int64.in_dll(h5, "H5T_NATIVE_DOUBLE_g").value
    fid = h5.H5Fopen(str(f).encode(), 0, 0)
    assert fid >= 0
    did = h5.H5Dopen2(fid, b"/3DArray", 0)
    assert did >= 0
    out = np.empty_like(a)
    assert h5.H5Dread(did, native, 0, 0, 0, out.ctypes.data) >= 0
    h5.H5Dclose(did)
    h5.H5Fclose(fid)
    assert np.array_equal(out, a)
