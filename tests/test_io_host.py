"""Host I/O either side of the kernel (no GPU): HDF5 readers (libsrt_io, replacing
snapshot-raytracer.cu:22-186), the HEALPix FITS writer (replacing chealpix
write_healpix_map at .cu:947-987, read back here by an independent Python
parser and by cfitsio), and the command line's argument and file validation
(.cu:600-713), which all happen before any device call."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "snapshot-raytracer_amd")
IOLIB = os.path.join(PKG, "libsrt_io.so")
CLI = os.path.join(PKG, "bin", "snapshot-raytracer")
CFITSIO = "/opt/conda/lib/libcfitsio.so"

pytestmark = pytest.mark.skipif(not os.path.exists(IOLIB), reason="libsrt_io.so not built (no HDF5)")


@pytest.fixture(scope="module")
def sio():
    from snapshot_raytracer_amd import snapshot_io
    return snapshot_io


@pytest.fixture(scope="module")
def iolib():
    L = ctypes.CDLL(IOLIB)
    P = ctypes.c_void_p
    for f in ("srtio_read_scalar_field", "srtio_read_vector_field"):
        getattr(L, f).argtypes = [ctypes.c_char_p, ctypes.c_char_p, P, ctypes.c_uint64, P, ctypes.c_char_p,
                                  ctypes.c_size_t]
    L.srtio_write_healpix_map.argtypes = [ctypes.c_char_p, P, ctypes.c_int64, ctypes.c_int, ctypes.c_char_p,
                                          ctypes.c_char_p, ctypes.c_size_t]
    return L


def read_field(L, fn, path, comps):
    dims = np.zeros(3, np.uint64)
    err = ctypes.create_string_buffer(256)
    rc = getattr(L, fn)(path.encode(), b"/data", None, 0, dims.ctypes.data, err, 256)
    if rc:
        return None, err.value.decode()
    out = np.empty(int(np.prod(dims)) * comps, np.float32)
    rc = getattr(L, fn)(path.encode(), b"/data", out.ctypes.data, out.size, dims.ctypes.data, err, 256)
    assert rc == 0
    return out.reshape(tuple(int(d) for d in dims) + ((comps,) if comps > 1 else ())), ""


def test_hdf5_round_trip(sio, iolib, tmp_path):
    rng = np.random.default_rng(0)
    phi = rng.standard_normal((6, 5, 4)).astype(np.float32)
    shift = rng.standard_normal((6, 5, 4, 3)).astype(np.float32)
    sio.write_potential(str(tmp_path / "p.h5"), phi)
    sio.write_shift(str(tmp_path / "b.h5"), shift)
    got, _ = read_field(iolib, "srtio_read_scalar_field", str(tmp_path / "p.h5"), 1)
    assert np.array_equal(got, phi)
    got, _ = read_field(iolib, "srtio_read_vector_field", str(tmp_path / "b.h5"), 3)
    assert np.array_equal(got, shift)
    # a double potential is converted to float by the reader (H5T_NATIVE_FLOAT, .cu:72)
    sio.write_potential(str(tmp_path / "pd.h5"), phi.astype(np.float64) * 1.000000001)
    got, _ = read_field(iolib, "srtio_read_scalar_field", str(tmp_path / "pd.h5"), 1)
    assert np.array_equal(got, (phi.astype(np.float64) * 1.000000001).astype(np.float32))


def write_h5(sio, path, arr, file_type, mem_type, chunk=None, userblock=0, external=None):
    """One rank-3 dataset /data with optional chunked layout, user block or external raw-data file, through the
    HDF5 C library."""
    L = sio._lib()
    hid = ctypes.c_int64
    L.H5Pcreate.restype = hid
    L.H5Pcreate.argtypes = [hid]
    L.H5Pset_chunk.argtypes = [hid, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    L.H5Pset_userblock.argtypes = [hid, ctypes.c_uint64]
    L.H5Pclose.argtypes = [hid]
    L.H5Pset_external.argtypes = [hid, ctypes.c_char_p, ctypes.c_int64, ctypes.c_uint64]
    fcpl = dcpl = 0
    if userblock:
        fcpl = L.H5Pcreate(sio._global("H5P_CLS_FILE_CREATE_ID_g"))
        assert L.H5Pset_userblock(fcpl, userblock) >= 0
    if chunk or external:
        dcpl = L.H5Pcreate(sio._global("H5P_CLS_DATASET_CREATE_ID_g"))
    if chunk:
        assert L.H5Pset_chunk(dcpl, 3, (ctypes.c_uint64 * 3)(*chunk)) >= 0
    if external:
        assert L.H5Pset_external(dcpl, external.encode(), 0, arr.nbytes) >= 0
    fid = L.H5Fcreate(path.encode(), 0x0002, fcpl, 0)
    assert fid >= 0
    sid = L.H5Screate_simple(3, (ctypes.c_uint64 * 3)(*arr.shape[:3]), None)
    did = L.H5Dcreate2(fid, b"/data", file_type, sid, 0, dcpl, 0)
    assert did >= 0
    assert L.H5Dwrite(did, mem_type, 0, 0, 0, np.ascontiguousarray(arr).ctypes.data_as(ctypes.c_void_p)) >= 0
    L.H5Dclose(did)
    L.H5Sclose(sid)
    L.H5Fclose(fid)
    for pl in (fcpl, dcpl):
        if pl:
            L.H5Pclose(pl)


def test_hdf5_raw_and_converted_reads(sio, iolib, tmp_path):
    """Contiguous little-endian float datasets are read straight from the file by several threads (160^3: a
    16 MB potential, a 49 MB shift); big-endian floats, chunked datasets, files with a user block and data in an
    external raw file go through H5Dread.  Every way returns the same floats."""
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((160, 160, 160)).astype(np.float32)
    shift = rng.standard_normal((160, 160, 160, 3)).astype(np.float32)
    sio.write_potential(str(tmp_path / "p.h5"), phi)
    sio.write_shift(str(tmp_path / "b.h5"), shift)
    got, _ = read_field(iolib, "srtio_read_scalar_field", str(tmp_path / "p.h5"), 1)
    assert np.array_equal(got, phi)
    got, _ = read_field(iolib, "srtio_read_vector_field", str(tmp_path / "b.h5"), 3)
    assert np.array_equal(got, shift)
    L = sio._lib()
    native, be = sio._global("H5T_NATIVE_FLOAT_g"), sio._global("H5T_IEEE_F32BE_g")
    small, small3 = phi[:24, :20, :16], shift[:24, :20, :16]
    three = (ctypes.c_uint64 * 1)(3)
    arr_native, arr_be = L.H5Tarray_create2(native, 1, three), L.H5Tarray_create2(be, 1, three)
    cases = [("be.h5", small, be, native, None, 0, 1), ("chunk.h5", small, native, native, (8, 10, 16), 0, 1),
             ("ub.h5", small, native, native, None, 512, 1), ("be3.h5", small3, arr_be, arr_native, None, 0, 3),
             ("chunk3.h5", small3, arr_native, arr_native, (12, 5, 16), 0, 3),
             ("ub3.h5", small3, arr_native, arr_native, None, 1024, 3)]
    for name, arr, ft, mt, chunk, ub, comps in cases:
        write_h5(sio, str(tmp_path / name), arr, ft, mt, chunk, ub)
    for name, arr, t, comps in (("ext.h5", small, native, 1), ("ext3.h5", small3, arr_native, 3)):
        write_h5(sio, str(tmp_path / name), arr, t, t, external=str(tmp_path / (name + ".raw")))
        cases.append((name, arr, t, t, None, 0, comps))
    for name, arr, ft, mt, chunk, ub, comps in cases:
        fn = "srtio_read_scalar_field" if comps == 1 else "srtio_read_vector_field"
        got, e = read_field(iolib, fn, str(tmp_path / name), comps)
        assert e == "" and np.array_equal(got, arr), name
    L.H5Tclose(arr_native)
    L.H5Tclose(arr_be)


def test_hdf5_reader_buffer_checks(sio, iolib, tmp_path):
    """A null buffer returns the dimensions only; a buffer smaller than the dataset is refused untouched."""
    phi = np.arange(4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    sio.write_potential(str(tmp_path / "p.h5"), phi)
    dims = np.zeros(3, np.uint64)
    err = ctypes.create_string_buffer(256)
    assert iolib.srtio_read_scalar_field(str(tmp_path / "p.h5").encode(), b"/data", None, 0, dims.ctypes.data,
                                         err, 256) == 0
    assert list(dims) == [4, 5, 6]
    out = np.full(phi.size - 1, -7.0, np.float32)
    assert iolib.srtio_read_scalar_field(str(tmp_path / "p.h5").encode(), b"/data", out.ctypes.data, out.size,
                                         dims.ctypes.data, err, 256) == 1
    assert err.value.decode() == "output buffer too small" and np.all(out == -7.0)


def test_hdf5_reader_errors(sio, iolib, tmp_path):
    _, e = read_field(iolib, "srtio_read_scalar_field", str(tmp_path / "missing.h5"), 1)
    assert e.startswith("Failed to open HDF5 file")
    phi = np.zeros((4, 4, 4), np.float32)
    sio.write_potential(str(tmp_path / "p.h5"), phi)
    # a scalar dataset is not an array[3] shift (.cu:130)
    _, e = read_field(iolib, "srtio_read_vector_field", str(tmp_path / "p.h5"), 3)
    assert e.startswith("Dataset is not an array type")
    # per-voxel arrays of length 2 (.cu:140)
    L = sio._lib()
    two = (ctypes.c_uint64 * 1)(2)
    t = L.H5Tarray_create2(sio._global("H5T_NATIVE_FLOAT_g"), 1, two)
    arr = np.zeros((4, 4, 4, 2), np.float32)
    sio._write(str(tmp_path / "b2.h5"), arr, t, t, (4, 4, 4))
    L.H5Tclose(t)
    _, e = read_field(iolib, "srtio_read_vector_field", str(tmp_path / "b2.h5"), 3)
    assert e.startswith("Expected per-voxel array length 3")


def test_fits_writer_read_back(iolib, sio, tmp_path):
    nside = 8
    m = np.random.default_rng(1).standard_normal(12 * nside * nside).astype(np.float32)
    m[5] = 0.0
    m[6] = -0.0
    m[7] = 1e-30
    path = str(tmp_path / "m.fits")
    err = ctypes.create_string_buffer(256)
    assert iolib.srtio_write_healpix_map(path.encode(), m.ctypes.data, nside, 0, b"C", err, 256) == 0
    assert os.path.getsize(path) % 2880 == 0
    got, hdr = sio.read_healpix_fits(path)
    assert got.tobytes() == m.tobytes()
    assert hdr["PIXTYPE"] == "HEALPIX" and hdr["ORDERING"] == "RING" and hdr["NSIDE"] == nside
    assert hdr["COORDSYS"] == "C" and hdr["FIRSTPIX"] == 0 and hdr["LASTPIX"] == 12 * nside * nside - 1
    assert hdr["NAXIS2"] == 12 * nside * nside and hdr["TFORM1"] == "1E"


@pytest.mark.skipif(not os.path.exists(CFITSIO), reason="cfitsio not present")
def test_fits_writer_read_by_cfitsio(iolib, tmp_path):
    nside = 4
    m = np.arange(12 * nside * nside, dtype=np.float32) * 0.25 - 7
    path = str(tmp_path / "c.fits")
    assert iolib.srtio_write_healpix_map(path.encode(), m.ctypes.data, nside, 0, b"C", None, 0) == 0
    F = ctypes.CDLL(CFITSIO)
    fptr = ctypes.c_void_p()
    status = ctypes.c_int(0)
    F.ffopen(ctypes.byref(fptr), path.encode(), 0, ctypes.byref(status))
    assert status.value == 0
    hdutype = ctypes.c_int()
    F.ffmahd(fptr, 2, ctypes.byref(hdutype), ctypes.byref(status))
    assert status.value == 0 and hdutype.value == 2  # BINARY_TBL
    val = ctypes.c_long()
    F.ffgkyj(fptr, b"NSIDE", ctypes.byref(val), None, ctypes.byref(status))
    assert val.value == nside
    s = ctypes.create_string_buffer(80)
    F.ffgkys(fptr, b"ORDERING", s, None, ctypes.byref(status))
    assert s.value == b"RING"
    out = np.zeros(m.size, np.float32)
    anynul = ctypes.c_int()
    F.ffgcv.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong,
                        ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(ctypes.c_int)]
    F.ffgcv(fptr, 42, 1, 1, 1, m.size, None, out.ctypes.data, ctypes.byref(anynul), ctypes.byref(status))
    assert status.value == 0
    F.ffclos(fptr, ctypes.byref(status))
    assert np.array_equal(out, m)


def run_cli(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60,
                          env=dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", "")))


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_argument_errors(sio, tmp_path):
    r = run_cli([], tmp_path)
    assert r.returncode == 255 and "Usage:" in r.stderr
    for args, msg in ((["a", "b", "0", "4"], "Invalid Nside value"), (["a", "b", "x", "4"], "Invalid Nside value"),
                      (["a", "b", "4", "0"], "Invalid n_steps value"), (["a", "b", "4", "-3"], "Invalid n_steps"),
                      (["a", "b", "4", "4", "-1"], "Invalid batch value"),
                      (["missing.h5", "b", "4", "4"], "Failed to open HDF5 file")):
        r = run_cli(args, tmp_path)
        assert r.returncode == 255, args
        assert msg in r.stderr, (args, r.stderr)


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_file_errors(sio, tmp_path):
    n = 8
    sio.write_potential(str(tmp_path / "p.h5"), np.zeros((n, n, n), np.float32))
    sio.write_shift(str(tmp_path / "b.h5"), np.zeros((n, n, n, 3), np.float32))
    sio.write_shift(str(tmp_path / "b6.h5"), np.zeros((n, n, 6, 3), np.float32))
    r = run_cli(["p.h5", "p.h5", "4", "4"], tmp_path)
    assert r.returncode == 255 and "not an array type" in r.stderr
    r = run_cli(["p.h5", "b6.h5", "4", "4"], tmp_path)
    assert r.returncode == 255 and "dimensions differ" in r.stderr
    r = run_cli(["p.h5", "b.h5", "4", "4", "12"], tmp_path)  # .cu:713
    assert r.returncode == 255 and "Batch exceeds total number of pixels" in r.stderr
    assert "Data dimensions: 8 x 8 x 8" in r.stdout
