"""Snapshot and map files of the ray tracer, from Python.

* HDF5 writers for the two inputs the command line reads (dataset /data):
  the potential as a 3-D float (or double) array, the shift as a 3-D array
  whose element type is H5T_ARRAY{3} of float (snapshot-raytracer.cu:89-186
  reads exactly that).  They call the HDF5 C library through ctypes (h5py is
  not needed).
* A reader for the HEALPix FITS maps the command line writes: the header
  cards and the float32 column of the binary table.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HDF5_LIB = os.environ.get("SRT_HDF5_LIB", "/opt/conda/lib/libhdf5.so")

_h5 = None


def hdf5_available() -> bool:
    return os.path.exists(HDF5_LIB)


def _lib():
    global _h5
    if _h5 is None:
        L = ctypes.CDLL(HDF5_LIB)
        hid = ctypes.c_int64
        L.H5open.restype = ctypes.c_int
        L.H5Fcreate.restype = hid
        L.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid, hid]
        L.H5Screate_simple.restype = hid
        L.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]
        L.H5Tarray_create2.restype = hid
        L.H5Tarray_create2.argtypes = [hid, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint64)]
        L.H5Dcreate2.restype = hid
        L.H5Dcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]
        L.H5Dwrite.restype = ctypes.c_int
        L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Tclose", "H5Fclose"):
            getattr(L, f).argtypes = [hid]
            getattr(L, f).restype = ctypes.c_int
        if L.H5open() < 0:
            raise RuntimeError("H5open failed")
        _h5 = L
    return _h5


def _global(name: str) -> int:
    return ctypes.c_int64.in_dll(_lib(), name).value


def _write(path: str, arr: np.ndarray, file_type: int, mem_type: int, dims) -> None:
    L = _lib()
    fid = L.H5Fcreate(path.encode(), 0x0002, 0, 0)  # H5F_ACC_TRUNC, default property lists
    if fid < 0:
        raise OSError(f"cannot create {path}")
    d = (ctypes.c_uint64 * 3)(*dims)
    sid = L.H5Screate_simple(3, d, None)
    did = L.H5Dcreate2(fid, b"/data", file_type, sid, 0, 0, 0)
    rc = L.H5Dwrite(did, mem_type, 0, 0, 0, arr.ctypes.data_as(ctypes.c_void_p)) if did >= 0 else -1
    if did >= 0:
        L.H5Dclose(did)
    L.H5Sclose(sid)
    L.H5Fclose(fid)
    if rc < 0:
        raise OSError(f"writing {path} failed")


def write_potential(path: str, phi: np.ndarray) -> None:
    """phi: (nz, ny, nx) float32 or float64 (stored in that precision)."""
    arr = np.ascontiguousarray(phi)
    if arr.ndim != 3:
        raise ValueError("potential must be 3-D")
    if arr.dtype == np.float64:
        t = _global("H5T_NATIVE_DOUBLE_g")
        _write(path, arr, t, t, arr.shape)
    else:
        arr = arr.astype(np.float32, copy=False)
        t = _global("H5T_NATIVE_FLOAT_g")
        _write(path, arr, t, t, arr.shape)


def write_shift(path: str, shift_xyz: np.ndarray) -> None:
    """shift_xyz: (nz, ny, nx, 3) float32, interleaved (B_x, B_y, B_z) per voxel."""
    arr = np.ascontiguousarray(shift_xyz, dtype=np.float32)
    if arr.ndim != 4 or arr.shape[3] != 3:
        raise ValueError("shift must be (nz, ny, nx, 3)")
    L = _lib()
    three = (ctypes.c_uint64 * 1)(3)
    t = L.H5Tarray_create2(_global("H5T_NATIVE_FLOAT_g"), 1, three)
    try:
        _write(path, arr, t, t, arr.shape[:3])
    finally:
        L.H5Tclose(t)


def read_healpix_fits(path: str):
    """(map float32 array, header dict of the table extension)."""
    with open(path, "rb") as f:
        data = f.read()

    def parse(off):
        hdr = {}
        while True:
            block = data[off:off + 2880]
            if len(block) < 2880:
                raise ValueError("truncated FITS header")
            off += 2880
            done = False
            for i in range(36):
                card = block[80 * i:80 * i + 80].decode("ascii")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    val = card[10:].split(" / ")[0].strip()
                    if val.startswith("'"):
                        val = val[1:val.rindex("'")].rstrip()
                    elif val in ("T", "F"):
                        val = val == "T"
                    else:
                        val = int(val) if val.lstrip("-").isdigit() else float(val)
                    hdr[key] = val
            if done:
                return hdr, off

    primary, off = parse(0)
    assert primary.get("SIMPLE") is True
    ext, off = parse(off)
    if ext.get("XTENSION") != "BINTABLE" or ext.get("TFORM1") != "1E":
        raise ValueError("not a one-column float32 binary table")
    rows = ext["NAXIS2"]
    m = np.frombuffer(data, dtype=">f4", count=rows, offset=off).astype(np.float32)
    return m, ext
