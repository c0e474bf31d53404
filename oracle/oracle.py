"""ctypes binding of the host oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``, always as the reference results, never as the
thing measured for the GPU path or shipped with it.

The oracle restates snapshot-raytracer.cu:195-583 on the host (see
srt_oracle.hpp for the line-by-line map).  ``FORM_FP64`` is the accuracy
oracle; ``FORM_REF`` reproduces the reference GPU code's float/double mix.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FORM_FP64 = 0
FORM_REF = 1

_lib = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def bind(path: str):
    """Load an oracle library built from this directory's sources and declare its entry points.

    lib() binds the default build; tests that compile a variant of
    srt_oracle.hpp (a planted mistake, to show a check is sensitive to it)
    bind that variant with this and pass it as ``library=``.
    """
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    L.orc_pix2zphi_ring.argtypes = [i64, P, i64, P, P, P]
    L.orc_tricubic.argtypes = [ctypes.c_int, P, P, i64, i64, i64, P, i64, P]
    L.orc_rhs.argtypes = [ctypes.c_int, P, P, P, P, P, i64, i64, i64, P, i64, P]
    L.orc_trace.argtypes = [ctypes.c_int, P, P, P, P, P, i64, i64, i64, i64, P, i64,
                            ctypes.c_int, P, P, P, ctypes.c_int, P]
    L.orc_f32_to_f16_array.argtypes = [P, P, i64]
    L.orc_f16_to_f32_array.argtypes = [P, P, i64]
    L.orc_f16_check_range.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.orc_f16_check_range.restype = ctypes.c_uint64
    L.orc_max_threads.restype = ctypes.c_int
    if hasattr(L, "orc_count_ops"):
        L.orc_count_ops.argtypes = [P, P, P, P, i64, i64, i64, i64, i64, ctypes.c_int, P, P, ctypes.c_int]
        L.orc_tricubic_weights.argtypes = [P, i64, i64, i64, P, i64, P]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = bind(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def max_threads() -> int:
    return lib().orc_max_threads()


def pix2zphi_ring(nside: int, pix) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    pix = np.ascontiguousarray(pix, dtype=np.int64)
    z = np.empty(pix.shape, np.float64)
    s = np.empty_like(z)
    phi = np.empty_like(z)
    rc = lib().orc_pix2zphi_ring(nside, _p(pix), pix.size, _p(z), _p(s), _p(phi))
    if rc:
        raise ValueError("invalid nside or pixel index")
    return z, s, phi


def tricubic(grid, pos, form=FORM_FP64, grid16=None) -> np.ndarray:
    """Tricubic value/gradient/Hessian at positions pos (n x 3, grid units).

    grid is (nz, ny, nx) float32, x fastest; grid16 (same shape, uint16) holds
    B * 2^27 as binary16 instead.  Returns n x 10 = value, grad[3],
    Hessian (xx, xy, xz, yy, yz, zz).
    """
    ref = grid16 if grid16 is not None else grid
    nz, ny, nx = ref.shape
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    out = np.empty((pos.shape[0], 10), np.float64)
    g32 = None if grid16 is not None else _f32(grid)
    g16 = None if grid16 is None else np.ascontiguousarray(grid16, dtype=np.uint16)
    lib().orc_tricubic(form, _p(g32), _p(g16), nx, ny, nz, _p(pos), pos.shape[0], _p(out))
    return out


class Fields:
    """Host copies of the potential and the three shift grids, x fastest.

    phi: (nz, ny, nx) float32.  shift: (3, nz, ny, nx) float32 planar
    components (B_x, B_y, B_z).  b16: optional (3, nz, ny, nx) uint16
    binary16 of B * 2^27 (the reference's device storage); when given, the
    oracle reads the shift from it.
    """

    def __init__(self, phi, shift, b16=None):
        self.phi = _f32(phi)
        self.shift = _f32(shift)
        assert self.shift.shape[0] == 3 and self.shift.shape[1:] == self.phi.shape
        self.b16 = None if b16 is None else np.ascontiguousarray(b16, dtype=np.uint16)
        self.nz, self.ny, self.nx = self.phi.shape

    def ptrs(self):
        return (_p(self.phi), _p(self.shift[0]), _p(self.shift[1]), _p(self.shift[2]), _p(self.b16),
                self.nx, self.ny, self.nz)


def rhs(fields: Fields, y, form=FORM_FP64) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, 19)
    d = np.empty_like(y)
    lib().orc_rhs(form, *fields.ptrs(), _p(y), y.shape[0], _p(d))
    return d


def trace(fields: Fields, nside: int, pix, n_steps: int, observer=None, form=FORM_FP64,
          nthreads: int = 0, return_state: bool = False, library=None):
    """Trace the rays of RING pixels `pix`; returns (maps (6, npix) float32, seconds[, state]).

    library: an oracle build bound with bind() (default: lib())."""
    pix = np.ascontiguousarray(pix, dtype=np.int64)
    if observer is None:
        observer = (fields.nx * 0.5, fields.ny * 0.5, fields.nz * 0.5)
    obs = np.ascontiguousarray(observer, dtype=np.float64)
    out = np.empty((6, pix.size), np.float32)
    state = np.empty((pix.size, 19), np.float64) if return_state else None
    sec = np.zeros(1, np.float64)
    rc = (library or lib()).orc_trace(form, *fields.ptrs(), nside, _p(pix), pix.size, n_steps, _p(obs),
                                      _p(out), _p(state), nthreads, _p(sec))
    if rc:
        raise ValueError("invalid trace arguments")
    if return_state:
        return out, float(sec[0]), state
    return out, float(sec[0])


def count_ops(fields: Fields, nside: int, pix: int, n_steps: int, observer=None, variant: int = 0) -> dict:
    """Floating-point operations per ray-step of the FP64 oracle algorithm.

    variant 0: the reference's interpolation form (three expanded Catmull-Rom
    polynomials per row and pass); 1: the work-efficient weights form of the
    same interpolant (opcount.cpp, WeightsTricubic); 2: the weights form with
    the weight sets shared between the four fields of a right-hand side."""
    if observer is None:
        observer = (fields.nx * 0.5, fields.ny * 0.5, fields.nz * 0.5)
    obs = np.ascontiguousarray(observer, dtype=np.float64)
    c = np.zeros(4, np.float64)
    p = fields.ptrs()
    rc = lib().orc_count_ops(p[0], p[1], p[2], p[3], fields.nx, fields.ny, fields.nz, nside, pix, n_steps,
                             _p(obs), _p(c), variant)
    if rc:
        raise ValueError("invalid count_ops arguments")
    return {"add": c[0], "mul": c[1], "div": c[2], "sqrt": c[3], "total": float(c.sum())}


def tricubic_weights(grid, pos) -> np.ndarray:
    """The weights-form interpolation (FP64) at positions pos (n x 3); n x 10 like tricubic()."""
    nz, ny, nx = grid.shape
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    out = np.empty((pos.shape[0], 10), np.float64)
    lib().orc_tricubic_weights(_p(_f32(grid)), nx, ny, nz, _p(pos), pos.shape[0], _p(out))
    return out


def f32_to_f16(a) -> np.ndarray:
    a = _f32(a)
    out = np.empty(a.shape, np.uint16)
    lib().orc_f32_to_f16_array(_p(a), _p(out), a.size)
    return out


def f16_to_f32(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint16)
    out = np.empty(a.shape, np.float32)
    lib().orc_f16_to_f32_array(_p(a), _p(out), a.size)
    return out


def f16_check_range(lo: int, hi: int) -> int:
    """Mismatches of the binary16 RNE conversion against x86 F16C over bit patterns [lo, hi)."""
    return int(lib().orc_f16_check_range(lo, hi))


def shift_to_f16(shift) -> np.ndarray:
    """The reference's device shift format: binary16(B * 2^27) per component (.cu:585-597)."""
    return f32_to_f16(_f32(shift) * np.float32(134217728.0))
