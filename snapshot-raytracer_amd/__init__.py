"""MI355X snapshot ray tracer: Python binding of libsrt.so (include/srt.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  This
module binds it with ctypes and mirrors the reference's host flow
(snapshot-raytracer.cu:600-998): upload a snapshot, trace the rays of a RING
batch (or an explicit pixel list), get the six maps.  There is no CPU
fallback: every call goes to the HIP library and fails loudly if it is not
built or no GPU is present.

Import name: ``snapshot_raytracer_amd`` (the directory is
``snapshot-raytracer_amd``; ``load_package()`` in the repository root
modules registers it).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrt.so")

PREC_FP64 = 0
PREC_MIXED = 1
PREC_REF = 2
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED, "ref": PREC_REF}

SHIFT_F32 = 0
SHIFT_F16_SCALED = 1

ORDER_LINEAR = 0
ORDER_NESTED = 1

SCHEDULE_PERSISTENT = 0
SCHEDULE_DISPATCH = 1

LAYOUT_TAPS = 0
LAYOUT_BICUBIC = 1

MAP_NAMES = ("dAmap", "ellipticity1map", "ellipticity2map", "rotationmap", "deflection1map", "deflection2map")

# C-ABI functions declared in include/srt.h
EXPORTS = ("srt_create", "srt_destroy", "srt_strerror", "srt_last_error", "srt_upload", "srt_set_layout",
           "srt_get_layout", "srt_trace",
           "srt_trace_host", "srt_trace_adjoint", "srt_plan_create", "srt_plan_destroy", "srt_plan_set_order", "srt_plan_configure",
           "srt_plan_get_config", "srt_plan_set_schedule", "srt_plan_get_schedule", "srt_plan_autotune", "srt_plan_autotune_brick", "srt_plan_launch", "srt_plan_last_ms", "srt_last_trace_ms",
           "srt_grid_info", "srt_version", "srt_effective_order", "srt_trace_noalloc")


class BatchItem(ctypes.Structure):
    """srt_batch_item of include/srt.h."""
    _fields_ = [("ctx", ctypes.c_void_p), ("nside", ctypes.c_int64), ("pix_begin", ctypes.c_int64),
                ("count", ctypes.c_int64), ("n_steps", ctypes.c_int), ("observer_set", ctypes.c_int),
                ("observer", ctypes.c_double * 3), ("out", ctypes.c_void_p)]

_lib = None


class SrtError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: status {status}")


def lib():
    """Load libsrt.so (built by `make -C snapshot-raytracer_amd`).

    torch, when installed, is imported first: it carries its own HIP runtime,
    and libsrt.so then binds to that one instead of loading the system's.  Two
    HIP runtimes in one process do not share the GPU: whichever initialises
    second finds no device.
    """
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                               "or make -C snapshot-raytracer_amd")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        i32 = ctypes.c_int
        L.srt_create.argtypes = [i32, ctypes.POINTER(P)]
        L.srt_destroy.argtypes = [P]
        L.srt_strerror.argtypes = [i32]
        L.srt_strerror.restype = ctypes.c_char_p
        L.srt_last_error.restype = ctypes.c_char_p
        L.srt_version.restype = ctypes.c_char_p
        L.srt_upload.argtypes = [P, P, P, i64, i64, i64, i32, i32]
        L.srt_set_layout.argtypes = [P, i32, i32]
        L.srt_get_layout.argtypes = [P, P, P, P]
        L.srt_trace.argtypes = [P, i64, i64, i64, P, i32, P, i32, i32, P, P, P]
        L.srt_trace_host.argtypes = [P, i64, i64, i64, P, i32, P, i32, i32, P, P]
        L.srt_trace_noalloc.argtypes = [P, i64, i64, i64, P, i32, P, i32, i32, P, P, P]
        L.srt_trace_adjoint.argtypes = [P, i64, i64, i64, P, i32, P, i32, P, P, P, P, P]
        L.srt_plan_create.argtypes = [P, i32, i32, ctypes.POINTER(P)]
        L.srt_plan_destroy.argtypes = [P]
        L.srt_plan_configure.argtypes = [P, i32, i32]
        L.srt_plan_set_order.argtypes = [P, i32, i32]
        L.srt_plan_get_config.argtypes = [P, P, P, P]
        L.srt_plan_autotune.argtypes = [P, P, P, P]
        L.srt_plan_set_schedule.argtypes = [P, i32]
        L.srt_plan_get_schedule.argtypes = [P, P]
        L.srt_plan_autotune_brick.argtypes = [P, P, P]
        L.srt_plan_launch.argtypes = [P, i32, P]
        L.srt_plan_last_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
        L.srt_last_trace_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
        L.srt_grid_info.argtypes = [P, P, P, P, P]
        L.srt_effective_order.argtypes = [i64, i64, i64, i32, i32, ctypes.POINTER(i32)]
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        L = lib()
        msg = L.srt_last_error().decode(errors="replace")
        raise SrtError(rc, f"{what} failed: {L.srt_strerror(rc).decode()} ({msg})")


def batch_range(nside: int, batch: int | None) -> tuple[int, int]:
    """(first RING pixel, pixel count) of a batch, as .cu:708-717 defines it.

    batch None (or < 0) is the full sky.  Batch b covers RING pixels
    [b Nside^2, (b + 1) Nside^2): a range of rings, not a HEALPix base pixel.
    """
    npix = 12 * nside * nside
    if batch is None or batch < 0:
        return 0, npix
    start = batch * nside * nside
    if start >= npix:
        raise ValueError("Batch exceeds total number of pixels")
    return start, nside * nside


def effective_order(nside: int, pix_begin: int, count: int, has_pix_list: bool = False,
                    order: int = ORDER_NESTED) -> int:
    """The lane order a trace runs for a request (srt_effective_order; host only, no GPU)."""
    v = ctypes.c_int()
    _check(lib().srt_effective_order(nside, pix_begin, count, 1 if has_pix_list else 0, order, ctypes.byref(v)),
           "srt_effective_order")
    return v.value


def _ptr(t) -> ctypes.c_void_p:
    """Address of a numpy array or torch tensor (host or device)."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(t.data_ptr())


class Tracer:
    """One library context on one GPU holding one uploaded snapshot."""

    def __init__(self, device: int = 0):
        self._ctx = ctypes.c_void_p()
        _check(lib().srt_create(device, ctypes.byref(self._ctx)), "srt_create")
        self.device = device
        self.n = None
        self.uploads = 0  # snapshots uploaded so far (ray_maps uses it to notice a replaced snapshot)

    def close(self) -> None:
        if self._ctx:
            lib().srt_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, phi, shift_xyz, shift_format: int = SHIFT_F32) -> None:
        """phi: (n, n, n) float32; shift_xyz: (n, n, n, 3) float32 interleaved.

        numpy arrays are host inputs; torch tensors on the GPU are device inputs.
        """
        on_dev = 0
        if isinstance(phi, np.ndarray):
            phi = np.ascontiguousarray(phi, dtype=np.float32)
            shift_xyz = np.ascontiguousarray(shift_xyz, dtype=np.float32)
        else:
            on_dev = 1 if phi.is_cuda else 0
            phi = phi.contiguous()
            shift_xyz = shift_xyz.contiguous()
            if not on_dev:
                phi = phi.numpy()
                shift_xyz = shift_xyz.numpy()
        nz, ny, nx = phi.shape
        if tuple(shift_xyz.shape) != (nz, ny, nx, 3):
            raise ValueError("shift must be (nz, ny, nx, 3) interleaved")
        self.uploads += 1
        _check(lib().srt_upload(self._ctx, _ptr(phi), _ptr(shift_xyz), nx, ny, nz, on_dev, shift_format),
               "srt_upload")
        self.n = nx

    def set_layout(self, layout: int, brick: int = 16) -> None:
        """Device layout of the snapshot (LAYOUT_TAPS or LAYOUT_BICUBIC) and column brick edge.

        LAYOUT_BICUBIC adds per-cell bicubic coefficient records, which mixed
        traces read; they are built now if a snapshot is uploaded, else at the
        next upload.
        """
        _check(lib().srt_set_layout(self._ctx, layout, brick), "srt_set_layout")

    def layout(self) -> tuple[int, int, int]:
        """(layout, brick edge, bytes of bicubic records held)."""
        lay, brick, nbytes = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(lib().srt_get_layout(self._ctx, ctypes.byref(lay), ctypes.byref(brick), ctypes.byref(nbytes)),
               "srt_get_layout")
        return lay.value, brick.value, nbytes.value

    def trace_device(self, nside: int, n_steps: int, out, *, pix_begin: int = 0, count: int | None = None,
                     pix_list=None, observer=None, precision: int = PREC_FP64, order: int = ORDER_LINEAR,
                     final_state=None, stream=None) -> None:
        """Asynchronous trace into device buffers (torch tensors on the GPU).

        out: float32 tensor of 6 * count; pix_list: int64 tensor or None;
        final_state: float64 tensor of count * 19 or None; stream: torch stream.
        """
        if count is None:
            count = pix_list.numel() if pix_list is not None else 12 * nside * nside - pix_begin
        obs = None if observer is None else np.ascontiguousarray(observer, dtype=np.float64)
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _check(lib().srt_trace(self._ctx, nside, pix_begin, count, _ptr(pix_list), n_steps, _ptr(obs),
                               precision, order, _ptr(out), _ptr(final_state), st), "srt_trace")

    def trace_adjoint(self, nside: int, n_steps: int, map_grad=None, *, checkpoint_every: int = 16,
                      pix_begin: int = 0, count: int | None = None, pix_list=None, observer=None, stream=None):
        """FP64 maps and, if map_grad is given, dL/d(grids) by the adjoint kernel.

        map_grad: float64 CUDA tensor (6, count) = dL/d(maps).  Returns
        (maps64 (6, count), grad_phi (n, n, n) or None, grad_shift (n, n, n, 3) or None),
        CUDA float64 tensors.
        """
        import torch
        if count is None:
            count = pix_list.numel() if pix_list is not None else 12 * nside * nside - pix_begin
        dev = torch.device("cuda", self.device)
        maps64 = torch.empty((6, count), dtype=torch.float64, device=dev)
        gphi = gshift = None
        if map_grad is not None:
            map_grad = map_grad.to(device=dev, dtype=torch.float64).contiguous()
            if tuple(map_grad.shape) != (6, count):
                raise ValueError("map_grad must be (6, count)")
            gphi = torch.empty((self.n,) * 3, dtype=torch.float64, device=dev)
            gshift = torch.empty((self.n,) * 3 + (3,), dtype=torch.float64, device=dev)
        obs = None if observer is None else np.ascontiguousarray(observer, dtype=np.float64)
        st = ctypes.c_void_p((stream or torch.cuda.current_stream(dev)).cuda_stream)
        _check(lib().srt_trace_adjoint(self._ctx, nside, pix_begin, count, _ptr(pix_list), n_steps, _ptr(obs),
                                       checkpoint_every, _ptr(map_grad), _ptr(maps64), _ptr(gphi), _ptr(gshift),
                                       st), "srt_trace_adjoint")
        return maps64, gphi, gshift

    def last_trace_ms(self) -> float:
        ms = ctypes.c_float()
        _check(lib().srt_last_trace_ms(self._ctx, ctypes.byref(ms)), "srt_last_trace_ms")
        return float(ms.value)

    def trace(self, nside: int, n_steps: int, *, batch: int | None = None, pix=None, observer=None,
              precision: int = PREC_FP64, order: int = ORDER_LINEAR, return_state: bool = False):
        """Synchronous trace with host outputs.

        Returns maps (6, count) float32 in slot order (RING pixel order of the
        batch, or the order of `pix`), plus the final state (count, 19) when
        return_state is set.
        """
        pix_arr = None
        if pix is not None:
            pix_arr = np.ascontiguousarray(pix, dtype=np.int64)
            start, count = 0, pix_arr.size
        else:
            start, count = batch_range(nside, batch)
        out = np.empty((6, count), np.float32)
        state = np.empty((count, 19), np.float64) if return_state else None
        obs = None if observer is None else np.ascontiguousarray(observer, dtype=np.float64)
        _check(lib().srt_trace_host(self._ctx, nside, start, count, _ptr(pix_arr), n_steps, _ptr(obs),
                                    precision, order, _ptr(out), _ptr(state)), "srt_trace_host")
        return (out, state) if return_state else out


class Plan:
    """A batched persistent launch over many (snapshot, pixel range, steps, observer) instances.

    items: list of dicts with keys tracer (Tracer), nside, pix_begin, count,
    n_steps, observer (optional) and out (float32 CUDA tensor of 6 * count).
    The plan holds references to the items' tracers and outputs, so their
    contexts stay alive while the plan does.
    """

    def __init__(self, items, precision: int = PREC_MIXED):
        arr = (BatchItem * len(items))()
        self._keep = []
        for i, it in enumerate(items):
            b = arr[i]
            b.ctx = it["tracer"]._ctx.value
            b.nside = it["nside"]
            b.pix_begin = it.get("pix_begin", 0)
            b.count = it["count"]
            b.n_steps = it["n_steps"]
            obs = it.get("observer")
            b.observer_set = 0 if obs is None else 1
            if obs is not None:
                for k in range(3):
                    b.observer[k] = float(obs[k])
            b.out = it["out"].data_ptr()
            self._keep.append(it["out"])
            self._keep.append(it["tracer"])
        self._plan = ctypes.c_void_p()
        _check(lib().srt_plan_create(arr, len(items), precision, ctypes.byref(self._plan)), "srt_plan_create")

    def close(self) -> None:
        if self._plan:
            lib().srt_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_order(self, order: int = ORDER_NESTED, tile_groups: int = 8) -> None:
        """Lane order within each item and the number of XCD tile ranges (outputs do not change)."""
        _check(lib().srt_plan_set_order(self._plan, order, tile_groups), "srt_plan_set_order")

    def set_schedule(self, schedule: int) -> None:
        """SCHEDULE_PERSISTENT (work-stealing workgroups) or SCHEDULE_DISPATCH (one workgroup per tile)."""
        _check(lib().srt_plan_set_schedule(self._plan, schedule), "srt_plan_set_schedule")

    def schedule(self) -> int:
        v = ctypes.c_int()
        _check(lib().srt_plan_get_schedule(self._plan, ctypes.byref(v)), "srt_plan_get_schedule")
        return v.value

    def configure(self, tile_rays: int, blocks: int) -> None:
        _check(lib().srt_plan_configure(self._plan, tile_rays, blocks), "srt_plan_configure")

    def config(self):
        t, b, rs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(lib().srt_plan_get_config(self._plan, ctypes.byref(t), ctypes.byref(b), ctypes.byref(rs)),
               "srt_plan_get_config")
        return t.value, b.value, rs.value

    def autotune(self, stream=None):
        t, b = ctypes.c_int(), ctypes.c_int()
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _check(lib().srt_plan_autotune(self._plan, st, ctypes.byref(t), ctypes.byref(b)), "srt_plan_autotune")
        return t.value, b.value

    def autotune_brick(self, stream=None) -> int:
        """Rebuilds the contexts' bicubic records at each candidate brick edge, keeps the fastest, returns it."""
        b = ctypes.c_int()
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _check(lib().srt_plan_autotune_brick(self._plan, st, ctypes.byref(b)), "srt_plan_autotune_brick")
        return b.value

    def launch(self, stream=None, graph: bool = False) -> None:
        st = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _check(lib().srt_plan_launch(self._plan, 1 if graph else 0, st), "srt_plan_launch")

    def last_ms(self) -> float:
        ms = ctypes.c_float()
        _check(lib().srt_plan_last_ms(self._plan, ctypes.byref(ms)), "srt_plan_last_ms")
        return float(ms.value)


def full_sky_maps(nside: int, batch: int | None, maps: np.ndarray) -> np.ndarray:
    """Place batch maps (6, count) into zeroed full-sky RING maps (6, 12 nside^2), .cu:924-942."""
    start, count = batch_range(nside, batch)
    full = np.zeros((6, 12 * nside * nside), np.float32)
    full[:, start:start + count] = maps
    return full


def version() -> str:
    return lib().srt_version().decode()


def ray_maps(phi, shift_xyz, nside: int, n_steps: int, *, tracer: Tracer | None = None, checkpoint_every: int = 16,
             pix_begin: int = 0, count: int | None = None, pix_list=None, observer=None):
    """Differentiable FP64 maps (6, count) of a snapshot, for torch autograd.

    phi (n, n, n) and shift_xyz (n, n, n, 3) are CUDA tensors (stored on the
    device as float32, the precision the grids have in the ray tracer); the
    backward pass runs the adjoint kernel (srt_trace_adjoint).  With `tracer`
    given, the forward pass uploads phi and shift into it, replacing its
    snapshot; the backward pass uploads them again if the tracer holds another
    snapshot by then (e.g. after a second ray_maps call on the same tracer),
    so each output's gradient is taken at its own snapshot.
    """
    import torch

    class _RayMaps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, phi_t, shift_t):
            tr = tracer or Tracer(phi_t.device.index or 0)
            grids = (phi_t.detach().float().contiguous(), shift_t.detach().float().contiguous())
            tr.upload(*grids)
            maps, _, _ = tr.trace_adjoint(nside, n_steps, None, checkpoint_every=checkpoint_every,
                                          pix_begin=pix_begin, count=count, pix_list=pix_list, observer=observer)
            ctx.tr = tr
            ctx.grids = grids
            ctx.upload = tr.uploads
            ctx.dtypes = (phi_t.dtype, shift_t.dtype)
            return maps

        @staticmethod
        def backward(ctx, grad_maps):
            if ctx.tr.uploads != ctx.upload:  # the tracer was given another snapshot since forward
                ctx.tr.upload(*ctx.grids)
                ctx.upload = ctx.tr.uploads
            _, gphi, gshift = ctx.tr.trace_adjoint(nside, n_steps, grad_maps, checkpoint_every=checkpoint_every,
                                                   pix_begin=pix_begin, count=count, pix_list=pix_list,
                                                   observer=observer)
            return gphi.to(ctx.dtypes[0]), gshift.to(ctx.dtypes[1])

    return _RayMaps.apply(phi, shift_xyz)
