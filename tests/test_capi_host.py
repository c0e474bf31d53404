"""Host-side checks of the C ABI (no GPU): the library loads, exports every
entry point include/srt.h declares, and rejects bad arguments before any
device call.  Plus the batch/indexing semantics of .cu:708-717 and .cu:924-942."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "srt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(srt_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_the_bound_exports(srt):
    assert declared_functions() == sorted(srt.EXPORTS)


def test_library_exports_every_declared_symbol(srt):
    L = srt.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_status_strings_and_version(srt):
    L = srt.lib()
    assert L.srt_strerror(0) == b"success"
    for code in range(1, 7):
        assert L.srt_strerror(code) not in (b"", b"unknown status")
    assert L.srt_strerror(99) == b"unknown status"
    assert b"gfx950" in L.srt_version()


def test_null_arguments_rejected_without_device_calls(srt):
    L = srt.lib()
    assert L.srt_create(0, None) == 1
    assert L.srt_destroy(None) == 0
    assert L.srt_upload(None, None, None, 8, 8, 8, 0, 0) == 1
    assert L.srt_trace(None, 4, 0, 1, None, 4, None, 0, 0, None, None, None) == 1
    assert L.srt_trace_host(None, 4, 0, 1, None, 4, None, 0, 0, None, None) == 1
    ms = ctypes.c_float()
    assert L.srt_last_trace_ms(None, ctypes.byref(ms)) == 1
    assert L.srt_grid_info(None, None, None, None, None) == 1
    assert L.srt_set_layout(None, 1, 8) == 1
    assert L.srt_get_layout(None, None, None, None) == 1
    assert L.srt_plan_set_order(None, 1, 8) == 1
    assert L.srt_plan_autotune(None, None, None, None) == 1
    assert L.srt_plan_autotune_brick(None, None, None) == 1
    assert L.srt_plan_set_schedule(None, 1) == 1
    assert L.srt_plan_get_schedule(None, None) == 1
    assert b"null" in L.srt_last_error()


@pytest.mark.parametrize("nside", [1, 2, 16, 1024])
def test_batch_ranges(srt, nside):
    npix = 12 * nside * nside
    assert srt.batch_range(nside, None) == (0, npix)
    assert srt.batch_range(nside, -1) == (0, npix)
    covered = np.zeros(npix, np.int64)
    for b in range(12):
        lo, cnt = srt.batch_range(nside, b)
        assert (lo, cnt) == (b * nside * nside, nside * nside)
        covered[lo:lo + cnt] += 1
    assert np.all(covered == 1)
    with pytest.raises(ValueError):
        srt.batch_range(nside, 12)


def test_batch_is_a_ring_range_not_a_base_pixel(srt, oracle):
    """Batch 0 at Nside 1024 is the polar cap of rings 1..724 (partial last ring),
    not base pixel 0 (.cu:708-717; the README's 'one base pixel' wording)."""
    nside = 1024
    lo, cnt = srt.batch_range(nside, 0)
    z, s, phi = oracle.pix2zphi_ring(nside, np.array([lo, lo + cnt - 1]))
    ring_last = int(round(np.sqrt(3 * nside ** 2 * (1 - z[1]))))
    assert ring_last == 724
    assert phi[0] < np.pi / 2 < 2 * np.pi - 1e-3 < phi[1] + 4  # spans all longitudes


def test_full_sky_assembly_masks(srt):
    nside = 4
    maps = np.arange(6 * nside * nside, dtype=np.float32).reshape(6, -1) + 1
    full = srt.full_sky_maps(nside, 3, maps)
    lo = 3 * nside * nside
    assert np.all(full[:, :lo] == 0) and np.all(full[:, lo + nside * nside:] == 0)
    assert np.array_equal(full[:, lo:lo + nside * nside], maps)


def test_one_hip_runtime_per_process():
    """Loading the library through the package binds it to torch's HIP runtime
    (torch is imported first), so the process maps one libamdhip64, whichever
    of torch and the library touches the GPU first."""
    import ast
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, {root!r}); import __graft_entry__; __graft_entry__.load_package(); "
            "import snapshot_raytracer_amd as srt; srt.lib(); "
            "print(sorted({{l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}}))").format(root=ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    libs = ast.literal_eval(out.stdout.strip().splitlines()[-1])
    assert len(libs) == 1, libs


def test_effective_lane_order_host_only(srt):
    """NESTED traces sort their rays unless they cover the full sky at a power-of-two Nside; a sort
    covers at most 2^31 - 1 rays, and larger requests (the CLI's batches at Nside >= 46341, full skies
    at non-power-of-two Nside >= 13378) run in LINEAR order instead of failing."""
    N, L = srt.ORDER_NESTED, srt.ORDER_LINEAR
    big = 46341  # one RING batch of Nside^2 > 2^31 - 1 rays
    assert big * big > 2 ** 31 - 1 > (big - 1) ** 2
    assert srt.effective_order(big, 0, big * big) == L
    assert srt.effective_order(big - 1, 0, (big - 1) ** 2) == N
    assert srt.effective_order(13378, 0, 12 * 13378 ** 2) == L  # full sky, not a power of two
    assert srt.effective_order(13377, 0, 12 * 13377 ** 2) == N
    assert srt.effective_order(1 << 15, 0, 12 << 30) == N  # full sky at 2^15: lanes are NESTED pixels
    assert srt.effective_order(1 << 15, 0, 12 << 30, has_pix_list=True) == L
    assert srt.effective_order(1024, 5, 100, order=L) == L
    assert srt.effective_order(1024, 5, 100, has_pix_list=True) == N
    with pytest.raises(srt.SrtError):
        srt.effective_order(0, 0, 1)
    with pytest.raises(srt.SrtError):
        srt.effective_order(4, 0, 1, order=7)
