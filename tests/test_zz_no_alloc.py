"""srt_trace_noalloc (the srt_trace variant that allocates no device memory:
its sorted lane order's scratch lives in a fixed device array), through the C
ABI against srt_trace: bit-identical maps and final states on two calls in a
row, the second with a smaller input, in every precision and lane order; and
the requests it must refuse (a sorted order beyond the array's 2^20 rays, an
FP64 trace before the FP64 grid copies exist) fail before touching the output.

Workload: periodic 64^3 snapshot (seed 1), 8 steps; call 1: a list of 2^20
distinct random pixels at Nside 1024 (the largest sorted order the array
holds); call 2: RING batch 3 at Nside 32 (1,024 rays); plus the full sky at
Nside 32 (lanes are NESTED pixels, no scratch).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, STEPS = 64, 8
MAX_SORTED = 1 << 20


@pytest.fixture(scope="module")
def tracers(srt, fields):
    import torch
    phi, shift = fields.snapshot(N, 1)
    phi_d, shift_d = phi.to("cuda:0"), fields.interleave_shift(shift).to("cuda:0")
    taps = srt.Tracer(0)
    taps.upload(phi_d, shift_d)
    rec = srt.Tracer(0)
    rec.set_layout(srt.LAYOUT_BICUBIC, 16)
    rec.upload(phi_d, shift_d)
    fresh = srt.Tracer(0)  # no FP64 grid copies: never traced in FP64
    fresh.upload(phi_d, shift_d)
    torch.cuda.synchronize()
    yield dict(taps=taps, rec=rec, fresh=fresh)
    for t in (taps, rec, fresh):
        t.close()


def call(srt, fn, tr, nside, pix_begin, count, pix_list, prec, order, out, state):
    return getattr(srt.lib(), fn)(tr._ctx, nside, pix_begin, count,
                                  None if pix_list is None else ctypes.c_void_p(pix_list.data_ptr()), STEPS, None,
                                  prec, order, ctypes.c_void_p(out.data_ptr()),
                                  None if state is None else ctypes.c_void_p(state.data_ptr()), None)


def requests():
    import torch
    rng = np.random.default_rng(17)
    big = torch.from_numpy(rng.choice(12 * 1024 * 1024, MAX_SORTED, replace=False).astype(np.int64)).to("cuda:0")
    return [(1024, 0, MAX_SORTED, big), (32, 3 * 32 * 32, 32 * 32, None), (32, 0, 12 * 32 * 32, None)]


@pytest.mark.parametrize("layout,prec", [("taps", "fp64"), ("taps", "mixed"), ("taps", "ref"), ("rec", "mixed")])
def test_noalloc_trace_equals_trace(srt, tracers, layout, prec):
    import torch
    print(f"[gpu] srt_trace_noalloc vs srt_trace: {layout}, {prec}", flush=True)
    tr = tracers[layout]
    p = srt.PRECISIONS[prec]
    for order in (srt.ORDER_NESTED, srt.ORDER_LINEAR):
        for nside, begin, count, pix in requests():  # large, then smaller, then the full sky
            outs = []
            for fn in ("srt_trace", "srt_trace_noalloc"):
                out = torch.full((6 * count,), float("nan"), dtype=torch.float32, device="cuda:0")
                state = torch.full((19 * count,), float("nan"), dtype=torch.float64, device="cuda:0")
                rc = call(srt, fn, tr, nside, begin, count, pix, p, order, out, state)
                assert rc == 0, (fn, srt.lib().srt_last_error())
                torch.cuda.synchronize()
                outs.append((out, state))
            assert torch.equal(outs[0][0], outs[1][0]), (order, nside, count)
            assert torch.equal(outs[0][1], outs[1][1]), (order, nside, count)
            assert not bool(torch.isnan(outs[1][0]).any())


def test_noalloc_refuses_what_needs_allocation(srt, tracers):
    import torch
    print("[gpu] srt_trace_noalloc refusals", flush=True)
    L = srt.lib()
    count = MAX_SORTED + 1
    pix = torch.arange(count, dtype=torch.int64, device="cuda:0")
    out = torch.zeros(6 * count, dtype=torch.float32, device="cuda:0")
    rc = call(srt, "srt_trace_noalloc", tracers["rec"], 1024, 0, count, pix, srt.PREC_MIXED, srt.ORDER_NESTED,
              out, None)
    assert rc == 6 and b"scratch" in L.srt_last_error()
    # a RING range of the same size needs the same sort
    rc = call(srt, "srt_trace_noalloc", tracers["rec"], 1024, 0, count, None, srt.PREC_MIXED, srt.ORDER_NESTED,
              out, None)
    assert rc == 6
    # the FP64 kernel's grid copies are not built by this entry point
    small = out[:6 * 3072]
    rc = call(srt, "srt_trace_noalloc", tracers["fresh"], 16, 0, 3072, None, srt.PREC_FP64, srt.ORDER_LINEAR,
              small, None)
    assert rc == 6 and b"FP64" in L.srt_last_error()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
    # LINEAR order needs no scratch at any size
    rc = call(srt, "srt_trace_noalloc", tracers["rec"], 1024, 0, count, None, srt.PREC_MIXED, srt.ORDER_LINEAR,
              out, None)
    assert rc == 0
    torch.cuda.synchronize()
    ref = torch.empty_like(out)
    assert call(srt, "srt_trace", tracers["rec"], 1024, 0, count, None, srt.PREC_MIXED, srt.ORDER_NESTED,
                ref, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
