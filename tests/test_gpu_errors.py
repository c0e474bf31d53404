"""The C ABI's argument checks, call by call: every rejected call returns its status (include/srt.h) without
touching its outputs, and the context keeps tracing exactly as before.  Among the cases: pixel ranges whose end
would overflow int64, a step count below zero, unknown enums, a context without a snapshot, the adjoint's own
rules, plan items out of range, and uploads that must leave the previous snapshot in place."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OK, INVALID, NOT_UPLOADED, SHAPE, UNSUPPORTED = 0, 1, 4, 5, 6
I64_MAX = (1 << 63) - 1
NSIDE, STEPS = 4, 6
NPIX = 12 * NSIDE * NSIDE


@pytest.fixture(scope="module")
def ctx(srt, fields):
    import torch
    phi, shift = fields.snapshot(16, 81)
    tr = srt.Tracer(0)
    tr.upload(phi.numpy(), fields.interleave_shift(shift).numpy())
    ref = torch.empty(6 * NPIX, dtype=torch.float32, device="cuda:0")
    tr.trace_device(NSIDE, STEPS, ref, precision=srt.PREC_MIXED)
    torch.cuda.synchronize()
    yield tr, ref
    tr.close()


def still_traces(srt, tr, ref):
    import torch
    o = torch.empty_like(ref)
    tr.trace_device(NSIDE, STEPS, o, precision=srt.PREC_MIXED)
    torch.cuda.synchronize()
    assert torch.equal(o, ref)


def test_trace_argument_checks(srt, ctx):
    import torch
    tr, ref = ctx
    L = srt.lib()
    out = torch.zeros(6 * NPIX, dtype=torch.float32, device="cuda:0")
    o = ctypes.c_void_p(out.data_ptr())
    # (nside, pix_begin, count, n_steps, precision, order, out) -> status
    cases = [((0, 0, 1, STEPS, 1, 0, o), INVALID), ((-4, 0, 1, STEPS, 1, 0, o), INVALID),
             (((1 << 20) + 1, 0, 1, STEPS, 1, 0, o), INVALID), ((NSIDE, 0, 1, -1, 1, 0, o), INVALID),
             ((NSIDE, 0, -1, STEPS, 1, 0, o), INVALID), ((NSIDE, -1, 5, STEPS, 1, 0, o), INVALID),
             ((NSIDE, I64_MAX - 2, 5, STEPS, 1, 0, o), INVALID), ((NSIDE, 0, NPIX + 1, STEPS, 1, 0, o), INVALID),
             ((NSIDE, 5, NPIX - 4, STEPS, 1, 0, o), INVALID), ((NSIDE, 0, I64_MAX, STEPS, 1, 0, o), INVALID),
             ((NSIDE, 0, NPIX, STEPS, 3, 0, o), INVALID), ((NSIDE, 0, NPIX, STEPS, -1, 0, o), INVALID),
             ((NSIDE, 0, NPIX, STEPS, 1, 2, o), INVALID), ((NSIDE, 0, NPIX, STEPS, 1, 0, None), INVALID)]
    for (nside, lo, cnt, steps, prec, order, dst), want in cases:
        for fn in (L.srt_trace, L.srt_trace_noalloc):
            rc = fn(tr._ctx, nside, lo, cnt, None, steps, None, prec, order, dst, None, None)
            assert rc == want, (fn.__name__, nside, lo, cnt, steps, prec, order, rc)
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
    # a context without a snapshot
    fresh = srt.Tracer(0)
    assert L.srt_trace(fresh._ctx, NSIDE, 0, NPIX, None, STEPS, None, 1, 0, o, None, None) == NOT_UPLOADED
    fresh.close()
    still_traces(srt, tr, ref)


def test_adjoint_argument_checks(srt, ctx):
    import torch
    tr, ref = ctx
    L = srt.lib()
    n = 16
    maps = torch.zeros((6, NPIX), dtype=torch.float64, device="cuda:0")
    g = torch.ones((6, NPIX), dtype=torch.float64, device="cuda:0")
    gphi = torch.zeros((n, n, n), dtype=torch.float64, device="cuda:0")
    gsh = torch.zeros((n, n, n, 3), dtype=torch.float64, device="cuda:0")
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
    # (pix_begin, count, n_steps, checkpoint_every, map_grad, maps64, grad_phi, grad_shift) -> status
    cases = [((0, NPIX, 0, 4, g, maps, gphi, gsh), INVALID), ((0, NPIX, STEPS, 0, g, maps, gphi, gsh), INVALID),
             ((0, NPIX, STEPS, 4, g, maps, None, gsh), INVALID), ((0, NPIX, STEPS, 4, g, maps, gphi, None), INVALID),
             ((0, NPIX, STEPS, 4, None, None, gphi, gsh), INVALID), ((-1, 5, STEPS, 4, g, maps, gphi, gsh), INVALID),
             ((I64_MAX - 2, 5, STEPS, 4, g, maps, gphi, gsh), INVALID), ((0, -1, STEPS, 4, g, maps, gphi, gsh), INVALID),
             ((0, NPIX + 1, STEPS, 4, g, maps, gphi, gsh), INVALID)]
    for (lo, cnt, steps, k, mg, m64, gp, gs), want in cases:
        rc = L.srt_trace_adjoint(tr._ctx, NSIDE, lo, cnt, None, steps, None, k, P(mg), P(m64), P(gp), P(gs), None)
        assert rc == want, (lo, cnt, steps, k, rc)
    torch.cuda.synchronize()
    assert float(maps.abs().max()) == 0.0 and float(gphi.abs().max()) == 0.0 and float(gsh.abs().max()) == 0.0
    still_traces(srt, tr, ref)


def test_plan_item_checks(srt, ctx):
    import torch
    tr, ref = ctx
    L = srt.lib()
    out = torch.zeros(6 * NPIX, dtype=torch.float32, device="cuda:0")

    def item(**kw):
        b = srt.BatchItem()
        b.ctx, b.nside, b.pix_begin, b.count, b.n_steps = tr._ctx.value, NSIDE, 0, NPIX, STEPS
        b.observer_set, b.out = 0, out.data_ptr()
        for k, v in kw.items():
            setattr(b, k, v)
        return b
    bad = [dict(nside=0), dict(nside=(1 << 20) + 1), dict(n_steps=-1), dict(count=-1), dict(pix_begin=-1),
           dict(pix_begin=I64_MAX - 2, count=5), dict(count=NPIX + 1), dict(pix_begin=1, count=NPIX),
           dict(out=None), dict(ctx=None)]
    for kw in bad:
        arr = (srt.BatchItem * 1)(item(**kw))
        plan = ctypes.c_void_p()
        assert L.srt_plan_create(arr, 1, srt.PREC_MIXED, ctypes.byref(plan)) == INVALID, kw
        assert not plan.value
    arr = (srt.BatchItem * 1)(item())
    plan = ctypes.c_void_p()
    assert L.srt_plan_create(arr, 0, srt.PREC_MIXED, ctypes.byref(plan)) == INVALID
    assert L.srt_plan_create(arr, 1, 3, ctypes.byref(plan)) == INVALID
    assert L.srt_plan_create(None, 1, srt.PREC_MIXED, ctypes.byref(plan)) == INVALID
    fresh = srt.Tracer(0)
    arr = (srt.BatchItem * 1)(item(ctx=fresh._ctx.value))
    assert L.srt_plan_create(arr, 1, srt.PREC_MIXED, ctypes.byref(plan)) == NOT_UPLOADED
    fresh.close()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
    still_traces(srt, tr, ref)


def test_rejected_uploads_keep_the_snapshot(srt, ctx):
    tr, ref = ctx
    L = srt.lib()
    z = np.zeros((16, 16, 16), np.float32)
    z3 = np.zeros((16, 16, 16, 3), np.float32)
    p, p3 = z.ctypes.data_as(ctypes.c_void_p), z3.ctypes.data_as(ctypes.c_void_p)
    # (nx, ny, nz, shift_format, phi, shift) -> status
    cases = [((16, 16, 8, 0, p, p3), SHAPE), ((3, 3, 3, 0, p, p3), SHAPE), ((2049, 2049, 2049, 0, p, p3), SHAPE),
             ((16, 16, 16, 2, p, p3), INVALID), ((16, 16, 16, 0, None, p3), INVALID),
             ((16, 16, 16, 0, p, None), INVALID)]
    for (nx, ny, nz, fmt, a, b), want in cases:
        assert L.srt_upload(tr._ctx, a, b, nx, ny, nz, 0, fmt) == want, (nx, ny, nz, fmt)
    for layout, brick in ((2, 16), (-1, 16), (1, 3), (1, 128), (1, 0)):
        assert L.srt_set_layout(tr._ctx, layout, brick) == INVALID, (layout, brick)
    still_traces(srt, tr, ref)


def test_noalloc_fp64_needs_the_fp64_copies(srt, fields):
    import torch
    phi, shift = fields.snapshot(16, 82)
    tr = srt.Tracer(0)
    tr.upload(phi.numpy(), fields.interleave_shift(shift).numpy())
    out = torch.zeros(6 * NPIX, dtype=torch.float32, device="cuda:0")
    o = ctypes.c_void_p(out.data_ptr())
    L = srt.lib()
    assert L.srt_trace_noalloc(tr._ctx, NSIDE, 0, NPIX, None, STEPS, None, srt.PREC_FP64, 0, o, None, None) == UNSUPPORTED
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
    want = torch.empty_like(out)
    tr.trace_device(NSIDE, STEPS, want, precision=srt.PREC_FP64)  # builds the FP64 copies
    assert L.srt_trace_noalloc(tr._ctx, NSIDE, 0, NPIX, None, STEPS, None, srt.PREC_FP64, 0, o, None, None) == OK
    torch.cuda.synchronize()
    assert torch.equal(out, want)
    tr.close()
