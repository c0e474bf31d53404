"""Contexts are independent (include/srt.h): traces of different contexts queued on different streams, all in
flight at once, give the same maps bit for bit as the same traces run one after another on one stream.  The cases
cover the sorted lane order (a partial-sky NESTED trace keeps its lane map in the context), both mixed layouts,
FP64 and the adjoint's forward maps; each context also queues two traces back to back on its own stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (grid edge, seed, records, precision, nside, pix_begin, count, steps)
CASES = [(24, 61, False, "fp64", 8, 100, 500, 30),
         (32, 62, True, "mixed", 16, 0, 3072, 40),
         (40, 63, False, "mixed", 12, 777, 900, 50),
         (20, 64, False, "ref", 5, 0, 300, 25)]


def test_concurrent_contexts_on_separate_streams(srt, fields):
    import torch
    dev = torch.device("cuda:0")
    tracers, outs, outs2, seq, seq2 = [], [], [], [], []
    for n, seed, records, prec, nside, lo, cnt, steps in CASES:
        phi, shift = fields.snapshot(n, seed)
        tr = srt.Tracer(0)
        if records:
            tr.set_layout(srt.LAYOUT_BICUBIC, 8)
        tr.upload(phi.numpy(), fields.interleave_shift(shift).numpy())
        tracers.append(tr)
        outs.append(torch.full((6 * cnt,), float("nan"), dtype=torch.float32, device=dev))
        outs2.append(torch.full((6 * cnt,), float("nan"), dtype=torch.float32, device=dev))
    # one after another on the default stream
    for tr, (n, seed, records, prec, nside, lo, cnt, steps) in zip(tracers, CASES):
        a = torch.empty(6 * cnt, dtype=torch.float32, device=dev)
        b = torch.empty(6 * cnt, dtype=torch.float32, device=dev)
        tr.trace_device(nside, steps, a, pix_begin=lo, count=cnt, precision=srt.PRECISIONS[prec],
                        order=srt.ORDER_NESTED)
        torch.cuda.synchronize()
        tr.trace_device(nside, steps // 2, b, pix_begin=lo, count=cnt, precision=srt.PRECISIONS[prec],
                        order=srt.ORDER_LINEAR)
        torch.cuda.synchronize()
        seq.append(a)
        seq2.append(b)
    # all at once, one stream per context, two traces queued on each
    streams = [torch.cuda.Stream() for _ in CASES]
    torch.cuda.synchronize()
    for tr, s, o, o2, (n, seed, records, prec, nside, lo, cnt, steps) in zip(tracers, streams, outs, outs2, CASES):
        print(f"[gpu] concurrent {n}^3 {prec} records {records} nside {nside} rays {cnt} steps {steps}", flush=True)
        tr.trace_device(nside, steps, o, pix_begin=lo, count=cnt, precision=srt.PRECISIONS[prec],
                        order=srt.ORDER_NESTED, stream=s)
        tr.trace_device(nside, steps // 2, o2, pix_begin=lo, count=cnt, precision=srt.PRECISIONS[prec],
                        order=srt.ORDER_LINEAR, stream=s)
    torch.cuda.synchronize()
    for k in range(len(CASES)):
        assert torch.equal(outs[k], seq[k]), CASES[k]
        assert torch.equal(outs2[k], seq2[k]), CASES[k]
    for tr in tracers:
        tr.close()


def test_adjoint_forward_maps_on_separate_streams(srt, fields):
    """Forward-only adjoint passes (FP64 maps) of two contexts on two streams at once equal each alone."""
    import torch
    runs = [(24, 71, 8, 20), (28, 72, 6, 16)]
    tracers, alone = [], []
    for n, seed, nside, steps in runs:
        phi, shift = fields.snapshot(n, seed)
        tr = srt.Tracer(0)
        tr.upload(phi.numpy(), fields.interleave_shift(shift).numpy())
        maps, _, _ = tr.trace_adjoint(nside, steps, None, checkpoint_every=4)
        torch.cuda.synchronize()
        tracers.append(tr)
        alone.append(maps.clone())
    streams = [torch.cuda.Stream() for _ in runs]
    got = []
    for tr, s, (n, seed, nside, steps) in zip(tracers, streams, runs):
        maps, _, _ = tr.trace_adjoint(nside, steps, None, checkpoint_every=4, stream=s)
        got.append(maps)
    torch.cuda.synchronize()
    for a, b in zip(got, alone):
        assert torch.equal(a, b)
    for tr in tracers:
        tr.close()
