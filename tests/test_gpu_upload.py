"""Uploads from host memory (srt_upload with host buffers) and host outputs
(srt_trace_host), both through the library's pinned transfer ring, against
the device paths.

Workload: periodic 260^3 snapshots (seeds 41 and 42; the potential is 4.19
chunks of the 16 MiB ring, the interleaved shift 12.57 chunks, so both end in
a partial chunk), Nside 16, 20 steps; a 72^3 snapshot (seed 43) for a change
of grid edge; Nside 512 full sky (3,145,728 rays), 2 steps, for the host
outputs (maps 75.5 MB and final states 478 MB: 5 and 29 chunks).

Bar: bit-identical.  The host path places each element into the padded
cubes' interiors and fills the periodic halo from them; the device path pads
in one pass.  Both give every padded cell the same float (after the same
binary16 quantisation in the FP16 storage mode), so every map, in every
precision, must match bit for bit; and re-uploading into a context that keeps
its buffers must give what a fresh context gives.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NSIDE, STEPS = 260, 16, 20


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.fixture(scope="module")
def snaps(fields):
    out = {}
    for seed in (41, 42):
        phi, shift = fields.snapshot(N, seed, device="cuda:0")  # generated on the GPU (faster), used from host copies
        out[seed] = (phi.cpu().numpy(), fields.interleave_shift(shift).cpu().numpy())
    return out


def traced(srt, tr, precisions=("mixed", "fp64")):
    return {p: tr.trace(NSIDE, STEPS, precision=srt.PRECISIONS[p], order=srt.ORDER_NESTED) for p in precisions}


def from_device(srt, phi, shift_xyz, shift_format, layout=None):
    import torch
    dev = torch.device("cuda:0")
    tr = srt.Tracer(0)
    if layout is not None:
        tr.set_layout(*layout)
    tr.upload(torch.from_numpy(phi).to(dev), torch.from_numpy(shift_xyz).to(dev), shift_format)
    return tr


@pytest.mark.parametrize("shift_format", ["f32", "f16_scaled"])
def test_host_upload_matches_device_upload(srt, snaps, shift_format):
    fmt = srt.SHIFT_F32 if shift_format == "f32" else srt.SHIFT_F16_SCALED
    print(f"[gpu] host vs device upload, {N}^3, shift {shift_format}", flush=True)
    phi, shift = snaps[41]
    a = srt.Tracer(0)
    a.set_layout(srt.LAYOUT_BICUBIC, 16)
    a.upload(phi, shift, fmt)
    b = from_device(srt, phi, shift, fmt, layout=(srt.LAYOUT_BICUBIC, 16))
    ma, mb = traced(srt, a), traced(srt, b)
    for p in ma:
        assert np.array_equal(bits(ma[p]), bits(mb[p])), p
    # the REF form reads the padded cubes directly
    ra = a.trace(NSIDE, STEPS, precision=srt.PREC_REF, order=srt.ORDER_NESTED)
    rb = b.trace(NSIDE, STEPS, precision=srt.PREC_REF, order=srt.ORDER_NESTED)
    assert np.array_equal(bits(ra), bits(rb))
    a.close()
    b.close()


def test_reupload_into_kept_buffers(srt, fields, snaps):
    """A second snapshot of the same edge goes into the buffers of the first
    (padded cubes, FP64 copies, records); a snapshot of another edge replaces
    them.  Each must trace exactly as on a fresh context."""
    print("[gpu] re-upload into kept buffers", flush=True)
    a = srt.Tracer(0)
    a.set_layout(srt.LAYOUT_BICUBIC, 16)
    a.upload(*snaps[41])
    first = traced(srt, a)  # builds the FP64 copies and records of snapshot 41
    a.upload(*snaps[42])
    second = traced(srt, a)
    fresh = from_device(srt, *snaps[42], srt.SHIFT_F32, layout=(srt.LAYOUT_BICUBIC, 16))
    ref = traced(srt, fresh)
    for p in ref:
        assert np.array_equal(bits(second[p]), bits(ref[p])), p
        assert not np.array_equal(bits(first[p]), bits(second[p])), p
    phi3, shift3 = fields.snapshot(72, 43)
    phi3, shift3 = phi3.numpy(), fields.interleave_shift(shift3).numpy()
    a.upload(phi3, shift3)
    small = traced(srt, a)
    fresh3 = from_device(srt, phi3, shift3, srt.SHIFT_F32, layout=(srt.LAYOUT_BICUBIC, 16))
    ref3 = traced(srt, fresh3)
    for p in ref3:
        assert np.array_equal(bits(small[p]), bits(ref3[p])), p
    for t in (a, fresh, fresh3):
        t.close()


def test_host_outputs_in_chunks(srt, snaps):
    """srt_trace_host copies maps and final states back chunk by chunk: the
    same bits as the device outputs of srt_trace."""
    import torch
    print("[gpu] host outputs in chunks, Nside 512 full sky, 2 steps", flush=True)
    nside, steps = 512, 2
    tr = srt.Tracer(0)
    tr.set_layout(srt.LAYOUT_BICUBIC, 16)
    tr.upload(*snaps[41])
    count = 12 * nside * nside
    maps_h, state_h = tr.trace(nside, steps, precision=srt.PREC_MIXED, order=srt.ORDER_NESTED, return_state=True)
    dev = torch.device("cuda:0")
    out = torch.empty(6 * count, dtype=torch.float32, device=dev)
    state = torch.empty(19 * count, dtype=torch.float64, device=dev)
    tr.trace_device(nside, steps, out, precision=srt.PREC_MIXED, order=srt.ORDER_NESTED, final_state=state)
    torch.cuda.synchronize()
    assert np.array_equal(bits(maps_h), bits(out.view(6, count).cpu().numpy()))
    assert np.array_equal(bits(state_h), bits(state.view(count, 19).cpu().numpy()))
    tr.close()


@pytest.mark.parametrize("n", [4, 5, 33])
def test_small_grids_host_upload(srt, fields, n):
    """Edges where the halo is most of the padded cube (n = 4: 343 padded
    cells, 279 of them halo): host and device uploads trace alike."""
    print(f"[gpu] host vs device upload, {n}^3", flush=True)
    phi, shift = fields.snapshot(n, 50 + n)
    phi, shift = phi.numpy(), fields.interleave_shift(shift).numpy()
    for fmt in (srt.SHIFT_F32, srt.SHIFT_F16_SCALED):
        a = srt.Tracer(0)
        a.set_layout(srt.LAYOUT_BICUBIC, 16)
        a.upload(phi, shift, fmt)
        b = from_device(srt, phi, shift, fmt, layout=(srt.LAYOUT_BICUBIC, 16))
        ma, mb = traced(srt, a, ("mixed", "fp64", "ref")), traced(srt, b, ("mixed", "fp64", "ref"))
        for p in ma:
            assert np.array_equal(bits(ma[p]), bits(mb[p])), (p, fmt)
        a.close()
        b.close()
