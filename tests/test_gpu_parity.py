"""GPU parity of the HIP ray kernels against the host oracle, through the C ABI.

Workload: periodic 64^3 snapshot (potential rms 1e-5, shift rms 1e-7,
P(k) ~ k^-3, seed 1), observer at the grid centre, Nside 16 (3072 rays),
32 steps: the reference's small CPU-runnable case.

Bars:
  * pixel indexing, batch masks, lane orders: bit-exact;
  * FP64 kernel vs FP64 oracle: final state within 1e-10 relative to each
    component's scale, maps within 2 float ulps;
  * mixed kernel vs FP64 oracle: dA max per-pixel relative error <= 1e-5,
    shear rms(|d eps|)/rms(|eps|) <= 1e-4, other maps rms(d)/rms <= 1e-3;
  * reference-arithmetic kernel vs the oracle's REF form: dA within 1e-6
    relative, other maps within 2e-5 of each map's rms (the two evaluate the
    float interpolation in a different order; the reference itself is built
    with fast math).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NSIDE, STEPS, SEED = 64, 16, 32, 1


@pytest.fixture(scope="module")
def snap(fields):
    phi, shift = fields.snapshot(N, SEED)
    return phi.numpy(), shift.numpy(), fields.interleave_shift(shift).numpy()


@pytest.fixture(scope="module")
def tracer(srt, snap):
    print(f"[gpu] upload {N}^3 snapshot", flush=True)
    tr = srt.Tracer(0)
    tr.upload(snap[0], snap[2])
    yield tr
    tr.close()


@pytest.fixture(scope="module")
def ref64(oracle, snap):
    F = oracle.Fields(snap[0], snap[1])
    pix = np.arange(12 * NSIDE * NSIDE)
    maps, _, state = oracle.trace(F, NSIDE, pix, STEPS, form=oracle.FORM_FP64, return_state=True)
    return maps, state


def state_error(a, b):
    scale = np.abs(b).max(axis=0)
    scale[scale == 0] = 1.0
    return float((np.abs(a - b) / scale).max())


def test_fp64_kernel_matches_oracle(srt, tracer, ref64):
    print("[gpu] fp64 full sky", flush=True)
    maps, state = tracer.trace(NSIDE, STEPS, precision=srt.PREC_FP64, return_state=True)
    err = state_error(state, ref64[1])
    print(f"fp64 state error {err:.3e}")
    assert err < 1e-10
    np.testing.assert_allclose(maps, ref64[0], rtol=2.5e-7, atol=1e-30)


def test_batches_masks_and_sum(srt, tracer):
    print("[gpu] 12 batches", flush=True)
    full = tracer.trace(NSIDE, STEPS, precision=srt.PREC_FP64)
    acc = np.zeros((6, 12 * NSIDE * NSIDE), np.float32)
    for b in range(12):
        part = tracer.trace(NSIDE, STEPS, batch=b, precision=srt.PREC_FP64)
        sky = srt.full_sky_maps(NSIDE, b, part)
        lo, cnt = srt.batch_range(NSIDE, b)
        assert lo == b * NSIDE * NSIDE and cnt == NSIDE * NSIDE
        outside = np.ones(12 * NSIDE * NSIDE, bool)
        outside[lo:lo + cnt] = False
        assert np.all(sky[:, outside] == 0.0)
        assert np.array_equal(sky[:, lo:lo + cnt], full[:, lo:lo + cnt])
        acc += sky
    assert np.array_equal(acc, full)


def test_pixel_list_and_nested_order(srt, tracer):
    print("[gpu] pixel list + nested order", flush=True)
    full = tracer.trace(NSIDE, STEPS, precision=srt.PREC_FP64)
    rng = np.random.default_rng(3)
    pix = np.sort(rng.choice(12 * NSIDE * NSIDE, 500, replace=False))
    sub = tracer.trace(NSIDE, STEPS, pix=pix, precision=srt.PREC_FP64)
    assert np.array_equal(sub, full[:, pix])
    nested = tracer.trace(NSIDE, STEPS, precision=srt.PREC_FP64, order=srt.ORDER_NESTED)
    assert np.array_equal(nested, full)
    nested_m = tracer.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, order=srt.ORDER_NESTED)
    linear_m = tracer.trace(NSIDE, STEPS, precision=srt.PREC_MIXED)
    assert np.array_equal(nested_m, linear_m)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_nested_order_partial_sky(srt, tracer, prec):
    """NESTED lane order over RING batches, pixel lists (with duplicates and
    invalid entries) and a Nside that is not a power of two: the rays are
    sorted on the device, results equal LINEAR order bit for bit."""
    print(f"[gpu] nested order over partial skies ({prec})", flush=True)
    p = srt.PRECISIONS[prec]
    for b in (0, 5, 11):
        lin = tracer.trace(NSIDE, STEPS, batch=b, precision=p)
        nes = tracer.trace(NSIDE, STEPS, batch=b, precision=p, order=srt.ORDER_NESTED)
        assert np.array_equal(lin, nes), b
    rng = np.random.default_rng(12)
    pix = rng.integers(-5, 12 * NSIDE * NSIDE + 5, 900)
    pix[:10] = pix[10:20]  # duplicates
    lin = tracer.trace(NSIDE, STEPS, pix=pix, precision=p)
    nes = tracer.trace(NSIDE, STEPS, pix=pix, precision=p, order=srt.ORDER_NESTED)
    assert np.array_equal(lin, nes, equal_nan=True)
    bad = (pix < 0) | (pix >= 12 * NSIDE * NSIDE)
    assert bad.any() and np.all(np.isnan(nes[:, bad]))
    for nside in (3, 12):  # NESTED numbering is undefined here; the sort key still orders the lanes
        lin = tracer.trace(nside, 8, precision=p)
        nes = tracer.trace(nside, 8, precision=p, order=srt.ORDER_NESTED)
        assert np.array_equal(lin, nes), nside


def map_errors(test, ref):
    out = {}
    out["dA_max_rel"] = float(np.max(np.abs(test[0] - ref[0]) / np.abs(ref[0])))
    de = np.hypot(test[1] - ref[1], test[2] - ref[2])
    out["shear_rms"] = float(np.sqrt(np.mean(de ** 2)) / np.sqrt(np.mean(ref[1] ** 2 + ref[2] ** 2)))
    for m, name in ((3, "rotation"), (4, "deflection1"), (5, "deflection2")):
        rms = np.sqrt(np.mean(ref[m].astype(np.float64) ** 2))
        d = (test[m].astype(np.float64) - ref[m])
        out[name + "_rms_rel"] = float(np.sqrt(np.mean(d ** 2)) / rms)
        out[name + "_max_rel"] = float(np.max(np.abs(d)) / rms)
    return out


def test_mixed_kernel_error_bounds(srt, tracer, ref64):
    print("[gpu] mixed full sky", flush=True)
    maps = tracer.trace(NSIDE, STEPS, precision=srt.PREC_MIXED)
    e = map_errors(maps, ref64[0])
    print("mixed vs fp64 oracle:", e)
    assert e["dA_max_rel"] <= 1e-5
    assert e["shear_rms"] <= 1e-4
    for name in ("rotation", "deflection1", "deflection2"):
        assert e[name + "_rms_rel"] <= 1e-3, name


def test_ref_kernel_matches_ref_oracle(srt, oracle, tracer, snap):
    print("[gpu] ref form full sky", flush=True)
    maps = tracer.trace(NSIDE, STEPS, precision=srt.PREC_REF)
    F = oracle.Fields(snap[0], snap[1])
    ref, _ = oracle.trace(F, NSIDE, np.arange(12 * NSIDE * NSIDE), STEPS, form=oracle.FORM_REF)
    da = float(np.max(np.abs(maps[0].astype(np.float64) - ref[0]) / np.abs(ref[0])))
    print(f"ref dA: max relative difference = {da:.3e}")
    assert da < 1e-6
    for m in range(1, 6):
        rms = np.sqrt(np.mean(ref[m].astype(np.float64) ** 2))
        d = np.abs(maps[m].astype(np.float64) - ref[m]).max() / rms
        print(f"ref map {m}: max |d|/rms = {d:.3e}")
        assert d < 2e-5, m


def test_f16_shift_storage(srt, oracle, snap):
    print("[gpu] f16 shift storage", flush=True)
    tr = srt.Tracer(0)
    tr.upload(snap[0], snap[2], shift_format=srt.SHIFT_F16_SCALED)
    maps, state = tr.trace(NSIDE, STEPS, precision=srt.PREC_FP64, return_state=True)
    tr.close()
    b16 = oracle.shift_to_f16(snap[1])
    F = oracle.Fields(snap[0], snap[1], b16=b16)
    ref, _, ref_state = oracle.trace(F, NSIDE, np.arange(12 * NSIDE * NSIDE), STEPS, return_state=True)
    assert state_error(state, ref_state) < 1e-10


def test_zero_steps_and_invalid_pixels(srt, oracle, tracer, snap):
    print("[gpu] zero steps + invalid pixels", flush=True)
    maps = tracer.trace(NSIDE, 0, precision=srt.PREC_FP64)
    assert np.all(maps[0] == 0.0) and np.all(maps[1:4] == 0.0)
    assert np.all(np.isnan(maps[4:]))  # direction of a zero displacement is undefined
    pix = np.array([5, -1, 12 * NSIDE * NSIDE, 7], np.int64)
    out = tracer.trace(NSIDE, 4, pix=pix, precision=srt.PREC_FP64)
    assert np.all(np.isnan(out[:, 1])) and np.all(np.isnan(out[:, 2]))
    good = tracer.trace(NSIDE, 4, pix=np.array([5, 7]), precision=srt.PREC_FP64)
    assert np.array_equal(out[:, [0, 3]], good)


def test_empty_trace_and_errors(srt, tracer):
    out = tracer.trace(NSIDE, 4, pix=np.zeros(0, np.int64))
    assert out.shape == (6, 0)
    with pytest.raises(ValueError):
        tracer.trace(NSIDE, 4, batch=12)  # batch_range rejects, like .cu:713
    with pytest.raises(srt.SrtError) as ei:
        tr = srt.Tracer(0)
        tr.upload(np.zeros((8, 8, 4), np.float32), np.zeros((8, 8, 4, 3), np.float32))
    assert ei.value.status == 5
    with pytest.raises(srt.SrtError) as ei:
        tracer.trace_device(NSIDE, 4, None)
    assert ei.value.status == 1
    import torch
    o = torch.zeros(6 * 12 * NSIDE * NSIDE, dtype=torch.float32, device="cuda:0")
    with pytest.raises(srt.SrtError) as ei:
        tracer.trace_device(NSIDE, 4, o, precision=7)
    assert ei.value.status == 1
    assert float(o.abs().max()) == 0.0  # rejected before any work
    # an Nside above 2^20 and a count beyond the sky are refused (SRT_ERR_INVALID_ARGUMENT) before the
    # output is touched (a valid count is bounded only by the caller's output buffer, which the library
    # cannot see: traces of 2^32 rays and more are split into several launches)
    for nside, count in (((1 << 20) + 1, 12), (NSIDE, 12 * NSIDE * NSIDE + 1)):
        rc = srt.lib().srt_trace(tracer._ctx, nside, 0, count, None, 1, None, srt.PREC_MIXED,
                                 srt.ORDER_LINEAR, ctypes.c_void_p(o.data_ptr()), None, None)
        assert rc == 1
    assert float(o.abs().max()) == 0.0


@pytest.mark.parametrize("n", [16, 20])
def test_periodic_wrap_pow2_and_general_grids(srt, oracle, fields, n):
    """Rays that cross the periodic box several times, on a power-of-two grid
    (16^3: the kernels wrap cell indices with a mask) and on a general one
    (20^3: floor division): observer near three faces, 3 n steps, Nside 4.
    FP64 kernel against the FP64 oracle (final state 1e-10), mixed kernel on the
    padded cubes and on the bicubic records against the same oracle (the mixed
    bars above)."""
    print(f"[gpu] periodic wrap, {n}^3", flush=True)
    phi_t, shift_t = fields.snapshot(n, 91)
    phi, shift = phi_t.numpy(), shift_t.numpy()
    tr = srt.Tracer(0)
    tr.upload(phi, fields.interleave_shift(shift_t).numpy())
    obs = (0.3, n - 0.6, 0.45 * n)
    nside, steps = 4, 3 * n
    F = oracle.Fields(phi, shift)
    ref_maps, _, ref_state = oracle.trace(F, nside, np.arange(12 * nside * nside), steps, observer=obs,
                                          form=oracle.FORM_FP64, return_state=True)
    maps, state = tr.trace(nside, steps, observer=obs, precision=srt.PREC_FP64, return_state=True)
    assert np.abs(state[:, :3] - np.asarray(obs)).max() > 2 * n  # rays crossed the box more than once
    assert state_error(state, ref_state) < 1e-10
    for layout in (srt.LAYOUT_TAPS, srt.LAYOUT_BICUBIC):
        tr.set_layout(layout, 8)
        e = map_errors(tr.trace(nside, steps, observer=obs, precision=srt.PREC_MIXED), ref_maps)
        print(f"mixed (layout {layout}) vs fp64 oracle:", e)
        assert e["dA_max_rel"] <= 1e-5
        assert e["shear_rms"] <= 1e-4
    tr.close()


@pytest.mark.parametrize("layout", ["taps", "bicubic"])
def test_trace_split_into_launches(srt, tracer, snap, layout, monkeypatch):
    """A trace larger than the launch limit runs as several launches (one launch may not hold 2^32
    lanes): a context created with SRT_MAX_LAUNCH_RAYS = 512 (a 3072-ray sky in 6 launches, a
    1000-ray pixel list in 2, the last one partial) gives the same maps as one launch, in every
    form and lane order."""
    import torch
    print(f"[gpu] split launches, {layout}", flush=True)
    monkeypatch.setenv("SRT_MAX_LAUNCH_RAYS", "512")
    small = srt.Tracer(0)
    monkeypatch.delenv("SRT_MAX_LAUNCH_RAYS")
    one = tracer
    if layout == "bicubic":
        one = srt.Tracer(0)
        one.set_layout(srt.LAYOUT_BICUBIC, 16)
        one.upload(snap[0], snap[2])
        small.set_layout(srt.LAYOUT_BICUBIC, 16)
    small.upload(snap[0], snap[2])
    rng = np.random.default_rng(5)
    pix = rng.choice(12 * NSIDE * NSIDE, 1000, replace=False).astype(np.int64)
    forms = ("fp64", "mixed", "ref") if layout == "taps" else ("mixed",)
    for prec in forms:
        for order in (srt.ORDER_LINEAR, srt.ORDER_NESTED):
            p = srt.PRECISIONS[prec]
            for kw in (dict(), dict(batch=5), dict(pix=pix)):
                a = one.trace(NSIDE, STEPS, precision=p, order=order, **kw)
                b = small.trace(NSIDE, STEPS, precision=p, order=order, **kw)
                assert np.array_equal(a, b), (prec, order, kw.keys())
    small.close()
    if one is not tracer:
        one.close()
    torch.cuda.synchronize()
