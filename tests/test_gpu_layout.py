"""Bicubic cell records (SRT_LAYOUT_BICUBIC): the mixed kernel reading
per-cell monomial coefficients instead of raw taps.

Workload: periodic 64^3 snapshot (potential rms 1e-5, shift rms 1e-7,
P(k) ~ k^-3, seed 1), observer at the grid centre, Nside 16, 32 steps; plus
a 36^3 snapshot whose edge is not a multiple of the brick edge.

Bars:
  * against the FP64 oracle: the mixed-form bars (dA max relative <= 1e-5,
    shear rms(|d eps|)/rms(|eps|) <= 1e-4, other maps rms(d)/rms <= 1e-3) and
    the final position within 1e-6 cells;
  * the brick edge only changes addresses: maps bit-identical for every edge;
  * batched plans over contexts with records: bit-identical to single traces.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NSIDE, STEPS, SEED = 64, 16, 32, 1


@pytest.fixture(scope="module")
def snap(fields):
    phi, shift = fields.snapshot(N, SEED)
    return phi.numpy(), shift.numpy(), fields.interleave_shift(shift).numpy()


@pytest.fixture(scope="module")
def ref64(oracle, snap):
    F = oracle.Fields(snap[0], snap[1])
    pix = np.arange(12 * NSIDE * NSIDE)
    maps, _, state = oracle.trace(F, NSIDE, pix, STEPS, form=oracle.FORM_FP64, return_state=True)
    return maps, state


def map_errors(test, ref):
    out = {"dA_max_rel": float(np.max(np.abs(test[0] - ref[0]) / np.abs(ref[0])))}
    de = np.hypot(test[1] - ref[1], test[2] - ref[2])
    out["shear_rms"] = float(np.sqrt(np.mean(de ** 2)) / np.sqrt(np.mean(ref[1] ** 2 + ref[2] ** 2)))
    for m, name in ((3, "rotation"), (4, "deflection1"), (5, "deflection2")):
        rms = np.sqrt(np.mean(ref[m].astype(np.float64) ** 2))
        out[name] = float(np.sqrt(np.mean((test[m].astype(np.float64) - ref[m]) ** 2)) / rms)
    return out


def test_bicubic_mixed_matches_fp64_oracle(srt, snap, ref64):
    print("[gpu] bicubic records, mixed full sky", flush=True)
    tr = srt.Tracer(0)
    tr.set_layout(srt.LAYOUT_BICUBIC, 8)
    tr.upload(snap[0], snap[2])
    lay, brick, nbytes = tr.layout()
    assert (lay, brick) == (srt.LAYOUT_BICUBIC, 8)
    assert nbytes == 4 * 64 * N * N * (N + 3)
    maps, state = tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, return_state=True)
    e = map_errors(maps, ref64[0])
    print("bicubic mixed vs fp64 oracle:", e)
    assert e["dA_max_rel"] <= 1e-5
    assert e["shear_rms"] <= 1e-4
    for name in ("rotation", "deflection1", "deflection2"):
        assert e[name] <= 1e-3, name
    assert np.abs(state[:, :3] - ref64[1][:, :3]).max() < 1e-6
    # the FP64 and reference forms keep reading the padded cubes
    m64, s64 = tr.trace(NSIDE, STEPS, precision=srt.PREC_FP64, return_state=True)
    scale = np.maximum(np.abs(ref64[1]).max(axis=0), 1e-300)
    assert float((np.abs(s64 - ref64[1]) / scale).max()) < 1e-10
    # back to taps: records freed, mixed kernel on the cubes
    tr.set_layout(srt.LAYOUT_TAPS)
    assert tr.layout()[0] == srt.LAYOUT_TAPS and tr.layout()[2] == 0
    taps = tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED)
    assert map_errors(taps, ref64[0])["shear_rms"] <= 1e-4
    tr.close()


def test_brick_edge_changes_only_addresses(srt, snap):
    print("[gpu] bicubic records, brick edges", flush=True)
    tr = srt.Tracer(0)
    tr.upload(snap[0], snap[2])
    outs = []
    for brick in (1, 2, 8, 64):
        tr.set_layout(srt.LAYOUT_BICUBIC, brick)
        outs.append(tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, order=srt.ORDER_NESTED))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    # the layout persists across uploads
    tr.upload(snap[0], snap[2])
    assert tr.layout()[0] == srt.LAYOUT_BICUBIC
    assert np.array_equal(tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED), outs[0])
    tr.close()


def test_edge_not_multiple_of_brick(srt, oracle, fields):
    print("[gpu] bicubic records, 36^3 with 8-column bricks", flush=True)
    n, nside, steps = 36, 8, 20
    phi, shift = fields.snapshot(n, 7)
    tr = srt.Tracer(0)
    tr.set_layout(srt.LAYOUT_BICUBIC, 8)
    tr.upload(phi.numpy(), fields.interleave_shift(shift).numpy())
    obs = np.array([3.3, 30.1, 17.9])  # rays wrap across every face of the box
    maps = tr.trace(nside, steps, precision=srt.PREC_MIXED, observer=obs)
    tr.close()
    F = oracle.Fields(phi.numpy(), shift.numpy())
    ref, _ = oracle.trace(F, nside, np.arange(12 * nside * nside), steps, observer=obs, form=oracle.FORM_FP64)
    e = map_errors(maps, ref)
    print("36^3 bicubic mixed vs fp64 oracle:", e)
    assert e["dA_max_rel"] <= 1e-5 and e["shear_rms"] <= 1e-4


def test_invalid_layout_arguments(srt):
    tr = srt.Tracer(0)
    for lay, brick in ((2, 8), (srt.LAYOUT_BICUBIC, 3), (srt.LAYOUT_BICUBIC, 128), (srt.LAYOUT_BICUBIC, 0)):
        with pytest.raises(srt.SrtError) as ei:
            tr.set_layout(lay, brick)
        assert ei.value.status == 1
    tr.close()


def test_batch_plan_over_bicubic_contexts(srt, fields):
    import torch
    print("[gpu] batch plan over bicubic contexts", flush=True)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    items = []
    for i, n in enumerate([16, 24, 20]):
        phi, shift = fields.snapshot(n, 200 + i)
        tr = srt.Tracer(0)
        tr.set_layout(srt.LAYOUT_BICUBIC, 4)
        tr.upload(phi.to(dev), fields.interleave_shift(shift).to(dev))
        nside = 4
        lo, cnt = srt.batch_range(nside, int(rng.integers(0, 12)))
        items.append(dict(tracer=tr, nside=nside, pix_begin=lo, count=cnt, n_steps=int(rng.integers(4, n // 2)),
                          observer=rng.uniform(0, n, 3), out=torch.zeros(6 * cnt, dtype=torch.float32, device=dev)))
    ref = []
    for it in items:
        o = torch.empty_like(it["out"])
        it["tracer"].trace_device(it["nside"], it["n_steps"], o, pix_begin=it["pix_begin"], count=it["count"],
                                  observer=it["observer"], precision=srt.PREC_MIXED)
        ref.append(o)
    plan = srt.Plan(items, precision=srt.PREC_MIXED)
    s = torch.cuda.Stream()
    for use_graph in (False, True, True):
        for it in items:
            it["out"].zero_()
        torch.cuda.synchronize()
        plan.launch(stream=s, graph=use_graph)
        s.synchronize()
        for it, r in zip(items, ref):
            assert torch.equal(it["out"], r)
    plan.close()
    for it in items:
        it["tracer"].close()


def test_record_allocation_failure_falls_back(srt, snap, monkeypatch):
    """Records that do not fit (forced by a per-allocation limit of the context):
    srt_set_layout and srt_upload report SRT_ERR_OUT_OF_MEMORY, the context keeps
    the snapshot on the padded grids, and traces work and equal a padded-grid
    trace."""
    print("[gpu] record allocation failure", flush=True)
    ref_tr = srt.Tracer(0)
    ref_tr.upload(snap[0], snap[2])
    ref = ref_tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED)
    ref_tr.close()
    monkeypatch.setenv("SRT_MAX_ALLOC_BYTES", str(8 << 20))  # padded grids 4.9 MB fit, records 70 MB do not
    tr = srt.Tracer(0)
    monkeypatch.delenv("SRT_MAX_ALLOC_BYTES")
    tr.set_layout(srt.LAYOUT_BICUBIC, 8)  # no snapshot yet: nothing to build
    with pytest.raises(srt.SrtError) as ei:
        tr.upload(snap[0], snap[2])
    assert ei.value.status == 3
    assert tr.layout()[0] == srt.LAYOUT_TAPS and tr.layout()[2] == 0
    assert np.array_equal(tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED), ref)
    with pytest.raises(srt.SrtError) as ei:
        tr.set_layout(srt.LAYOUT_BICUBIC, 8)
    assert ei.value.status == 3
    assert tr.layout()[0] == srt.LAYOUT_TAPS
    assert np.array_equal(tr.trace(NSIDE, STEPS, precision=srt.PREC_MIXED), ref)
    tr.close()
