"""Parity over seeded random configurations: the HIP kernels against the host
FP64 oracle on shapes the fixed tests do not pin one by one.

Each of 30 cases draws (numpy default_rng(100 + case)): a grid edge in
[4, 40] (odd and non-power-of-two edges included), Nside in [1, 12] (most not
powers of two), n_steps in [0, 3n] (rays that cross the periodic box up to
three times), an observer anywhere in [-n, 2n]^3 or the default centre, the
arithmetic (FP64; FP64 on the reference's binary16 x 2^27 shift storage,
against the oracle on the same codes; mixed on the padded grids or on
bicubic records of a random brick edge; the reference's float/double mix,
against the oracle's REF form), the lane order (LINEAR or NESTED, sorted on the device
where Nside is not a power of two), and the rays (the full sky, one RING
batch, or a random pixel list with repeated and invalid entries).

Bars: FP64 final states within 1e-10 of each component's scale and maps within
2 float ulps; mixed dA within 1e-5 relative, shear rms(|d eps|)/rms(|eps|)
<= 1e-4, rotation and deflection rms(d)/rms <= 1e-3; REF dA within 1e-6
relative and the other maps within 2e-5 of their rms; with 0 steps every map
exactly 0; invalid pixel-list entries give NaN maps.

Ten random batch plans (default_rng(500 + case)) check the batched launch
against single traces of the same items, bit for bit.  The cases are fixed by their seeds and the kernels are deterministic, so
the outcome is too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 30


def draw(case):
    rng = np.random.default_rng(100 + case)
    n = int(rng.integers(4, 41))
    nside = int(rng.integers(1, 13))
    steps = int(rng.integers(0, 3 * n + 1))
    observer = None if rng.random() < 0.25 else rng.uniform(-n, 2 * n, 3)
    kind = ["fp64", "mixed_taps", "mixed_records", "ref", "fp64_f16"][case % 5]
    brick = int(rng.choice([1, 2, 4, 8, 16, 32]))
    order = ["linear", "nested"][int(rng.integers(0, 2))]
    npix = 12 * nside * nside
    sel = ["full", "batch", "list"][int(rng.integers(0, 3))]
    batch = int(rng.integers(0, 12)) if sel == "batch" else None
    pix = None
    if sel == "list":
        pix = rng.integers(0, npix, int(rng.integers(1, 2 * npix + 1))).astype(np.int64)
        bad = rng.random(pix.size) < 0.05
        pix[bad] = rng.choice([-1, npix, npix + 7, -3 * npix], int(bad.sum()))
    return dict(n=n, nside=nside, steps=steps, observer=observer, kind=kind, brick=brick, order=order,
                batch=batch, pix=pix, seed=1000 + case)


@pytest.mark.parametrize("case", range(CASES))
def test_random_configuration(srt, oracle, fields, case):
    c = draw(case)
    print(f"[gpu] random case {case}: n {c['n']} nside {c['nside']} steps {c['steps']} kind {c['kind']} "
          f"order {c['order']} batch {c['batch']} list {None if c['pix'] is None else c['pix'].size} "
          f"observer {c['observer']}", flush=True)
    phi, shift = fields.snapshot(c["n"], c["seed"])
    phi, shift = phi.numpy(), shift.numpy()
    tr = srt.Tracer(0)
    if c["kind"] == "mixed_records":
        tr.set_layout(srt.LAYOUT_BICUBIC, c["brick"])
    f16 = c["kind"] == "fp64_f16"
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)),
              shift_format=srt.SHIFT_F16_SCALED if f16 else srt.SHIFT_F32)
    prec = {"ref": srt.PREC_REF, "fp64": srt.PREC_FP64, "fp64_f16": srt.PREC_FP64}.get(c["kind"], srt.PREC_MIXED)
    order = srt.ORDER_NESTED if c["order"] == "nested" else srt.ORDER_LINEAR
    maps, state = tr.trace(c["nside"], c["steps"], batch=c["batch"], pix=c["pix"], observer=c["observer"],
                           precision=prec, order=order, return_state=True)
    tr.close()
    if c["pix"] is not None:
        pix = c["pix"]
    else:
        lo, cnt = srt.batch_range(c["nside"], c["batch"])
        pix = np.arange(lo, lo + cnt)
    valid = (pix >= 0) & (pix < 12 * c["nside"] ** 2)
    assert np.all(np.isnan(maps[:, ~valid]))
    maps, state, pix = maps[:, valid], state[valid], pix[valid]
    if pix.size == 0:
        return
    F = oracle.Fields(phi, shift, b16=oracle.shift_to_f16(shift) if f16 else None)
    form = oracle.FORM_REF if c["kind"] == "ref" else oracle.FORM_FP64
    ref, _, ref_state = oracle.trace(F, c["nside"], pix, c["steps"], observer=c["observer"], form=form,
                                     return_state=True)
    if c["steps"] == 0:
        assert np.all(maps == 0) and np.all(ref == 0)
        return
    if c["kind"] in ("fp64", "fp64_f16"):
        scale = np.abs(ref_state).max(axis=0)
        scale[scale == 0] = 1.0
        assert (np.abs(state - ref_state) / scale).max() < 1e-10
        np.testing.assert_allclose(maps, ref, rtol=2.5e-7, atol=1e-30)
        return
    if c["kind"] == "ref":
        assert np.max(np.abs(maps[0].astype(np.float64) - ref[0]) / np.abs(ref[0])) < 1e-6
        for m in range(1, 6):
            rms = np.sqrt(np.mean(ref[m].astype(np.float64) ** 2))
            if rms > 0:
                assert np.abs(maps[m].astype(np.float64) - ref[m]).max() / rms < 2e-5, m
        return
    assert np.max(np.abs(maps[0].astype(np.float64) - ref[0]) / np.abs(ref[0])) <= 1e-5
    de = np.hypot(maps[1].astype(np.float64) - ref[1], maps[2].astype(np.float64) - ref[2])
    eps = np.sqrt(np.mean(ref[1].astype(np.float64) ** 2 + ref[2].astype(np.float64) ** 2))
    if eps > 0:
        assert np.sqrt(np.mean(de ** 2)) / eps <= 1e-4
    for m in (3, 4, 5):
        rms = np.sqrt(np.mean(ref[m].astype(np.float64) ** 2))
        if rms > 0:
            assert np.sqrt(np.mean((maps[m].astype(np.float64) - ref[m]) ** 2)) / rms <= 1e-3, m


PLANS = 10


def draw_plan(case):
    rng = np.random.default_rng(500 + case)
    contexts = [dict(n=int(rng.integers(4, 41)), records=bool(rng.random() < 0.6), brick=int(rng.choice([1, 4, 16])),
                     seed=2000 + 10 * case + k) for k in range(int(rng.integers(1, 5)))]
    items = []
    for _ in range(int(rng.integers(1, 9))):
        k = int(rng.integers(0, len(contexts)))
        nside = int(rng.integers(1, 13))
        npix = 12 * nside * nside
        count = int(rng.integers(0, npix + 1))
        pix_begin = int(rng.integers(0, npix - count + 1))
        n = contexts[k]["n"]
        items.append(dict(ctx=k, nside=nside, pix_begin=pix_begin, count=count,
                          n_steps=int(rng.integers(0, 3 * n + 1)),
                          observer=None if rng.random() < 0.3 else rng.uniform(-n, 2 * n, 3)))
    precision = ["mixed", "fp64", "ref"][case % 3]
    schedule = int(rng.integers(0, 2))
    order = ["linear", "nested"][int(rng.integers(0, 2))]
    groups = int(rng.choice([1, 3, 8]))
    return dict(contexts=contexts, items=items, precision=precision, schedule=schedule, order=order,
                groups=groups)


@pytest.mark.parametrize("case", range(PLANS))
def test_random_plan(srt, fields, case):
    """A random plan: 1-4 snapshots (edges 4-40, bicubic records on some), 1-8 items over them (several items may
    share a snapshot; arbitrary RING ranges, empty ones included; 0 to 3n steps; observers anywhere), FP64, mixed
    or REF, either schedule, LINEAR or NESTED lanes, 1, 3 or 8 tile ranges, launched directly and replayed from a
    HIP graph: every item equals its own srt_trace bit for bit (the plan runs one launch per kernel its items need:
    record or padded-cube kernel by layout, power-of-two or general grid by edge)."""
    import torch
    p = draw_plan(case)
    print(f"[gpu] random plan {case}: {len(p['contexts'])} snapshots {[c['n'] for c in p['contexts']]}, "
          f"{len(p['items'])} items, {p['precision']}, schedule {p['schedule']}, {p['order']}, "
          f"{p['groups']} groups", flush=True)
    dev = torch.device("cuda:0")
    tracers = []
    for c in p["contexts"]:
        phi, shift = fields.snapshot(c["n"], c["seed"])
        tr = srt.Tracer(0)
        if c["records"]:
            tr.set_layout(srt.LAYOUT_BICUBIC, c["brick"])
        tr.upload(phi.numpy(), fields.interleave_shift(shift).numpy())
        tracers.append(tr)
    prec = srt.PRECISIONS[p["precision"]]
    items = [dict(tracer=tracers[it["ctx"]], nside=it["nside"], pix_begin=it["pix_begin"], count=it["count"],
                  n_steps=it["n_steps"], observer=it["observer"],
                  out=torch.full((6 * max(it["count"], 1),), float("nan"), dtype=torch.float32, device=dev))
             for it in p["items"]]
    want = []
    for it in items:
        o = torch.empty_like(it["out"])
        it["tracer"].trace_device(it["nside"], it["n_steps"], o, pix_begin=it["pix_begin"], count=it["count"],
                                  observer=it["observer"], precision=prec)
        want.append(o)
    torch.cuda.synchronize()
    plan = srt.Plan(items, precision=prec)
    plan.set_schedule(srt.SCHEDULE_PERSISTENT if p["schedule"] == 0 else srt.SCHEDULE_DISPATCH)
    plan.set_order(srt.ORDER_NESTED if p["order"] == "nested" else srt.ORDER_LINEAR, p["groups"])
    stream = torch.cuda.Stream()
    for graph in (False, True, True):
        for it in items:
            it["out"].fill_(float("nan"))
        torch.cuda.synchronize()
        plan.launch(stream=stream, graph=graph)
        stream.synchronize()
        for it, w in zip(items, want):
            c = 6 * it["count"]
            got, exp = it["out"][:c], w[:c]  # 0-step deflections are 0/0 = NaN in both
            same = (got == exp) | (torch.isnan(got) & torch.isnan(exp))
            assert bool(same.all()), (graph, it["nside"], it["pix_begin"], it["count"])
    plan.close()
    for tr in tracers:
        tr.close()
