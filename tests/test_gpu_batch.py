"""Batched launch (srt_plan_*): many snapshots with different grid sizes,
observers, pixel ranges and step counts in one launch, under the persistent
(work-stealing) and the dispatch (one workgroup per tile) schedules, direct or
replayed from a HIP graph, with an autotuned tile size, grid and schedule.
Every instance must equal its own single-instance srt_trace bit for bit (the
same per-ray code runs in both), on every replay, and the plan's maps must
match the host oracle: FP64 plans within 2 float ulps of the FP64 oracle,
mixed plans within dA 1e-5 / shear 1e-4 of it, REF plans within dA 1e-6 /
2e-5 of each map's rms of the oracle's REF form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances(srt, fields):
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(41)
    out = []
    for i, n in enumerate([16, 24, 32, 20, 12, 28, 32]):
        phi, shift = fields.snapshot(n, 100 + 4 * i)
        tr = srt.Tracer(0)
        tr.upload(phi.to(dev), fields.interleave_shift(shift).to(dev))
        nside = int(rng.choice([2, 4, 8]))
        b = int(rng.integers(0, 12))
        lo, cnt = srt.batch_range(nside, b)
        n_steps = int(rng.integers(4, n // 2 + 1))
        obs = rng.uniform(0, n, 3)
        out.append(dict(tracer=tr, nside=nside, pix_begin=lo, count=cnt, n_steps=n_steps, observer=obs,
                        out=torch.zeros(6 * cnt, dtype=torch.float32, device=dev),
                        host=(phi.numpy(), shift.numpy())))
    return out


def expected(srt, items, precision):
    import torch
    res = []
    for it in items:
        o = torch.empty_like(it["out"])
        it["tracer"].trace_device(it["nside"], it["n_steps"], o, pix_begin=it["pix_begin"], count=it["count"],
                                  observer=it["observer"], precision=precision)
        res.append(o)
    torch.cuda.synchronize()
    return [r.cpu().numpy() for r in res]


def check(items, ref):
    for it, r in zip(items, ref):
        assert np.array_equal(it["out"].cpu().numpy(), r)


def check_oracle(oracle, items, prec):
    """Each item's plan output against the host oracle over its whole pixel range."""
    form = oracle.FORM_REF if prec == "ref" else oracle.FORM_FP64
    for it in items:
        pix = np.arange(it["pix_begin"], it["pix_begin"] + it["count"])
        ref, _ = oracle.trace(oracle.Fields(*it["host"]), it["nside"], pix, it["n_steps"],
                              observer=it["observer"], form=form)
        got = it["out"].view(6, it["count"]).cpu().numpy()
        ref64, got64 = ref.astype(np.float64), got.astype(np.float64)
        da = np.max(np.abs(got64[0] - ref64[0]) / np.abs(ref64[0]))
        if prec == "fp64":
            np.testing.assert_allclose(got, ref, rtol=2.5e-7, atol=1e-30)
        elif prec == "mixed":
            de2 = (got64[1] - ref64[1]) ** 2 + (got64[2] - ref64[2]) ** 2
            assert da <= 1e-5
            assert np.sqrt(np.mean(de2) / np.mean(ref64[1] ** 2 + ref64[2] ** 2)) <= 1e-4
        else:
            assert da < 1e-6
            for m in range(1, 6):
                assert np.abs(got64[m] - ref64[m]).max() / np.sqrt(np.mean(ref64[m] ** 2)) < 2e-5, m


SCHEDULES = {"persistent": 0, "dispatch": 1}


@pytest.mark.parametrize("schedule", ["persistent", "dispatch"])
@pytest.mark.parametrize("prec", ["mixed", "fp64", "ref"])
def test_plan_direct_and_graph_replay(srt, oracle, instances, prec, schedule):
    import torch
    print(f"[gpu] batch plan {prec}, {schedule}", flush=True)
    p = srt.PRECISIONS[prec]
    ref = expected(srt, instances, p)
    plan = srt.Plan(instances, precision=p)
    plan.set_schedule(SCHEDULES[schedule])
    assert plan.schedule() == SCHEDULES[schedule]
    plan.launch()
    torch.cuda.synchronize()
    check(instances, ref)
    check_oracle(oracle, instances, prec)
    s = torch.cuda.Stream()
    for rep in range(3):
        for it in instances:  # eager work between replays
            it["out"].zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.launch(stream=s, graph=True)
        s.synchronize()
        check(instances, ref)
    for tile, blocks in ((64, 1), (128, 7), (1024, 600)):
        plan.configure(tile, blocks)
        for it in instances:
            it["out"].zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.launch(stream=s, graph=True)
        s.synchronize()
        check(instances, ref)
    plan.close()


@pytest.mark.parametrize("schedule", ["persistent", "dispatch"])
@pytest.mark.parametrize("order,groups", [("linear", 1), ("linear", 8), ("nested", 1), ("nested", 3),
                                          ("nested", 64)])
def test_plan_lane_order_and_tile_groups(srt, instances, order, groups, schedule):
    import torch
    print(f"[gpu] batch plan order {order}, {groups} tile groups, {schedule}", flush=True)
    ref = expected(srt, instances, srt.PREC_MIXED)
    plan = srt.Plan(instances, precision=srt.PREC_MIXED)
    plan.set_schedule(SCHEDULES[schedule])
    plan.set_order(srt.ORDER_NESTED if order == "nested" else srt.ORDER_LINEAR, groups)
    s = torch.cuda.Stream()
    for graph in (False, True, True):
        for it in instances:
            it["out"].zero_()
        torch.cuda.synchronize()
        plan.launch(stream=s, graph=graph)
        s.synchronize()
        check(instances, ref)
    plan.configure(64, 3)  # fewer blocks than tile groups: every range is still drained
    for it in instances:
        it["out"].zero_()
    torch.cuda.synchronize()
    plan.launch(stream=s)
    s.synchronize()
    check(instances, ref)
    with pytest.raises(srt.SrtError):
        plan.set_order(srt.ORDER_NESTED, 65)
    plan.close()


def test_plan_autotune(srt, instances):
    import torch
    ref = expected(srt, instances, srt.PREC_MIXED)
    plan = srt.Plan(instances, precision=srt.PREC_MIXED)
    tile, blocks = plan.autotune()
    assert tile % 64 == 0 and blocks > 0
    assert plan.schedule() in SCHEDULES.values()
    assert plan.config()[:2] == (tile, blocks)
    assert plan.config()[2] == sum(it["count"] * it["n_steps"] for it in instances)
    for it in instances:
        it["out"].zero_()
    plan.launch()
    torch.cuda.synchronize()
    check(instances, ref)
    assert plan.last_ms() > 0
    plan.close()


def test_plan_errors(srt, instances):
    import torch
    with pytest.raises(srt.SrtError):
        srt.Plan([dict(instances[0], count=10 ** 9)])
    plan = srt.Plan(instances[:2])
    with pytest.raises(srt.SrtError):
        plan.configure(100, 4)  # not a multiple of 64
    with pytest.raises(srt.SrtError):
        plan.launch(graph=True)  # graph replay needs a non-default stream
    with pytest.raises(srt.SrtError):
        plan.set_schedule(2)
    plan.close()


@pytest.mark.parametrize("tile,blocks", [(64, 1), (64, 5), (64, 512), (128, 37), (256, 2048), (0, 0)])
def test_plan_many_tiles_per_workgroup(srt, fields, tile, blocks):
    """One full-sky item and the 12 RING batches of a second sky as 13 items of
    one plan: thousands of 64-lane groups drawn from each workgroup's tile slot
    (one workgroup draws them all with blocks=1).  Every ray is traced exactly
    once: outputs start as NaN and must equal srt_trace's bit for bit."""
    import torch
    print(f"[gpu] batch plan many tiles: tile {tile}, blocks {blocks} (0: dispatch schedule)", flush=True)
    dev = torch.device("cuda:0")
    phi, shift = fields.snapshot(32, 77)
    tr = srt.Tracer(0)
    tr.upload(phi.to(dev), fields.interleave_shift(shift).to(dev))
    nside, n_steps = 16, 6
    rays = 12 * nside * nside
    full = torch.empty(6 * rays, dtype=torch.float32, device=dev)
    tr.trace_device(nside, n_steps, full, precision=srt.PREC_MIXED)
    torch.cuda.synchronize()
    ref_full = full.cpu().numpy()
    items = [dict(tracer=tr, nside=nside, pix_begin=0, count=rays, n_steps=n_steps,
                  out=torch.full((6 * rays,), float("nan"), dtype=torch.float32, device=dev))]
    per = nside * nside
    ref_b = []
    for b in range(12):
        o = torch.empty(6 * per, dtype=torch.float32, device=dev)
        tr.trace_device(nside, n_steps + 1, o, pix_begin=b * per, count=per, precision=srt.PREC_MIXED)
        ref_b.append(o)
        items.append(dict(tracer=tr, nside=nside, pix_begin=b * per, count=per, n_steps=n_steps + 1,
                          out=torch.full((6 * per,), float("nan"), dtype=torch.float32, device=dev)))
    torch.cuda.synchronize()
    plan = srt.Plan(items, precision=srt.PREC_MIXED)
    if tile:
        plan.set_schedule(srt.SCHEDULE_PERSISTENT)
        plan.configure(tile, blocks)
    else:
        plan.set_schedule(srt.SCHEDULE_DISPATCH)
    s = torch.cuda.Stream()
    for rep in range(3):
        for it in items:
            it["out"].fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.launch(stream=s, graph=rep > 0)
        s.synchronize()
        assert np.array_equal(items[0]["out"].cpu().numpy(), ref_full)
        for it, r in zip(items[1:], ref_b):
            assert np.array_equal(it["out"].cpu().numpy(), r.cpu().numpy())
    plan.close()


def _tracer(srt, fields, n, seed, layout=None, brick=16):
    import torch
    phi, shift = fields.snapshot(n, seed)
    tr = srt.Tracer(0)
    if layout is not None:
        tr.set_layout(layout, brick)
    tr.upload(phi.to("cuda:0"), fields.interleave_shift(shift).to("cuda:0"))
    return tr


def _items(tracers):
    import torch
    a, b = tracers
    return [dict(tracer=a, nside=4, pix_begin=0, count=192, n_steps=6, observer=None,
                 out=torch.zeros(6 * 192, dtype=torch.float32, device="cuda:0")),
            dict(tracer=b, nside=8, pix_begin=3 * 64, count=64, n_steps=9, observer=(9.5, 20.25, 11.0),
                 out=torch.zeros(6 * 64, dtype=torch.float32, device="cuda:0")),
            dict(tracer=a, nside=8, pix_begin=700, count=68, n_steps=5, observer=(3.0, 4.0, 5.0),
                 out=torch.zeros(6 * 68, dtype=torch.float32, device="cuda:0"))]


@pytest.mark.parametrize("schedule", ["persistent", "dispatch"])
@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_plan_follows_reupload_and_layout_changes(srt, fields, prec, schedule):
    """A plan built before a context gets a new snapshot (of another size) or
    another layout traces the context's current buffers at its next launch,
    direct or from a graph: the same outputs as fresh single traces, also while
    the contexts of a mixed plan differ in layout (the plan then runs the record
    kernel and the padded-cube kernel as two launches)."""
    import torch
    print(f"[gpu] batch plan after re-upload / layout change, {prec}, {schedule}", flush=True)
    p = srt.PRECISIONS[prec]
    a = _tracer(srt, fields, 24, 301, srt.LAYOUT_BICUBIC)
    b = _tracer(srt, fields, 32, 302, srt.LAYOUT_BICUBIC)
    items = _items((a, b))
    plan = srt.Plan(items, precision=p)
    plan.set_schedule(SCHEDULES[schedule])
    s = torch.cuda.Stream()

    def run_and_check(graph):
        ref = expected(srt, items, p)
        for rep in range(2):
            for it in items:
                it["out"].fill_(float("nan"))
            torch.cuda.synchronize()
            with torch.cuda.stream(s):
                plan.launch(stream=s, graph=graph)
            s.synchronize()
            check(items, ref)

    run_and_check(False)
    run_and_check(True)
    phi, shift = fields.snapshot(20, 303)  # new snapshot, new grid edge: every buffer of `a` is replaced
    a.upload(phi.to("cuda:0"), fields.interleave_shift(shift).to("cuda:0"))
    run_and_check(True)
    b.set_layout(srt.LAYOUT_BICUBIC, 4)  # records rebuilt in 4 x 4 bricks
    run_and_check(True)
    run_and_check(False)
    # records of `b` dropped: the contexts differ in layout (for a mixed plan,
    # two launches), then `a` drops its records too
    b.set_layout(srt.LAYOUT_TAPS)
    run_and_check(True)
    run_and_check(False)
    a.set_layout(srt.LAYOUT_TAPS)
    run_and_check(True)
    run_and_check(False)
    plan.close()
    a.close()
    b.close()


def test_plan_autotune_brick(srt, fields):
    import torch
    print("[gpu] batch plan brick autotune", flush=True)
    a = _tracer(srt, fields, 16, 311, srt.LAYOUT_BICUBIC, 1)
    b = _tracer(srt, fields, 32, 312, srt.LAYOUT_BICUBIC, 8)
    items = _items((a, b))
    plan = srt.Plan(items, precision=srt.PREC_MIXED)
    brick = plan.autotune_brick()
    assert brick in (1, 4, 8, 16)  # edges above the smallest grid edge (16) are not tried
    for tr in (a, b):
        lay, br, nbytes = tr.layout()
        assert (lay, br) == (srt.LAYOUT_BICUBIC, brick) and nbytes > 0
    ref = expected(srt, items, srt.PREC_MIXED)
    for it in items:
        it["out"].fill_(float("nan"))
    plan.launch()
    torch.cuda.synchronize()
    check(items, ref)
    plan.close()
    fp64 = srt.Plan(items, precision=srt.PREC_FP64)
    with pytest.raises(srt.SrtError):
        fp64.autotune_brick()  # only mixed plans read records
    fp64.close()
    a.close()
    b.close()


def test_plan_keeps_its_tracers_alive(srt, fields):
    """A plan built from tracers that the caller no longer references still traces their snapshots:
    the plan holds the tracers, so their contexts (and device grids) are not destroyed under it."""
    import gc
    import torch
    print("[gpu] batch plan with out-of-scope tracers", flush=True)

    def make():
        items, refs = [], []
        for i, n in enumerate((16, 24)):
            phi, shift = fields.snapshot(n, 321 + i)
            tr = srt.Tracer(0)
            tr.upload(phi.to("cuda:0"), fields.interleave_shift(shift).to("cuda:0"))
            it = dict(tracer=tr, nside=4, pix_begin=0, count=192, n_steps=5 + i,
                      out=torch.full((6 * 192,), float("nan"), dtype=torch.float32, device="cuda:0"))
            ref = torch.empty_like(it["out"])
            tr.trace_device(4, it["n_steps"], ref, precision=srt.PREC_MIXED)
            items.append(it)
            refs.append(ref)
        plan = srt.Plan(items, precision=srt.PREC_MIXED)
        return plan, [it["out"] for it in items], refs

    plan, outs, refs = make()
    gc.collect()
    for _ in range(2):
        for o in outs:
            o.fill_(float("nan"))
        plan.launch()
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            assert torch.equal(o, r)
    plan.close()
