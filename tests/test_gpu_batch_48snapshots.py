"""Batched launch at its full workload: 48 independent snapshot/observer
instances in one plan, checked against single traces bit for bit and against
the host FP64 oracle.

Workload (fields.uneven_batch_params: numpy default_rng(7), instance i with
snapshot seed 1000 + 4 i): grid edges from {128, 192, 256, 384} (potential rms
1e-5, shift rms 1e-7, P(k) ~ k^-3), observers uniform in the box, n_steps
uniform in [64, n / 2], Nside from {128, 256, 512}, one random RING batch each;
mixed precision on bicubic records (16 x 16 column bricks) when the 48
instances' records fit in device memory, else on the padded cubes.

* The plan under the dispatch schedule (the default), the persistent
  schedule, and replayed from a HIP graph three times with eager work in
  between: every instance's maps equal its own srt_trace bit for bit.
* A seeded subset of 128 pixels per instance (numpy default_rng(2000 + i)
  over the instance's batch) against the FP64 oracle: per instance, dA max
  relative error <= 1e-5 and shear rms(|d eps|) / rms(|eps|) <= 1e-4; over
  the pooled 6,144 pixels, rotation and deflection rms(d) / rms <= 1e-3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SUBSET = 128


@pytest.fixture(scope="module")
def c5(srt, fields, oracle, progress):
    import torch
    torch.cuda.empty_cache()
    dev = torch.device("cuda:0")
    params = fields.uneven_batch_params()
    rec_bytes = sum(4 * 64 * q["n"] ** 2 * (q["n"] + 3) for q in params)
    grid_bytes = sum(4 * 4 * (q["n"] + 3) ** 3 for q in params)
    free, _ = torch.cuda.mem_get_info()
    bicubic = rec_bytes + grid_bytes + 8e9 < free
    progress(f"[gpu] 48 snapshots: records {rec_bytes / 1e9:.0f} GB + grids {grid_bytes / 1e9:.0f} GB, free "
             f"{free / 1e9:.0f} GB -> {'bicubic records' if bicubic else 'padded cubes'}")
    items, refs = [], []
    for i, q in enumerate(params):
        phi, shift = fields.snapshot(q["n"], q["seed"], device=dev)
        tr = srt.Tracer(0)
        if bicubic:
            tr.set_layout(srt.LAYOUT_BICUBIC, 16)
        tr.upload(phi, fields.interleave_shift(shift))
        lo, cnt = srt.batch_range(q["nside"], q["batch"])
        pix = np.sort(np.random.default_rng(2000 + i).choice(cnt, SUBSET, replace=False)) + lo
        ref, _ = oracle.trace(oracle.Fields(phi.cpu().numpy(), shift.cpu().numpy()), q["nside"], pix,
                              q["n_steps"], observer=q["observer"])
        del phi, shift
        items.append(dict(tracer=tr, nside=q["nside"], pix_begin=lo, count=cnt, n_steps=q["n_steps"],
                          observer=q["observer"], out=torch.zeros(6 * cnt, dtype=torch.float32, device=dev)))
        refs.append((pix - lo, ref))
        if i % 8 == 7:
            progress(f"[gpu] 48 snapshots: {i + 1} generated, uploaded and oracle-traced")
    torch.cuda.empty_cache()
    single = []
    for it in items:
        o = torch.empty_like(it["out"])
        it["tracer"].trace_device(it["nside"], it["n_steps"], o, pix_begin=it["pix_begin"], count=it["count"],
                                  observer=it["observer"], precision=srt.PREC_MIXED, order=srt.ORDER_NESTED)
        single.append(o)
    torch.cuda.synchronize()
    yield dict(items=items, refs=refs, single=single, bicubic=bicubic)
    for it in items:
        it["tracer"].close()
    del items, single
    torch.cuda.empty_cache()


def equal_to_single(c5):
    import torch
    return [torch.equal(it["out"], s) for it, s in zip(c5["items"], c5["single"])]


def test_48_snapshots_plan_schedules_and_graph_replay(srt, c5, progress):
    import torch
    items = c5["items"]
    plan = srt.Plan(items, precision=srt.PREC_MIXED)
    assert plan.schedule() == srt.SCHEDULE_DISPATCH
    progress("[gpu] 48 snapshots: dispatch plan")
    plan.launch()
    torch.cuda.synchronize()
    assert all(equal_to_single(c5))
    progress(f"[gpu] 48 snapshots: dispatch plan {plan.last_ms():.1f} ms, equal to 48 single traces")
    plan.set_schedule(srt.SCHEDULE_PERSISTENT)
    for it in items:
        it["out"].zero_()
    plan.launch()
    torch.cuda.synchronize()
    assert all(equal_to_single(c5))
    progress(f"[gpu] 48 snapshots: persistent plan {plan.last_ms():.1f} ms, equal")
    plan.set_schedule(srt.SCHEDULE_DISPATCH)
    s = torch.cuda.Stream()
    for rep in range(3):
        for it in items:  # eager work between replays
            it["out"].zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.launch(stream=s, graph=True)
        s.synchronize()
        assert all(equal_to_single(c5)), rep
    progress(f"[gpu] 48 snapshots: graph replays {plan.last_ms():.1f} ms, equal")
    plan.close()


def test_48_snapshots_plan_matches_fp64_oracle(srt, c5):
    import torch
    items = c5["items"]
    plan = srt.Plan(items, precision=srt.PREC_MIXED)
    for it in items:
        it["out"].fill_(float("nan"))
    plan.launch()
    torch.cuda.synchronize()
    plan.close()
    pooled_t, pooled_r = [], []
    worst_da, worst_shear = 0.0, 0.0
    for i, (it, (idx, ref)) in enumerate(zip(items, c5["refs"])):
        got = it["out"].view(6, it["count"]).cpu().numpy()[:, idx].astype(np.float64)
        ref = ref.astype(np.float64)
        da = float(np.max(np.abs(got[0] - ref[0]) / np.abs(ref[0])))
        de2 = (got[1] - ref[1]) ** 2 + (got[2] - ref[2]) ** 2
        shear = float(np.sqrt(np.mean(de2) / np.mean(ref[1] ** 2 + ref[2] ** 2)))
        worst_da, worst_shear = max(worst_da, da), max(worst_shear, shear)
        assert da <= 1e-5, (i, da)
        assert shear <= 1e-4, (i, shear)
        pooled_t.append(got)
        pooled_r.append(ref)
    t, r = np.concatenate(pooled_t, axis=1), np.concatenate(pooled_r, axis=1)
    rms = {}
    for m, name in ((3, "rotation"), (4, "deflection1"), (5, "deflection2")):
        rms[name] = float(np.sqrt(np.mean((t[m] - r[m]) ** 2)) / np.sqrt(np.mean(r[m] ** 2)))
        assert rms[name] <= 1e-3, name
    print(f"48 snapshots ({'records' if c5['bicubic'] else 'padded cubes'}) vs FP64 oracle on "
          f"{t.shape[1]} pixels: worst dA max rel {worst_da:.2e}, worst shear rms {worst_shear:.2e}, {rms}")
