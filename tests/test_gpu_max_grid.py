"""The largest grid the library accepts (2048^3, padded cubes of 2051^3: every
voxel offset beyond 2^32), checked against a 128^3 grid holding the same
periodic field.

The field has period 128 along every axis, so the 2048^3 grid is the 128^3
grid tiled 16 x 16 x 16 with bit-identical values.  Rays from observers that
differ by a multiple of 128 cells then see the same field; the observer of
the large grid sits near its far corner, so the rays cross its periodic
boundary and read voxels at the largest offsets.  The two traces differ only
in the rounding of positions of different magnitude (FP64, ~1e-13 cells),
which now and then moves an FP32 cell offset by one ulp; the mixed kernel's
maps then agree within 2e-5 of each map's rms (measured: dA identical, the
ellipticities within 2e-6), where a wrong voxel offset would change them by
order one.

At this size the bicubic records (4.4 TB) and the FP64 grid copies do not
fit: setting the record layout and an FP64 trace fail with
SRT_ERR_OUT_OF_MEMORY, and the context keeps tracing from the padded cubes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PERIOD, BIG, NSIDE, STEPS = 128, 2048, 8, 48
OBS_SMALL = np.array([100.37, 99.61, 110.29])
OBS_BIG = OBS_SMALL + 15 * PERIOD  # 2020.37, 2019.61, 2030.29: 48 steps reach past the far faces
ERR_OUT_OF_MEMORY = 3


@pytest.fixture(scope="module")
def small(srt, fields):
    tr = srt.Tracer(0)
    phi, shift = fields.periodic_field(PERIOD, PERIOD, "cuda:0")
    tr.upload(phi, shift)
    yield tr
    tr.close()


def test_max_grid_matches_tiled_small_grid(srt, small, fields):
    import torch
    torch.cuda.empty_cache()
    free, total = torch.cuda.mem_get_info()
    if free < 290e9:
        pytest.skip(f"needs ~280 GB of free device memory, {free / 1e9:.0f} GB free")
    print(f"[gpu] {BIG}^3 periodic field: generate, upload", flush=True)
    big = srt.Tracer(0)
    phi, shift = fields.periodic_field(BIG, PERIOD, "cuda:0")
    big.upload(phi, shift)
    del phi, shift
    torch.cuda.empty_cache()
    try:
        with pytest.raises(srt.SrtError) as e:
            big.set_layout(srt.LAYOUT_BICUBIC, 16)
        assert e.value.status == ERR_OUT_OF_MEMORY
        assert big.layout()[0] == srt.LAYOUT_TAPS
        print("[gpu] max grid: mixed traces", flush=True)
        ref = small.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, observer=OBS_SMALL)
        got = big.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, observer=OBS_BIG)
        got_nested = big.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, observer=OBS_BIG, order=srt.ORDER_NESTED)
        assert np.array_equal(got, got_nested)
        assert np.all(np.isfinite(got))
        # rays moved ~48 cells from a corner observer: they crossed the far faces of the large grid
        assert np.all(OBS_BIG + STEPS > BIG)
        rel = []
        for m in range(6):
            rms = float(np.sqrt(np.mean(ref[m].astype(np.float64) ** 2)))
            rel.append(float(np.max(np.abs(got[m].astype(np.float64) - ref[m]))) / rms)
        print("max |big - small| / rms per map:", [f"{r:.2e}" for r in rel])
        assert max(rel) <= 2e-5
        print("[gpu] max grid: FP64 trace must report out of memory", flush=True)
        with pytest.raises(srt.SrtError) as e:
            big.trace(NSIDE, 2, precision=srt.PREC_FP64, observer=OBS_BIG)
        assert e.value.status == ERR_OUT_OF_MEMORY
        again = big.trace(NSIDE, STEPS, precision=srt.PREC_MIXED, observer=OBS_BIG)
        assert np.array_equal(again, got)
    finally:
        big.close()
        torch.cuda.empty_cache()
