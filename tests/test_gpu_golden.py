"""Every kernel form against the committed oracle fixtures (tests/golden/oracle_16cube.npz, written by
tests/golden/make_golden.py), through the C ABI and without running the oracle, in both lane orders:
  * FP64: final states within 1e-10 of each component's scale, maps within 2 float ulps;
  * mixed, on the padded grids and on bicubic records (brick edge 4): within the mixed bars of the FP64 maps
    (dA 1e-5 relative, shear rms(|d eps|)/rms(|eps|) 1e-4, rotation and deflection rms 1e-3);
  * REF on the binary16 x 2^27 shift storage: within the REF bars of the oracle's REF form on the same codes
    (dA 1e-6 relative, other maps 2e-5 of their rms)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_16cube.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def run(srt, g, case, precision, records=False, f16=False, order=0):
    print(f"[gpu] golden {case}: precision {precision} records {records} f16 {f16} order {order}", flush=True)
    tr = srt.Tracer(0)
    if records:
        tr.set_layout(srt.LAYOUT_BICUBIC, 4)
    tr.upload(g["phi"], np.ascontiguousarray(np.moveaxis(g["shift"], 0, -1)),
              shift_format=srt.SHIFT_F16_SCALED if f16 else srt.SHIFT_F32)
    maps, state = tr.trace(int(g[f"{case}_nside"]), int(g[f"{case}_steps"]), pix=g[f"{case}_pix"],
                           observer=g[f"{case}_observer"], precision=precision, order=order, return_state=True)
    tr.close()
    return maps, state


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("case", ["fp64", "fp64_obs", "fp64_list"])
def test_fp64_kernel_matches_fixture(srt, golden, case, order):
    maps, state = run(srt, golden, case, srt.PREC_FP64, order=order)
    ref_state = golden[f"{case}_state"]
    scale = np.abs(ref_state).max(axis=0)
    scale[scale == 0] = 1.0
    assert (np.abs(state - ref_state) / scale).max() < 1e-10
    np.testing.assert_allclose(maps, golden[f"{case}_maps"], rtol=2.5e-7, atol=1e-30)


@pytest.mark.parametrize("records", [False, True])
@pytest.mark.parametrize("case", ["fp64", "fp64_obs", "fp64_list"])
def test_mixed_kernels_within_bars_of_fixture(srt, golden, case, records):
    maps, _ = run(srt, golden, case, srt.PREC_MIXED, records=records, order=1)
    m, r = maps.astype(np.float64), golden[f"{case}_maps"].astype(np.float64)
    assert np.max(np.abs(m[0] - r[0]) / np.abs(r[0])) <= 1e-5
    de = np.hypot(m[1] - r[1], m[2] - r[2])
    assert np.sqrt(np.mean(de ** 2)) / np.sqrt(np.mean(r[1] ** 2 + r[2] ** 2)) <= 1e-4
    for k in (3, 4, 5):
        assert np.sqrt(np.mean((m[k] - r[k]) ** 2)) / np.sqrt(np.mean(r[k] ** 2)) <= 1e-3, k


@pytest.mark.parametrize("order", [0, 1])
def test_ref_kernel_on_binary16_storage_matches_fixture(srt, golden, order):
    maps, _ = run(srt, golden, "ref_f16", srt.PREC_REF, f16=True, order=order)
    m, r = maps.astype(np.float64), golden["ref_f16_maps"].astype(np.float64)
    assert np.max(np.abs(m[0] - r[0]) / np.abs(r[0])) < 1e-6
    for k in range(1, 6):
        assert np.max(np.abs(m[k] - r[k])) / np.sqrt(np.mean(r[k] ** 2)) < 2e-5, k
