"""The host oracle against its committed fixtures (tests/golden/oracle_16cube.npz, written by
tests/golden/make_golden.py): on the stored 16^3 inputs every case reproduces the stored maps and final states,
so a change to the oracle's arithmetic shows up here before it moves any GPU parity bar."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_16cube.npz")
CASES = ("fp64", "fp64_obs", "fp64_list", "ref_f16")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_fixture(oracle, golden, case):
    g = golden
    f16 = case == "ref_f16"
    F = oracle.Fields(g["phi"], g["shift"], b16=oracle.shift_to_f16(g["shift"]) if f16 else None)
    maps, _, state = oracle.trace(F, int(g[f"{case}_nside"]), g[f"{case}_pix"], int(g[f"{case}_steps"]),
                                  observer=g[f"{case}_observer"], form=oracle.FORM_REF if f16 else oracle.FORM_FP64,
                                  return_state=True)
    scale = np.abs(g[f"{case}_state"]).max(axis=0)
    scale[scale == 0] = 1.0
    assert (np.abs(state - g[f"{case}_state"]) / scale).max() <= 1e-13
    np.testing.assert_allclose(maps, g[f"{case}_maps"], rtol=1.2e-7, atol=0)


def test_fixture_shapes(golden):
    assert golden["phi"].shape == (16, 16, 16) and golden["shift"].shape == (3, 16, 16, 16)
    for case in CASES:
        n = golden[f"{case}_pix"].size
        assert golden[f"{case}_maps"].shape == (6, n) and golden[f"{case}_state"].shape == (n, 19)
        assert np.all(np.isfinite(golden[f"{case}_maps"]))
