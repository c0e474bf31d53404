"""Writes tests/golden/oracle_16cube.npz: small fixed inputs and the host oracle's results for them.

The reference repository holds no test vectors, so these fixtures are the oracle's own results on seeded inputs.
They hold the oracle to its present results (tests/test_golden.py), and they let the GPU tests check every kernel
form against stored numbers without running the oracle (tests/test_gpu_golden.py).

Inputs: a 16^3 potential and shift from snapshot_raytracer_amd.fields (seed 31, the benchmark's k^-3 spectrum,
rms 1e-5 and 1e-7), stored as float32, so the fixtures do not depend on the generator's floating-point path.
Cases:
  fp64      FP64 form, every RING pixel of Nside 2 (48 rays), 12 steps, observer at the grid centre
  fp64_obs  FP64 form, the same rays, 20 steps, observer (3.25, 12.5, 7.75)
  fp64_list FP64 form, Nside 3, the pixel list [0, 107, 53, 53, 12, 71], 9 steps (a repeated pixel)
  ref_f16   REF form (the reference's float/double mix) on the binary16 x 2^27 shift codes, Nside 2, 12 steps
Each case stores its pixels, steps and observer, the maps (6, rays) float32 and the final states (rays, 19).

Run from the repository root after building the oracle:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

N, SEED = 16, 31
CASES = {
    "fp64": dict(form="fp64", nside=2, pix=np.arange(48), steps=12, observer=None),
    "fp64_obs": dict(form="fp64", nside=2, pix=np.arange(48), steps=20, observer=(3.25, 12.5, 7.75)),
    "fp64_list": dict(form="fp64", nside=3, pix=np.array([0, 107, 53, 53, 12, 71]), steps=9, observer=None),
    "ref_f16": dict(form="ref_f16", nside=2, pix=np.arange(48), steps=12, observer=None),
}


def main():
    __graft_entry__.load_package()
    from oracle import oracle
    from snapshot_raytracer_amd import fields
    phi, shift = fields.snapshot(N, SEED)
    phi, shift = phi.numpy().astype(np.float32), shift.numpy().astype(np.float32)
    out = {"phi": phi, "shift": shift}
    for name, c in CASES.items():
        f16 = c["form"] == "ref_f16"
        F = oracle.Fields(phi, shift, b16=oracle.shift_to_f16(shift) if f16 else None)
        obs = c["observer"] if c["observer"] is not None else (0.5 * N,) * 3
        maps, _, state = oracle.trace(F, c["nside"], c["pix"], c["steps"], observer=obs,
                                      form=oracle.FORM_REF if f16 else oracle.FORM_FP64, return_state=True)
        out[f"{name}_nside"] = np.int64(c["nside"])
        out[f"{name}_pix"] = c["pix"].astype(np.int64)
        out[f"{name}_steps"] = np.int64(c["steps"])
        out[f"{name}_observer"] = np.asarray(obs, np.float64)
        out[f"{name}_maps"] = maps
        out[f"{name}_state"] = state
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_16cube.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
