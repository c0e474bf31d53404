"""Benchmark of the snapshot ray tracer on one MI355X.

Workload (the headline configuration): periodic 512^3 float32 snapshot
(potential rms 1e-5, shift rms 1e-7, P(k) ~ k^-3 with a Gaussian cutoff at half
the Nyquist wavenumber, seed 2), observer at the grid centre, HEALPix Nside
1024 full sky (12,582,912 rays), 256 RK4 steps.  One "step" of the benchmark is
one full-sky trace with the grids already resident in HBM.

Printed: one JSON line with ray-steps/s of the mixed-precision kernel (the
headline `value`), the plain FP64 kernel's rate beside it, the roofline of
the dominant kernel, the per-map error of both kernels against the host FP64
oracle on a seeded 1% pixel subset, and the oracle's own CPU rate.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

# Executed floating-point operations per ray-step of each kernel (FMA = 2),
# from the rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_{F32,F64} counters of
# the 64^3 / Nside 64 / 64-step profile committed under profiles/.
FLOPS_PER_RAY_STEP = {"fp64": None, "mixed": None}
# The reference algorithm's count (oracle with an operation-counting scalar).
REF_ALGO_FLOPS_PER_RAY_STEP = 29294.0
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--grid", type=int, default=512)
    p.add_argument("--nside", type=int, default=1024)
    p.add_argument("--n-steps", type=int, default=256)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--kernels", default="mixed,fp64")
    p.add_argument("--order", default="nested", choices=["nested", "linear"])
    p.add_argument("--subset-frac", type=float, default=0.01)
    p.add_argument("--no-oracle", action="store_true", help="skip the CPU oracle leg (accuracy + cpu baseline)")
    p.add_argument("--cpu-threads", type=int, default=16)
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first run one device copy of 1 GiB (a known-size kernel that scales the "
                        "FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    return p.parse_args()


def map_errors(test, ref):
    import numpy as np
    ref = ref.astype(np.float64)
    test = test.astype(np.float64)
    out = {"dA_max_rel": float(np.max(np.abs(test[0] - ref[0]) / np.abs(ref[0])))}
    de = np.hypot(test[1] - ref[1], test[2] - ref[2])
    out["shear_rms_rel"] = float(np.sqrt(np.mean(de ** 2)) / np.sqrt(np.mean(ref[1] ** 2 + ref[2] ** 2)))
    for m, name in ((1, "ellipticity1"), (2, "ellipticity2"), (3, "rotation"), (4, "deflection1"),
                    (5, "deflection2")):
        rms = float(np.sqrt(np.mean(ref[m] ** 2)))
        d = test[m] - ref[m]
        out[name + "_rms_rel"] = float(np.sqrt(np.mean(d ** 2)) / rms)
        out[name + "_max_rel"] = float(np.max(np.abs(d)) / rms)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the ray tracer shards rays "
              "across GPUs by RING batches with no exchange, and this build measures the one-GPU share.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    pkg = __graft_entry__.load_package()
    fields = __import__(__graft_entry__.PKG_NAME + ".fields", fromlist=["fields"])
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 1 GiB read + 1 GiB written", file=sys.stderr, flush=True)
        del src, dst
    n, nside, n_steps = args.grid, args.nside, args.n_steps
    rays = 12 * nside * nside
    t0 = time.time()
    phi, shift = fields.snapshot(n, args.seed, device=dev)
    shift_il = fields.interleave_shift(shift)
    torch.cuda.synchronize()
    print(f"[bench] snapshot {n}^3 generated in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    tr = pkg.Tracer(0)
    t0 = time.time()
    tr.upload(phi, shift_il)
    torch.cuda.synchronize()
    upload_s = time.time() - t0
    print(f"[bench] upload {upload_s:.2f} s", file=sys.stderr, flush=True)
    del shift_il
    out = torch.empty(6 * rays, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    order = pkg.ORDER_NESTED if args.order == "nested" else pkg.ORDER_LINEAR

    results = {}
    maps_sub = {}
    rng = np.random.default_rng(8)
    subset = np.sort(rng.choice(rays, max(1, int(round(args.subset_frac * rays))), replace=False))
    subset_t = torch.from_numpy(subset).to(dev)
    for name in [k for k in args.kernels.split(",") if k]:
        prec = pkg.PRECISIONS[name]
        for _ in range(args.warmup):
            tr.trace_device(nside, n_steps, out, precision=prec, order=order, stream=stream)
        torch.cuda.synchronize()
        kernel_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tr.trace_device(nside, n_steps, out, precision=prec, order=order, stream=stream)
            kernel_ms.append(tr.last_trace_ms())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms_per_step = 1000.0 * wall / args.steps
        avg_kernel_ms = float(np.mean(kernel_ms))
        results[name] = {
            "ray_steps_per_s": rays * n_steps / (ms_per_step / 1000.0),
            "ms_per_step": ms_per_step,
            "kernel_ms_avg": avg_kernel_ms,
            "kernel_ms_min": float(np.min(kernel_ms)),
        }
        maps_sub[name] = out.view(6, rays)[:, subset_t].cpu().numpy()
        print(f"[bench] {name}: {ms_per_step:.1f} ms/step, kernel {avg_kernel_ms:.1f} ms, "
              f"{results[name]['ray_steps_per_s']:.3e} ray-steps/s", file=sys.stderr, flush=True)

    cpu_baseline = None
    accuracy = {}
    if not args.no_oracle:
        from oracle import oracle
        Fh = oracle.Fields(phi.cpu().numpy(), shift.cpu().numpy())
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        print(f"[bench] oracle fp64 on {subset.size} rays x {n_steps} steps, {threads} threads",
              file=sys.stderr, flush=True)
        ref, sec = oracle.trace(Fh, nside, subset, n_steps, form=oracle.FORM_FP64, nthreads=threads)
        cpu_baseline = {"value": subset.size * n_steps / sec, "unit": "ray-steps/s", "cores": threads,
                        "kind": "port",
                        "sample": f"FP64 host oracle (OpenMP C++) on the seeded 1% pixel subset "
                                  f"({subset.size} rays x {n_steps} steps) of the same snapshot, {sec:.1f} s"}
        for name, m in maps_sub.items():
            accuracy[name] = map_errors(m, ref)

    head = "mixed" if "mixed" in results else next(iter(results))
    r = results[head]
    flops = FLOPS_PER_RAY_STEP.get(head)
    roofline = None
    if flops:
        achieved = flops * rays * n_steps / (r["kernel_ms_avg"] / 1000.0) / 1e12
        roofline = {"bound": "mfma", "compute_unit": "VALU (no matrix work on this path)",
                    "achieved": achieved, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                    "frac": achieved / PEAK_TFLOPS[head], "traffic": None,
                    "flops_per_ray_step": flops, "kernel": f"trace_kernel<{head}>"}
    line = {
        "metric": "ray-steps/s at 512^3, Nside 1024, 256 steps (mixed-precision kernel)",
        "value": r["ray_steps_per_s"],
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 interpolation / f64 state" if head == "mixed" else "f64",
        "data": f"synthetic: seeded P(k)~k^-3 Gaussian fields (seed {args.seed}), generated on the GPU",
        "config": {"workload": f"snapshot {n}^3, Nside {nside} full sky ({rays} rays), {n_steps} RK4 steps, "
                               f"{args.order} lane order", "grid": n, "nside": nside, "n_steps": n_steps,
                   "rays": rays},
        "kernels": results,
        "mixed_over_fp64": (results["mixed"]["ray_steps_per_s"] / results["fp64"]["ray_steps_per_s"]
                            if "mixed" in results and "fp64" in results else None),
        "map_error_vs_fp64_oracle": accuracy,
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "upload_s": upload_s,
    }
    print(json.dumps(line))
    tr.close()


if __name__ == "__main__":
    main()
