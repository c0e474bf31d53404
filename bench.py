"""Benchmark of the snapshot ray tracer on one MI355X.

Workload (the headline configuration): periodic 512^3 float32 snapshot
(potential rms 1e-5, shift rms 1e-7, P(k) ~ k^-3 with a Gaussian cutoff at half
the Nyquist wavenumber, seed 2), observer at the grid centre, HEALPix Nside
1024 full sky (12,582,912 rays), 256 RK4 steps.  One "step" of the benchmark is
one full-sky trace with the grids already resident in HBM.

Printed: one JSON line with ray-steps/s of the mixed-precision kernel (the
headline `value`, reading the bicubic cell records), the plain FP64 kernel's
rate and the mixed kernel on the raw padded cubes beside it, the roofline of
the dominant kernel, the per-map error of every kernel against the host FP64
oracle on a seeded 1% pixel subset, the full-sky dA and shear errors of the
other kernels against the FP64 kernel, and the oracle's own CPU rate.
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

# Floating-point operations per ray-step executed by each kernel (FMA = 2),
# from rocprofv3 counters SQ_INSTS_VALU_FLOPS_FP32 + SQ_INSTS_VALU_FLOPS_FP64
# (per wavefront instruction, x 64 lanes) of one launch of this workload,
# divided by its 3,221,225,472 ray-steps (profiles/trace_kernel_512cube_nside1024_counters.md).
FLOPS_PER_RAY_STEP = {"mixed": 8293.0, "mixed_taps": 12133.0, "fp64": 12793.0}
# HBM bytes per launch of each kernel: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# (the counters scaled by what they read for a 1 GiB device copy), same profile.
HBM_BYTES_PER_LAUNCH = {"mixed": 28.46e9, "mixed_taps": 110.2e9, "fp64": 8.75e9}
# The gather path per launch of each kernel, from rocprofv3 TCP, TA and TCC
# counters of one launch of this workload (same profile): L1 (TCP) cache-line
# tag accesses (TCP_TOTAL_CACHE_ACCESSES_sum), L2 read bytes
# (TCC_READ_SECTORS_sum x 32 B) and the share of cycles the texture-address
# unit was busy (TA_BUSY_avr over GRBM_GUI_ACTIVE / 8 XCDs).
GATHER_PER_LAUNCH = {"mixed": {"l1_lines": 2.7518e11, "l2_read_bytes": 6.5088e11, "ta_busy": 0.779},
                     "mixed_taps": {"l1_lines": 3.2305e11, "l2_read_bytes": 1.38832e12, "ta_busy": 0.724},
                     "fp64": {"l1_lines": 5.6283e11, "l2_read_bytes": 3.7126e11, "ta_busy": 0.589}}
# L2 read peak: 8 XCDs x 16 channels (TCC_CYCLE_sum / cycles = 128 channels) x 128 B per channel and clock at
# 2.4 GHz; the CUs and the clock of the L1 rate
L2_PEAK_TBPS = 128 * 128 * 2.4e9 / 1e12
CUS, CLOCK_HZ = 256, 2.4e9
# Algorithmic flops per ray-step, the roofline's work constant: the operation
# count of a work-efficient form of the reference's mathematics (per axis the
# Catmull-Rom weight vectors once, weight sets shared between the four fields of
# a right-hand side, the 4 x 4 x 4 taps contracted x -> y -> z; the rest of the
# right-hand side and RK4 as the reference has them), counted by running the FP64
# oracle with an operation-counting scalar (oracle/opcount.cpp, variant 2) on
# Nside 16, RING pixel 1000, 32 steps of the seed-1 64^3 snapshot.
ALGO_FLOPS_PER_RAY_STEP = 12290.5
# The same count for the reference's own form (three expanded Catmull-Rom
# polynomials per row and pass, .cu:195-312), for comparison.
REF_FORM_FLOPS_PER_RAY_STEP = 29298.5
# MI355X published vector peaks (FP32 counts packed FP32; the library disables
# packed FP32, so 78.6 TFLOP/s is the FP32 rate of the instructions it issues).
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3, "mixed_taps": 157.3, "ref": 157.3}
# kernel name -> (precision, device layout it reads); None = the padded cubes
# every context holds
KERNELS = {"mixed": ("mixed", "bicubic"), "mixed_taps": ("mixed", "taps"), "fp64": ("fp64", None),
           "ref": ("ref", None)}
UNPACKED_FP32_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes of one 128^3 / Nside 64 / 64-step / K = 16 gradient pass: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# summed over its dispatches (profiles/adjoint_kernels_128cube_nside64_kernel_trace.md)
ADJOINT_HBM_BYTES_PER_PASS = 39.1e9
# doubles of one stage record of the adjoint (state, then value, gradient, Hessian and third derivatives of
# the four fields) and of the interpolation adjoints one stage writes (10 per field)
ADJ_STAGE_RECORD_DOUBLES = 19 + 4 * 20
ADJ_STAGE_BARS_DOUBLES = 4 * 10


def adjoint_pass_bytes(rays, n_steps, K):
    """Algorithmic device-memory bytes of one gradient pass: every stage record written once (by the forward
    sweep or a segment's recompute) and read once (reverse sweep); the interpolation adjoints written once
    (reverse sweep) and read once with the stage positions (scatter, per field); checkpoints written once and
    read once; dL/dy and the maps."""
    stages = 4 * n_steps
    nseg = (n_steps + K - 1) // K
    per_ray = 8 * (2 * stages * ADJ_STAGE_RECORD_DOUBLES + 2 * stages * ADJ_STAGE_BARS_DOUBLES
                   + stages * 4 * 3 + 2 * (nseg + 1) * 19 + 2 * nseg * 19 + 12)
    return rays * per_ray


def host_cpu():
    """(threads the CPU baseline uses, CPU model): the process's CPU share (OMP_NUM_THREADS when set,
    else the CPUs this process may run on) and the model name from /proc/cpuinfo."""
    threads = 0
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        pass
    if threads <= 0:
        threads = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, model


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--grid", type=int, default=512)
    p.add_argument("--nside", type=int, default=1024)
    p.add_argument("--n-steps", type=int, default=256)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--kernels", default="mixed,fp64,mixed_taps,ref",
                   help="comma list of: mixed (FP32 interpolation from bicubic cell records), "
                        "mixed_taps (same arithmetic on the padded cubes), fp64, ref (the reference GPU "
                        "code's float/double mix, FP32 shift)")
    p.add_argument("--brick", type=int, default=16, help="column brick edge of the bicubic records")
    p.add_argument("--order", default="nested", choices=["nested", "linear"])
    p.add_argument("--subset-frac", type=float, default=0.01)
    p.add_argument("--no-oracle", action="store_true", help="skip the CPU oracle leg (accuracy + cpu baseline)")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="threads of the CPU baseline (default: OMP_NUM_THREADS, else every CPU this process "
                        "may run on)")
    p.add_argument("--workload", default="headline",
                   choices=["headline", "large", "adjoint", "batch", "layouts", "cli", "max"],
                   help="headline: 512^3 / Nside 1024 / 256 steps (the benchmark's JSON line); "
                        "large: 1024^3 / Nside 2048 / 512 steps, one full-sky launch; adjoint: 128^3 / "
                        "Nside 64 / 64 steps gradient with finite-difference check; batch: 48 uneven "
                        "instances in one plan launch (persistent and dispatch schedules); layouts: the headline workload's mixed "
                        "kernel on the padded cubes and on bicubic records of several brick edges; cli: the "
                        "command-line drop-in end to end on the headline workload (HDF5 files in, FITS maps out); "
                        "max: the largest grid, 2048^3 / Nside 2048 / 512 steps, on the padded cubes")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first run one device copy of 1 GiB (a known-size kernel that scales the "
                        "FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    return p.parse_args()


def full_sky_errors(test, ref):
    """dA and shear errors of a full-sky trace (maps 0-2, on the device) against the FP64 kernel's."""
    t, r = test.double(), ref.double()
    rel = (t[0] - r[0]).abs() / r[0].abs()
    de2 = (t[1] - r[1]) ** 2 + (t[2] - r[2]) ** 2
    return {"dA_max_rel": rel.max().item(), "dA_pixels_above_1e-5": int((rel > 1e-5).sum().item()),
            "shear_rms_rel": (de2.mean() / (r[1] ** 2 + r[2] ** 2).mean()).sqrt().item()}


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


def synth(fields, n, seed, dev):
    phi, shift = fields.snapshot(n, seed, device=dev)
    return phi, fields.interleave_shift(shift), shift


def workload_large(args, pkg, fields, dev):
    """1024^3 potential + 3 x 1024^3 shift, Nside 2048 full sky, 512 steps, mixed, one launch per
    layout and lane order: bicubic cell records in 16x16 column bricks (also as the 12 RING batches in
    one plan launch, persistent and dispatch schedules) and in row-major columns (when they fit in device memory), then the
    padded grids."""
    import torch
    n, nside, n_steps = 1024, 2048, 512
    rays = 12 * nside * nside
    t0 = time.time()
    tr = pkg.Tracer(0)
    phi, shift_il, shift = synth(fields, n, 3, dev)
    del shift
    tr.upload(phi, shift_il)
    del phi, shift_il
    torch.cuda.empty_cache()
    free, total = torch.cuda.mem_get_info()
    print(f"[bench] large: snapshot + upload {time.time() - t0:.1f} s; device memory free {free / 1e9:.1f} of "
          f"{total / 1e9:.1f} GB", file=sys.stderr, flush=True)
    out = torch.empty(6 * rays, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    res = {"device_memory_free_gb": free / 1e9, "device_memory_total_gb": total / 1e9}

    def run(name, orders):
        res[name] = {"record_bytes": tr.layout()[2]}
        for order_name, order in orders:
            tr.trace_device(nside, n_steps, out, precision=pkg.PREC_MIXED, order=order, stream=stream)
            ms = tr.last_trace_ms()
            res[name][order_name] = {"kernel_ms": ms, "ray_steps_per_s": rays * n_steps / (ms / 1000.0)}
            print(f"[bench] large {name} {order_name}: {ms:.0f} ms, "
                  f"{res[name][order_name]['ray_steps_per_s']:.3e} ray-steps/s", file=sys.stderr, flush=True)

    both = (("nested", pkg.ORDER_NESTED), ("linear", pkg.ORDER_LINEAR))
    try:
        tr.set_layout(pkg.LAYOUT_BICUBIC, 16)
        run("bicubic_brick16", both)
        # the 12 RING batches of the sky as 12 items of one persistent launch (64-ray tiles, 2 blocks
        # per CU, NESTED lanes within each batch, 8 XCD tile ranges)
        per = nside * nside
        items = [dict(tracer=tr, nside=nside, pix_begin=b * per, count=per, n_steps=n_steps,
                      out=out[6 * per * b: 6 * per * (b + 1)]) for b in range(12)]
        plan = pkg.Plan(items, precision=pkg.PREC_MIXED)
        plan.set_schedule(pkg.SCHEDULE_PERSISTENT)
        plan.launch(stream=stream)
        ms = plan.last_ms()
        res["bicubic_brick16"]["plan_12_batches"] = {"kernel_ms": ms, "ray_steps_per_s": rays * n_steps / (ms / 1000.0),
                                                    "config": plan.config()[:2]}
        print(f"[bench] large bicubic_brick16, 12 batches in one persistent launch: {ms:.0f} ms",
              file=sys.stderr, flush=True)
        # the same plan under the dispatch schedule (one 64-lane workgroup per tile)
        plan.set_schedule(pkg.SCHEDULE_DISPATCH)
        plan.launch(stream=stream)
        ms = plan.last_ms()
        res["bicubic_brick16"]["plan_12_batches_dispatch"] = {"kernel_ms": ms,
                                                             "ray_steps_per_s": rays * n_steps / (ms / 1000.0)}
        print(f"[bench] large bicubic_brick16, 12 batches in one dispatch-schedule launch: {ms:.0f} ms",
              file=sys.stderr, flush=True)
        plan.close()
        # the full sky as one item, under both schedules (what the 12 batches cost beyond one item)
        plan = pkg.Plan([dict(tracer=tr, nside=nside, pix_begin=0, count=rays, n_steps=n_steps, out=out)],
                        precision=pkg.PREC_MIXED)
        for sched_name, sched in (("persistent", pkg.SCHEDULE_PERSISTENT), ("dispatch", pkg.SCHEDULE_DISPATCH)):
            plan.set_schedule(sched)
            plan.launch(stream=stream)
            ms = plan.last_ms()
            res["bicubic_brick16"][f"plan_full_sky_{sched_name}"] = {
                "kernel_ms": ms, "ray_steps_per_s": rays * n_steps / (ms / 1000.0)}
            print(f"[bench] large bicubic_brick16, full sky as one {sched_name} plan item: {ms:.0f} ms",
                  file=sys.stderr, flush=True)
        plan.close()
        del items, plan
        tr.set_layout(pkg.LAYOUT_BICUBIC, 1)
        run("bicubic_brick1", (("nested", pkg.ORDER_NESTED),))
    except pkg.SrtError as e:
        res["bicubic"] = f"not built: {e}"
        print(f"[bench] large: bicubic records not built ({e})", file=sys.stderr, flush=True)
    tr.set_layout(pkg.LAYOUT_TAPS)
    run("taps", both)
    roof = None
    if isinstance(res.get("bicubic_brick16"), dict):
        ms = res["bicubic_brick16"]["nested"]["kernel_ms"]
        ach = ALGO_FLOPS_PER_RAY_STEP * rays * n_steps / (ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s",
                "frac_kernel_time_only": ach / PEAK_TFLOPS["mixed"],
                "algorithmic_flops_per_ray_step": ALGO_FLOPS_PER_RAY_STEP,
                "kernel": "trace_kernel<mixed, bicubic records>, 16x16 bricks, nested lanes"}
    print(json.dumps({"workload": "large", "config": {"grid": n, "nside": nside, "n_steps": n_steps,
                                                      "rays": rays, "precision": "mixed"}, "results": res,
                      "roofline": roof}))


def workload_max(args, pkg, fields, dev):
    """The largest grid the library accepts: 2048^3 (padded cubes of 2051^3, 138 GB of device memory), a field of
    period 128 cells (fields.periodic_field: separable modes at the other workloads' amplitudes; the spectral
    generator would need more memory than the card has beside the grids), Nside 2048 full sky, 512 steps, NESTED
    lanes.  Neither the bicubic records (4.4 TB) nor the FP64 grid copies (276 GB beside the cubes) fit, so the
    mixed kernel reads the padded cubes; the REF kernel is timed too."""
    import torch
    n, nside, n_steps = 2048, 2048, 512
    rays = 12 * nside * nside
    torch.cuda.empty_cache()
    free, total = torch.cuda.mem_get_info()
    if free < 290e9:
        print(json.dumps({"workload": "max", "skipped": f"needs ~280 GB of free device memory, {free / 1e9:.0f} GB free"}))
        return
    t0 = time.time()
    tr = pkg.Tracer(0)
    phi, shift = fields.periodic_field(n, 128, device=dev)
    tr.upload(phi, shift)
    del phi, shift
    torch.cuda.empty_cache()
    free, total = torch.cuda.mem_get_info()
    print(f"[bench] max: field + upload {time.time() - t0:.1f} s; device memory free {free / 1e9:.1f} of "
          f"{total / 1e9:.1f} GB", file=sys.stderr, flush=True)
    out = torch.empty(6 * rays, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    res = {"device_memory_free_gb": free / 1e9, "device_memory_total_gb": total / 1e9}
    for name, prec in (("mixed_taps", pkg.PREC_MIXED), ("ref", pkg.PREC_REF)):
        tr.trace_device(nside, n_steps, out, precision=prec, order=pkg.ORDER_NESTED, stream=stream)
        ms = tr.last_trace_ms()
        finite = bool(torch.isfinite(out).all())
        res[name] = {"kernel_ms": ms, "ray_steps_per_s": rays * n_steps / (ms / 1000.0), "maps_finite": finite}
        print(f"[bench] max {name}: {ms:.0f} ms, {res[name]['ray_steps_per_s']:.3e} ray-steps/s",
              file=sys.stderr, flush=True)
    ms = res["mixed_taps"]["kernel_ms"]
    ach = ALGO_FLOPS_PER_RAY_STEP * rays * n_steps / (ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s",
            "frac_kernel_time_only": ach / PEAK_TFLOPS["mixed"], "algorithmic_flops_per_ray_step": ALGO_FLOPS_PER_RAY_STEP,
            "kernel": "trace_kernel<mixed, padded cubes>, nested lanes"}
    print(json.dumps({"workload": "max", "config": {"grid": n, "nside": nside, "n_steps": n_steps, "rays": rays,
                                                    "field": "period 128 cells (fields.periodic_field)"},
                      "results": res, "roofline": roof}))
    tr.close()


def workload_adjoint(args, pkg, fields, dev):
    """128^3, Nside 64, 64 steps, checkpoints every 16; loss vs a second seed's maps + noise;
    central differences on 64 voxels within 60 cells of the observer."""
    import numpy as np
    import torch
    n, nside, n_steps, K = 128, 64, 64, 16
    count = 12 * nside * nside
    tr = pkg.Tracer(0)
    phi_t, shift_t, _ = synth(fields, n, 5, dev)
    tr.upload(phi_t, shift_t)
    tmaps, _, _ = tr.trace_adjoint(nside, n_steps)
    rng = np.random.default_rng(6)
    sel = [1, 2, 0]
    target = tmaps[sel] + torch.from_numpy(rng.standard_normal((3, count)) * 0.01).to(dev)
    phi, shift_il, _ = synth(fields, n, 4, dev)
    phi_h, shift_h = phi.cpu().numpy(), shift_il.cpu().numpy()

    def loss(maps):
        r = maps[sel] - target
        return float((r * r).sum()), r

    tr.upload(phi_h, shift_h)
    maps, _, _ = tr.trace_adjoint(nside, n_steps)
    L0, r = loss(maps)
    g = torch.zeros_like(maps)
    g[sel] = 2 * r
    tr.trace_adjoint(nside, n_steps, g, checkpoint_every=K)  # warm-up
    t0 = time.perf_counter()
    _, gphi, gshift = tr.trace_adjoint(nside, n_steps, g, checkpoint_every=K)
    torch.cuda.synchronize()
    adj_s = time.perf_counter() - t0
    adj_kernel_ms = tr.last_trace_ms()
    gphi, gshift = gphi.cpu().numpy(), gshift.cpu().numpy()
    rng = np.random.default_rng(9)
    checks = []
    while len(checks) < 64:
        v = rng.integers(0, n, 3)
        if np.linalg.norm(v - n / 2) >= 60:
            continue
        i, j, k = (int(a) for a in v)
        comp = len(checks) % 4  # 0 = potential, 1..3 = shift components
        base = phi_h if comp == 0 else shift_h[..., comp - 1]
        # steps: 1e-3 x rms for the potential, 1e-1 x rms for the shift, the steps of least difference error
        # (a smaller shift step only adds the rounding noise of the perturbed traces, which falls as 1/h)
        h = (1e-3 if comp == 0 else 1e-1) * float(np.sqrt(np.mean(base.astype(np.float64) ** 2)))
        vals = []
        for sgn in (1, -1):
            p2, s2 = phi_h.copy(), shift_h.copy()
            if comp == 0:
                p2[k, j, i] = np.float32(float(phi_h[k, j, i]) + sgn * h)
                step = float(p2[k, j, i])
            else:
                s2[k, j, i, comp - 1] = np.float32(float(shift_h[k, j, i, comp - 1]) + sgn * h)
                step = float(s2[k, j, i, comp - 1])
            tr.upload(p2, s2)
            m, _, _ = tr.trace_adjoint(nside, n_steps)
            vals.append((m[sel], step))
        # L+ - L- = sum (m+ - m-) (m+ + m- - 2 t): no cancellation between two large sums
        dm = vals[0][0] - vals[1][0]
        dL = float((dm * (vals[0][0] + vals[1][0] - 2 * target)).sum())
        fd = dL / (vals[0][1] - vals[1][1])
        ad = float(gphi[k, j, i] if comp == 0 else gshift[k, j, i, comp - 1])
        checks.append((comp, fd, ad))
    scale = {c: max(abs(x[1]) for x in checks if x[0] == c) for c in range(4)}
    worst = {c: max(abs(x[1] - x[2]) for x in checks if x[0] == c) / scale[c] for c in range(4)}
    algo_bytes = adjoint_pass_bytes(count, n_steps, K)
    print(json.dumps({"workload": "adjoint", "config": {"grid": n, "nside": nside, "n_steps": n_steps,
                                                        "checkpoint_every": K, "rays": count},
                      "adjoint_wall_s": adj_s, "adjoint_kernel_ms": adj_kernel_ms,
                      "roofline": {"bound": "hbm", "achieved": algo_bytes / (adj_kernel_ms * 1e-3) / 1e9,
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": algo_bytes / (adj_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "algorithmic_bytes_per_pass": algo_bytes,
                                   "traffic": ADJOINT_HBM_BYTES_PER_PASS,
                                   "kernel": "the gradient pass (forward sweep, recomputes, reverse sweeps, "
                                             "scatters): stage records through HBM"},
                      "ray_steps_per_s": count * n_steps / adj_s,
                      "fd_voxels": len(checks),
                      "max_abs_diff_over_max_fd": {"potential": worst[0], "shift_x": worst[1], "shift_y": worst[2],
                                                   "shift_z": worst[3]}}))


def workload_cli(args, pkg, fields, dev):
    """The drop-in end to end: the 512^3 snapshot (seed 2) written as the two HDF5 files the reference reads,
    then the command-line tool at Nside 1024 / 256 steps, full sky (HDF5 in, six FITS maps out), timed as a
    process and by phase; its dA map is checked bit for bit against a library trace of the same snapshot."""
    import re
    import subprocess
    import tempfile
    import numpy as np
    from snapshot_raytracer_amd import snapshot_io
    n, nside, n_steps = args.grid, args.nside, args.n_steps
    cli = os.path.join(ROOT, "snapshot-raytracer_amd", "bin", "snapshot-raytracer")
    if not os.path.exists(cli) or not snapshot_io.hdf5_available():
        print(json.dumps({"workload": "cli", "skipped": "CLI not built or HDF5 library absent"}))
        return
    phi, shift_il, _ = synth(fields, n, args.seed, dev)
    phi_h, shift_h = phi.cpu().numpy(), shift_il.cpu().numpy()
    del phi, shift_il
    with tempfile.TemporaryDirectory(prefix="srt_cli_") as d:
        t0 = time.perf_counter()
        snapshot_io.write_potential(os.path.join(d, "phi.h5"), phi_h)
        snapshot_io.write_shift(os.path.join(d, "B.h5"), shift_h)
        write_s = time.perf_counter() - t0
        print(f"[bench] cli: HDF5 inputs written in {write_s:.2f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        r = subprocess.run([cli, "phi.h5", "B.h5", str(nside), str(n_steps)], cwd=d, capture_output=True,
                           text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            print(r.stdout[-2000:], r.stderr[-2000:], file=sys.stderr)
            raise SystemExit(f"cli failed with status {r.returncode}")
        got, _ = snapshot_io.read_healpix_fits(os.path.join(d, "dAmap_batch0.fits"))
    timing = re.search(r"Timing: HDF5 read ([0-9.e+-]+) s, device setup ([0-9.e+-]+) s, trace with map download "
                       r"([0-9.e+-]+) s, FITS write ([0-9.e+-]+) s, total ([0-9.e+-]+) s", r.stdout)
    kernel = re.search(r"kernel ([0-9.e+-]+) ms", r.stdout)
    tr = pkg.Tracer(0)
    tr.set_layout(pkg.LAYOUT_BICUBIC, 16)
    tr.upload(phi_h, shift_h)
    ref = tr.trace(nside, n_steps, precision=pkg.PREC_MIXED, order=pkg.ORDER_NESTED)
    tr.close()
    rays = 12 * nside * nside
    phases = dict(zip(["hdf5_read_s", "device_setup_s", "trace_with_map_download_s", "fits_write_s", "total_s"],
                      (float(x) for x in timing.groups()))) if timing else None
    print(json.dumps({"workload": "cli", "config": {"grid": n, "nside": nside, "n_steps": n_steps, "rays": rays,
                                                    "precision": "mixed", "layout": "auto (bicubic records)"},
                      "process_wall_s": wall, "phases": phases,
                      "kernel_ms": float(kernel.group(1)) if kernel else None,
                      "ray_steps_per_s_process": rays * n_steps / wall,
                      "dA_map_equals_library_trace": bool(np.array_equal(got, ref[0]))}))


def workload_batch(args, pkg, fields, dev):
    """48 instances, grids from {128,192,256,384}^3, observers uniform in the box, n_steps in
    [64, N/2], Nside from {128,256,512}, one random RING batch each (rng 7, seeds 1000 + 4 i).
    Mixed precision; bicubic cell records for every instance when they all fit in device memory.
    Times 48 separate launches, the plan under both schedules, then the autotuned plan."""
    import numpy as np
    import torch
    params = [(q["n"], q["nside"], q["batch"], q["n_steps"], q["observer"])
              for q in fields.uneven_batch_params()]
    rec_bytes = sum(4 * 64 * n * n * (n + 3) for n, *_ in params)
    taps_bytes = sum(4 * 4 * (n + 3) ** 3 for n, *_ in params)
    free, total = torch.cuda.mem_get_info()
    # room for the records next to the grids, the outputs and the generator's buffers
    layout = "bicubic" if rec_bytes + taps_bytes + 8e9 < free else "taps"
    print(f"[bench] batch: records {rec_bytes / 1e9:.1f} GB + grids {taps_bytes / 1e9:.1f} GB, free "
          f"{free / 1e9:.1f} GB -> {layout}", file=sys.stderr, flush=True)
    items = []
    t0 = time.time()
    for i, (n, nside, b, n_steps, obs) in enumerate(params):
        tr = pkg.Tracer(0)
        if layout == "bicubic":
            tr.set_layout(pkg.LAYOUT_BICUBIC, 16)
        phi, shift_il, _ = synth(fields, n, 1000 + 4 * i, dev)
        tr.upload(phi, shift_il)
        del phi, shift_il
        torch.cuda.empty_cache()
        lo, cnt = pkg.batch_range(nside, b)
        items.append(dict(tracer=tr, nside=nside, pix_begin=lo, count=cnt, n_steps=n_steps, observer=obs,
                          out=torch.empty(6 * cnt, dtype=torch.float32, device=dev)))
    torch.cuda.synchronize()
    print(f"[bench] batch: 48 snapshots in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    total_rs = sum(it["count"] * it["n_steps"] for it in items)
    stream = torch.cuda.Stream()

    def sequential(order):
        with torch.cuda.stream(stream):
            for it in items:
                it["tracer"].trace_device(it["nside"], it["n_steps"], it["out"], pix_begin=it["pix_begin"],
                                          count=it["count"], observer=it["observer"], precision=pkg.PREC_MIXED,
                                          order=order, stream=stream)
        stream.synchronize()

    seq = {}
    for name, order in (("linear", pkg.ORDER_LINEAR), ("nested", pkg.ORDER_NESTED)):
        sequential(order)
        t0 = time.perf_counter()
        sequential(order)
        wall = time.perf_counter() - t0
        seq[name] = {"ms": wall * 1000, "ray_steps_per_s": total_rs / wall}
        print(f"[bench] batch: 48 launches, {name} lane order: {wall * 1000:.1f} ms", file=sys.stderr, flush=True)
    plan = pkg.Plan(items, precision=pkg.PREC_MIXED)
    default_cfg = plan.config()[:2]
    variants = {}
    plan.set_schedule(pkg.SCHEDULE_PERSISTENT)
    for name, order, groups in (("linear_1group", pkg.ORDER_LINEAR, 1), ("nested_1group", pkg.ORDER_NESTED, 1),
                                ("nested_8groups", pkg.ORDER_NESTED, 8)):
        plan.set_order(order, groups)
        plan.launch(stream=stream)
        stream.synchronize()
        plan.launch(stream=stream)
        stream.synchronize()
        variants[name] = plan.last_ms()
        print(f"[bench] batch plan {name}: {variants[name]:.1f} ms", file=sys.stderr, flush=True)
    plan.set_schedule(pkg.SCHEDULE_DISPATCH)
    for name, groups in (("dispatch_nested_1group", 1), ("dispatch_nested_8groups", 8)):
        plan.set_order(pkg.ORDER_NESTED, groups)
        plan.launch(stream=stream)
        stream.synchronize()
        plan.launch(stream=stream)
        stream.synchronize()
        variants[name] = plan.last_ms()
        print(f"[bench] batch plan {name}: {variants[name]:.1f} ms", file=sys.stderr, flush=True)
    plan.set_schedule(pkg.SCHEDULE_PERSISTENT)
    plan.set_order(pkg.ORDER_NESTED, 8)
    tile, blocks = plan.autotune(stream=stream)
    schedule = "dispatch" if plan.schedule() == pkg.SCHEDULE_DISPATCH else "persistent"
    brick = None
    if layout == "bicubic":
        t0 = time.time()
        brick = plan.autotune_brick(stream=stream)
        print(f"[bench] batch plan brick autotune: edge {brick} ({time.time() - t0:.1f} s)", file=sys.stderr,
              flush=True)
    res = {}
    for graph in (False, True):
        plan.launch(stream=stream, graph=graph)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.launch(stream=stream, graph=graph)
        stream.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        res["graph" if graph else "direct"] = {"ms": wall * 1000, "ray_steps_per_s": total_rs / wall,
                                               "kernel_ms": plan.last_ms()}
        print(f"[bench] batch plan autotuned ({schedule}, {tile}, {blocks}, brick {brick}) graph={graph}: "
              f"{wall * 1000:.1f} ms",
              file=sys.stderr, flush=True)
    print(json.dumps({"workload": "batch", "instances": 48, "ray_steps": total_rs, "layout": layout,
                      "sequential_launches": seq,
                      "plan_default_config": {"tile_rays": default_cfg[0], "blocks": default_cfg[1]},
                      "plan_variants_ms": variants,
                      "autotuned": {"schedule": schedule, "tile_rays": tile, "blocks": blocks, "brick": brick},
                      "plan": res,
                      "roofline": {"bound": "valu", "unit": "TFLOP/s", "peak": PEAK_TFLOPS["mixed"],
                                   "achieved": ALGO_FLOPS_PER_RAY_STEP * total_rs / (res["direct"]["kernel_ms"] * 1e-3) / 1e12,
                                   "frac_kernel_time_only": ALGO_FLOPS_PER_RAY_STEP * total_rs
                                   / (res["direct"]["kernel_ms"] * 1e-3) / 1e12 / PEAK_TFLOPS["mixed"],
                                   "algorithmic_flops_per_ray_step": ALGO_FLOPS_PER_RAY_STEP,
                                   "kernel": "the autotuned plan launch (" + schedule + " schedule)"}}))
    plan.close()


def workload_layouts(args, pkg, fields, dev):
    """512^3 / Nside 1024 / 256 steps, mixed kernel: padded cubes vs bicubic records with column
    bricks of edge 1 (row-major columns), 4, 8, 16 and 32, each in NESTED and linear lane order."""
    import torch
    n, nside, n_steps = args.grid, args.nside, args.n_steps
    rays = 12 * nside * nside
    tr = pkg.Tracer(0)
    phi, shift_il, shift = synth(fields, n, args.seed, dev)
    del shift
    tr.upload(phi, shift_il)
    del phi, shift_il
    torch.cuda.empty_cache()
    out = torch.empty(6 * rays, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    res = {}
    for lay, brick in (("taps", 8), ("bicubic", 1), ("bicubic", 4), ("bicubic", 8), ("bicubic", 16), ("bicubic", 32)):
        t0 = time.time()
        tr.set_layout(pkg.LAYOUT_TAPS if lay == "taps" else pkg.LAYOUT_BICUBIC, brick)
        torch.cuda.synchronize()
        build_s = time.time() - t0
        name = lay if lay == "taps" else f"bicubic_brick{brick}"
        res[name] = {"build_s": build_s, "record_bytes": tr.layout()[2]}
        for order_name, order in (("nested", pkg.ORDER_NESTED), ("linear", pkg.ORDER_LINEAR)):
            tr.trace_device(nside, n_steps, out, precision=pkg.PREC_MIXED, order=order, stream=stream)
            ms = tr.last_trace_ms()
            res[name][order_name] = {"kernel_ms": ms, "ray_steps_per_s": rays * n_steps / (ms / 1000.0)}
            print(f"[bench] layouts {name} {order_name}: {ms:.1f} ms", file=sys.stderr, flush=True)
    print(json.dumps({"workload": "layouts", "config": {"grid": n, "nside": nside, "n_steps": n_steps,
                                                        "rays": rays, "precision": "mixed"}, "results": res}))


def device_info() -> dict:
    """The GPU the line was measured on: what the runtime reports (name, architecture, CUs, memory, L2), and the
    shader clock and L2 channel count the trace kernels' counters show (GRBM_GUI_ACTIVE / 8 XCDs over the kernel
    time; TCC_CYCLE_sum over those cycles), which the runtime does not report."""
    import torch
    p = torch.cuda.get_device_properties(0)
    fields = {"name": "name", "arch": "gcnArchName", "cus": "multi_processor_count", "memory_bytes": "total_memory",
              "l2_bytes": "L2_cache_size", "wavefront": "warp_size"}
    info = {k: getattr(p, a) for k, a in fields.items() if hasattr(p, a)}
    info.update({"shader_clock_ghz_from_counters": 2.37, "l2_channels_from_counters": 128})
    return info


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the ray tracer shards rays "
              "across GPUs by RING batches with no exchange, and the benchmark measures the one-GPU share.",
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
    if args.workload != "headline":
        {"large": workload_large, "adjoint": workload_adjoint, "batch": workload_batch,
         "layouts": workload_layouts, "cli": workload_cli, "max": workload_max}[args.workload](
            args, pkg, fields, dev)
        return
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
    names = [k for k in args.kernels.split(",") if k]
    # kernels on the padded cubes first, so the bicubic records are built once
    names.sort(key=lambda k: KERNELS[k][1] == "bicubic")
    layout_s = {}
    fp64_maps = None
    full_sky = {}
    for name in names:
        prec_name, layout = KERNELS[name]
        prec = pkg.PRECISIONS[prec_name]
        if layout is not None:
            t0 = time.time()
            tr.set_layout(pkg.LAYOUT_BICUBIC if layout == "bicubic" else pkg.LAYOUT_TAPS, args.brick)
            torch.cuda.synchronize()
            layout_s[layout] = time.time() - t0
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
        if name == "fp64":
            fp64_maps = out.view(6, rays)[:3].clone()
        elif fp64_maps is not None:
            full_sky[name] = full_sky_errors(out.view(6, rays)[:3], fp64_maps)
        print(f"[bench] {name}: {ms_per_step:.1f} ms/step, kernel {avg_kernel_ms:.1f} ms, "
              f"{results[name]['ray_steps_per_s']:.3e} ray-steps/s", file=sys.stderr, flush=True)

    # the drop-in path with host buffers (what the CLI does after reading HDF5): host -> device upload
    # with the record build, then a synchronous trace with host outputs (device -> host copy of the maps)
    end_to_end = None
    if "mixed" in results:
        phi_h = phi.cpu().numpy()
        shift_h = fields.interleave_shift(shift).cpu().numpy()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr.upload(phi_h, shift_h)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        maps_h = tr.trace(nside, n_steps, precision=pkg.PREC_MIXED, order=order)
        t2 = time.perf_counter()
        end_to_end = {"upload_from_host_s": t1 - t0, "trace_to_host_s": t2 - t1,
                      "kernel_ms": tr.last_trace_ms(),
                      "ray_steps_per_s_incl_transfers": rays * n_steps / (t2 - t0),
                      "maps_equal_device_run": bool(np.array_equal(maps_h[:, subset], maps_sub["mixed"]))}
        print(f"[bench] end to end from host buffers: upload {t1 - t0:.2f} s, trace {t2 - t1:.2f} s",
              file=sys.stderr, flush=True)
        del phi_h, shift_h, maps_h

    cpu_baseline = None
    accuracy = {}
    if not args.no_oracle:
        from oracle import oracle
        Fh = oracle.Fields(phi.cpu().numpy(), shift.cpu().numpy())
        threads, cpu_model = host_cpu()
        if args.cpu_threads > 0:
            threads = args.cpu_threads
        visible = len(os.sched_getaffinity(0))
        print(f"[bench] oracle fp64 and ref forms on {subset.size} rays x {n_steps} steps, {threads} threads "
              f"({visible} CPUs visible, {cpu_model})", file=sys.stderr, flush=True)
        ref, sec = oracle.trace(Fh, nside, subset, n_steps, form=oracle.FORM_FP64, nthreads=threads)
        _, sec_ref = oracle.trace(Fh, nside, subset, n_steps, form=oracle.FORM_REF, nthreads=threads)
        # the 64^3 / Nside 16 / 32-step snapshot (seed 1), full sky, both forms
        phi1, shift1 = fields.snapshot(64, 1)
        F1 = oracle.Fields(phi1.numpy(), shift1.numpy())
        pix1 = np.arange(12 * 16 * 16)
        small = {}
        for form_name, form in (("fp64", oracle.FORM_FP64), ("ref", oracle.FORM_REF)):
            _, s1 = oracle.trace(F1, 16, pix1, 32, form=form, nthreads=threads)
            small[form_name] = pix1.size * 32 / s1
        rs = subset.size * n_steps
        cpu_baseline = {"value": rs / sec_ref, "unit": "ray-steps/s", "cores": threads, "kind": "port",
                        "sample": f"host oracle in the reference's arithmetic (REF form: float interpolation, "
                                  f"double state; OpenMP C++, {threads} threads) on the seeded 1% pixel subset "
                                  f"({subset.size} rays) x {n_steps} steps of the same snapshot, {sec_ref:.1f} s",
                        "fp64_form": {"value": rs / sec, "seconds": sec},
                        "cpu_model": cpu_model, "cpus_visible": visible,
                        "snapshot_64cube_nside16_32steps_full_sky": small}
        for name, m in maps_sub.items():
            accuracy[name] = map_errors(m, ref)

    head = "mixed" if "mixed" in results else next(iter(results))
    r = results[head]
    roofline = None
    if head in PEAK_TFLOPS:
        rs_launch = rays * n_steps
        # the mixed kernel reads bicubic records built once per snapshot: one build is charged to the trace
        build_ms = 1000.0 * layout_s.get("bicubic", 0.0) if KERNELS[head][1] == "bicubic" else 0.0
        charged_ms = r["kernel_ms_avg"] + build_ms
        achieved = ALGO_FLOPS_PER_RAY_STEP * rs_launch / (charged_ms / 1000.0) / 1e12
        roofline = {"bound": "valu",
                    "compute_unit": "vector ALU (FP32 interpolation, FP64 state); nothing on this path is a "
                                    "matrix product, so MFMA does not apply",
                    "achieved": achieved, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                    "frac": achieved / PEAK_TFLOPS[head],
                    "traffic": HBM_BYTES_PER_LAUNCH.get(head),
                    "hbm_frac": (HBM_BYTES_PER_LAUNCH[head] / (r["kernel_ms_avg"] / 1000.0) / 8e12
                                 if head in HBM_BYTES_PER_LAUNCH else None),
                    "algorithmic_flops_per_ray_step": ALGO_FLOPS_PER_RAY_STEP,
                    "reference_form_flops_per_ray_step": REF_FORM_FLOPS_PER_RAY_STEP,
                    "ray_steps_per_launch": rs_launch,
                    "kernel_ms": r["kernel_ms_avg"], "record_build_ms": build_ms, "charged_ms": charged_ms,
                    "frac_kernel_time_only": ALGO_FLOPS_PER_RAY_STEP * rs_launch / (r["kernel_ms_avg"] / 1e3)
                                             / 1e12 / PEAK_TFLOPS[head],
                    "kernel": "trace_kernel<mixed, bicubic records>" if head == "mixed" else f"trace_kernel<{head}>"}
        if head in FLOPS_PER_RAY_STEP:
            ex = FLOPS_PER_RAY_STEP[head] * rs_launch / (r["kernel_ms_avg"] / 1000.0) / 1e12
            roofline["executed"] = {"flops_per_ray_step": FLOPS_PER_RAY_STEP[head], "tflops": ex,
                                    "frac_of_peak": ex / PEAK_TFLOPS[head],
                                    "frac_of_unpacked_fp32_rate": ex / UNPACKED_FP32_TFLOPS}
        if head in GATHER_PER_LAUNCH:
            g, t = GATHER_PER_LAUNCH[head], r["kernel_ms_avg"] / 1000.0
            roofline["gather"] = {
                "l1_lines_per_ray_step": g["l1_lines"] / rs_launch,
                "l1_lines_per_cu_clock": g["l1_lines"] / (CUS * CLOCK_HZ * t),
                "ta_busy": g["ta_busy"],
                "l2_read_bytes": g["l2_read_bytes"], "l2_read_TBps": g["l2_read_bytes"] / t / 1e12,
                "l2_peak_TBps": L2_PEAK_TBPS, "l2_frac": g["l2_read_bytes"] / t / 1e12 / L2_PEAK_TBPS,
                "stencil_bytes_per_ray_step": 4096, "stencil_TBps": 4096 * rs_launch / t / 1e12}
        if "fp64" in results and head != "fp64":
            a64 = ALGO_FLOPS_PER_RAY_STEP * rs_launch / (results["fp64"]["kernel_ms_avg"] / 1000.0) / 1e12
            roofline["fp64_kernel"] = {"achieved": a64, "peak": PEAK_TFLOPS["fp64"], "frac": a64 / PEAK_TFLOPS["fp64"]}
    line = {
        "metric": "ray-steps/s at 512^3, Nside 1024, 256 steps (mixed-precision kernel)",
        "value": r["ray_steps_per_s"],
        "value_excludes": "upload of the snapshot and its per-snapshot bicubic record build (see "
                          "value_with_record_build), like the metric's definition",
        "value_with_record_build": (rays * n_steps / ((r["ms_per_step"] + 1000.0 * layout_s.get("bicubic", 0.0))
                                                      / 1000.0) if KERNELS[head][1] == "bicubic" else None),
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
        "layout": {"mixed": f"bicubic cell records, {args.brick}x{args.brick} column bricks",
                   "mixed_taps": "padded periodic cubes", "fp64": "padded periodic cubes (FP64 copies)",
                   "ref": "padded periodic cubes", "build_s": layout_s},
        "map_error_vs_fp64_oracle": accuracy,
        "full_sky_error_vs_fp64_kernel": full_sky,
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "upload_s": upload_s,
        "end_to_end_host_buffers": end_to_end,
        "device": device_info(),
    }
    print(json.dumps(line))
    tr.close()


if __name__ == "__main__":
    main()
