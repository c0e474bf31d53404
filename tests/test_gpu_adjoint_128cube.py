"""Adjoint kernel at its full workload: the map-loss gradient with respect to
both grids, checked against central finite differences of the GPU FP64
forward trace on 64 voxels and of the host FP64 oracle on 2 voxels.

Workload: periodic 128^3 snapshot (potential rms 1e-5, shift rms 1e-7,
P(k) ~ k^-3, seed 4), observer at the grid centre, Nside 64 (49,152 rays),
64 RK4 steps, checkpoints every 16 steps.  Loss = sum over pixels of
(eps1 - t1)^2 + (eps2 - t2)^2 + (dA - tA)^2, the targets being the maps of a
second snapshot (seed 5) plus Gaussian noise of sigma 0.01 (numpy
default_rng(6)).  Voxels: default_rng(9) draws inside 60 cells of the
observer, cycling potential, B_x, B_y, B_z.  Step: 1e-3 x the rms of the
potential, 1e-1 x the rms of the shift (FD_STEP); the quotient uses the
float32 values the grid stores, and
L(+) - L(-) is summed per pixel as (m+ - m-)(m+ + m- - 2 t), free of the
cancellation of two large sums.  Bar: |FD - adjoint| <= 1e-6 x the largest
|FD| of the same grid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NSIDE, STEPS, K = 128, 64, 64, 16
COUNT = 12 * NSIDE * NSIDE
SEL = [1, 2, 0]  # eps1, eps2, dA rows of the map array
BAR = 1e-6
# Central-difference steps per field, as fractions of the field's rms, from a sweep over 1e-4 .. 1e-1 on this
# workload: for the shift the differences' error falls as 1/h all the way (B_x worst 5.5e-6, 7.9e-7, 6.5e-8,
# 4.1e-9), the rounding of the perturbed FP64 traces over a step that moves the maps very little; the maps stay
# linear in a shift value at 1e-1.  For the potential the error is smallest at 1e-3 (8.6e-10) and truncation
# shows at 1e-1 (4.2e-6).
FD_STEP = (1e-3, 1e-1, 1e-1, 1e-1)  # potential, B_x, B_y, B_z


@pytest.fixture(scope="module")
def c4(srt, fields):
    import torch
    print("[gpu] adjoint 128^3 / Nside 64 / 64 steps: snapshots", flush=True)
    dev = torch.device("cuda:0")
    tr = srt.Tracer(0)
    phi_t, shift_t = fields.snapshot(N, 5, device=dev)
    tr.upload(phi_t, fields.interleave_shift(shift_t))
    tmaps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    noise = np.random.default_rng(6).standard_normal((3, COUNT)) * 0.01
    target = tmaps[SEL] + torch.from_numpy(noise).to(dev)
    phi, shift = fields.snapshot(N, 4, device=dev)
    shift_il = fields.interleave_shift(shift)
    del phi_t, shift_t, tmaps
    yield dict(tr=tr, phi=phi.contiguous(), shift_il=shift_il, target=target, dev=dev)
    tr.close()
    torch.cuda.empty_cache()


def loss_grad(maps, target):
    import torch
    r = maps[SEL] - target
    g = torch.zeros_like(maps)
    g[SEL] = 2 * r
    return g


def voxels(count=64):
    rng = np.random.default_rng(9)
    out = []
    while len(out) < count:
        v = rng.integers(0, N, 3)
        if np.linalg.norm(v - N / 2) >= 60:
            continue
        out.append((len(out) % 4, tuple(int(a) for a in v)))  # 0 = potential, 1..3 = shift component
    return out


def perturbed_maps(c4, comp, ijk, h):
    """FP64 maps (selected rows) with the grid value at ijk moved by +h and -h; the float32 values stored."""
    tr, phi, shift_il = c4["tr"], c4["phi"], c4["shift_il"]
    i, j, k = ijk
    arr, idx = (phi, (k, j, i)) if comp == 0 else (shift_il, (k, j, i, comp - 1))
    base = float(arr[idx])
    res = []
    for sgn in (1, -1):
        arr[idx] = base + sgn * h
        stored = float(arr[idx])
        tr.upload(phi, shift_il)
        m, _, _ = tr.trace_adjoint(NSIDE, STEPS)
        res.append((m[SEL].clone(), stored))
    arr[idx] = base
    return res


def test_adjoint_128cube_matches_gpu_central_differences(srt, c4):
    tr, phi, shift_il, target = c4["tr"], c4["phi"], c4["shift_il"], c4["target"]
    tr.upload(phi, shift_il)
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    g = loss_grad(maps, target)
    print("[gpu] adjoint 128^3: gradient pass", flush=True)
    _, gphi, gshift = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=K)
    gphi, gshift = gphi.cpu().numpy(), gshift.cpu().numpy()
    rms = [float(phi.double().pow(2).mean().sqrt())] + [float(shift_il[..., a].double().pow(2).mean().sqrt())
                                                        for a in range(3)]
    checks = []
    for comp, (i, j, k) in voxels():
        (mp, vp), (mm, vm) = perturbed_maps(c4, comp, (i, j, k), FD_STEP[comp] * rms[comp])
        fd = float(((mp - mm) * (mp + mm - 2 * target)).sum()) / (vp - vm)
        ad = float(gphi[k, j, i] if comp == 0 else gshift[k, j, i, comp - 1])
        checks.append((comp, fd, ad))
    print(f"[gpu] adjoint 128^3: {len(checks)} voxels differenced", flush=True)
    for comp, name in enumerate(("potential", "B_x", "B_y", "B_z")):
        sub = [c for c in checks if c[0] == comp]
        scale = max(abs(c[1]) for c in sub)
        worst = max(abs(c[1] - c[2]) for c in sub) / scale
        print(f"{name}: {len(sub)} voxels, max |FD| {scale:.3e}, worst |FD - adjoint| / max |FD| = {worst:.3e}")
        assert scale > 0
        assert worst < BAR, (name, worst)


def test_adjoint_128cube_matches_oracle_central_differences(srt, oracle, c4):
    """Two voxels (the first potential and the first B_x voxel of the draw) differenced by the host FP64
    oracle over the full sky; the bar is relative to the largest adjoint gradient of that grid over its
    16 voxels of the draw, as in the GPU differences."""
    tr, phi, shift_il, target = c4["tr"], c4["phi"], c4["shift_il"], c4["target"]
    tr.upload(phi, shift_il)
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    g = loss_grad(maps, target)
    _, gphi, gshift = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=K)
    gphi, gshift = gphi.cpu().numpy(), gshift.cpu().numpy()
    phi_h = phi.cpu().numpy()
    shift_h = np.ascontiguousarray(np.moveaxis(shift_il.cpu().numpy(), -1, 0))  # planar (3, n, n, n)
    tgt = target.cpu().numpy()
    pix = np.arange(COUNT)
    draw = voxels()
    picks = [v for v in draw if v[0] == 0][:1] + [v for v in draw if v[0] == 1][:1]

    def adjoint_at(comp, ijk):
        i, j, k = ijk
        return float(gphi[k, j, i] if comp == 0 else gshift[k, j, i, comp - 1])
    for comp, (i, j, k) in picks:
        base = phi_h if comp == 0 else shift_h[comp - 1]
        h = FD_STEP[comp] * float(np.sqrt(np.mean(base.astype(np.float64) ** 2)))
        vals = []
        for sgn in (1, -1):
            p, s = phi_h.copy(), shift_h.copy()
            arr = p if comp == 0 else s[comp - 1]
            arr[k, j, i] = np.float32(float(base[k, j, i]) + sgn * h)
            _, sec, st = oracle.trace(oracle.Fields(p, s), NSIDE, pix, STEPS, return_state=True)
            print(f"[gpu] oracle full-sky trace {sec:.1f} s", flush=True)
            vals.append((np.stack([st[:, 16], st[:, 17], st[:, 12]]), float(arr[k, j, i])))
        (mp, vp), (mm, vm) = vals
        fd = float(((mp - mm) * (mp + mm - 2 * tgt)).sum()) / (vp - vm)
        ad = adjoint_at(comp, (i, j, k))
        scale = max(abs(adjoint_at(c, v)) for c, v in draw if c == comp)
        rel = abs(fd - ad) / scale
        print(f"comp {comp} voxel {(i, j, k)}: oracle FD {fd:+.10e} adjoint {ad:+.10e} "
              f"|FD - adjoint| / max |adjoint| = {rel:.2e}")
        assert abs(fd) > 0
        assert rel < BAR, (comp, rel)
