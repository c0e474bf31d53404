"""Adjoint kernel (srt_trace_adjoint): gradients of a map loss with respect to
the potential and shift grids, checked against central finite differences of
the FP64 forward trace (GPU) and of the FP64 host oracle, in FP64.

Workload: periodic 32^3 snapshot (seed 31), Nside 8 (768 rays), 12 steps,
checkpoints every 4 steps; loss = sum over pixels of (eps1 - t1)^2 +
(eps2 - t2)^2 + (dA - tA)^2 with targets from a second seed's maps plus
Gaussian noise (sigma = 0.01 x map rms).  Finite differences use the step
1e-3 x the potential's rms and 1e-1 x the shift's rms (FD_STEP, the steps of
least difference error, as measured on the 128^3 workload in
test_gpu_adjoint_128cube.py) on voxels within 8 cells of the observer; the perturbed
value is the float32 the grid stores, the difference quotient uses the
exact float32 step, and L(+) - L(-) is summed as (m+ - m-)(m+ + m- - 2t) per
pixel.  Bar: 1e-6 relative to the largest checked gradient of each grid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NSIDE, STEPS, K = 32, 8, 12, 4
FD_STEP = {"phi": 1e-3, 0: 1e-1, 1: 1e-1, 2: 1e-1}  # fractions of each field's rms


@pytest.fixture(scope="module")
def setup(srt, fields):
    import torch
    phi, shift = fields.snapshot(N, 31)
    phi2, shift2 = fields.snapshot(N, 32)
    dev = torch.device("cuda:0")
    tr = srt.Tracer(0)
    tr.upload(phi2.numpy(), fields.interleave_shift(shift2).numpy())
    tmaps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    rng = np.random.default_rng(33)
    sel = [1, 2, 0]  # eps1, eps2, dA rows of the map array
    target = tmaps.cpu().numpy()[sel]
    target = target + rng.standard_normal(target.shape) * 0.01 * target.std(axis=1, keepdims=True)
    return tr, phi.numpy(), shift.numpy(), torch.from_numpy(target).to(dev)


def loss_and_grad(maps, target):
    import torch
    sel = [1, 2, 0]
    r = maps[sel] - target
    L = float((r * r).sum())
    g = torch.zeros_like(maps)
    g[sel] = 2 * r
    return L, g


def maps_of(tr, phi, shift_planar):
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift_planar, 0, -1)))
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    return maps[[1, 2, 0]]


def loss_diff(mp, mm, target):
    """L(+) - L(-) = sum (m+ - m-)(m+ + m- - 2 t), free of the cancellation of two large sums."""
    return float(((mp - mm) * (mp + mm - 2 * target)).sum())


def voxels(rng, count):
    out = []
    while len(out) < count:
        v = rng.integers(0, N, 3)
        if np.linalg.norm(v - N / 2) < 8:
            out.append(tuple(int(a) for a in v))
    return out


def test_adjoint_matches_central_differences(srt, setup, fields):
    import torch
    tr, phi, shift, target = setup
    print("[gpu] adjoint gradient", flush=True)
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS, checkpoint_every=K)
    L0, g = loss_and_grad(maps, target)
    _, gphi, gshift = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=K)
    gphi = gphi.cpu().numpy()
    gshift = gshift.cpu().numpy()
    rng = np.random.default_rng(34)
    checks = []
    for comp, vs in (("phi", voxels(rng, 8)), (0, voxels(rng, 4)), (1, voxels(rng, 4)), (2, voxels(rng, 4))):
        base = phi if comp == "phi" else shift[comp]
        h = FD_STEP[comp] * float(np.sqrt(np.mean(base.astype(np.float64) ** 2)))
        for (i, j, k) in vs:
            vals = []
            for sgn in (+1, -1):
                p = phi.copy()
                s = shift.copy()
                arr = p if comp == "phi" else s[comp]
                arr[k, j, i] = np.float32(float(base[k, j, i]) + sgn * h)
                vals.append((maps_of(tr, p, s), float(arr[k, j, i])))
            fd = loss_diff(vals[0][0], vals[1][0], target) / (vals[0][1] - vals[1][1])
            ad = gphi[k, j, i] if comp == "phi" else gshift[k, j, i, comp]
            checks.append((comp, (i, j, k), fd, ad))
    for c in checks:
        print(f"{c[0]} {c[1]}: fd {c[2]:+.10e} adjoint {c[3]:+.10e}")
    for group in (("phi",), (0, 1, 2)):
        sub = [c for c in checks if c[0] in group]
        scale = max(abs(c[2]) for c in sub)
        worst = max(abs(c[2] - c[3]) for c in sub) / scale
        print(f"{group}: worst |fd - adjoint| / max|fd| = {worst:.3e}")
        assert scale > 0
        assert worst < 1e-6


def test_adjoint_checkpoint_interval_invariance(srt, setup):
    tr, phi, shift, target = setup
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    _, g = loss_and_grad(maps, target)
    ref = None
    for k in (1, 5, STEPS):
        m, gp, gs = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=k)
        assert np.array_equal(m.cpu().numpy(), maps.cpu().numpy())
        cur = (gp.cpu().numpy(), gs.cpu().numpy())
        if ref is None:
            ref = cur
        else:  # same arithmetic, only the atomic summation order differs
            for a, b in zip(cur, ref):
                np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max())


def test_adjoint_forward_maps_match_fp64_kernel(srt, setup):
    tr, phi, shift, target = setup
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    maps64, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    maps32 = tr.trace(NSIDE, STEPS, precision=srt.PREC_FP64)
    np.testing.assert_allclose(maps64.cpu().numpy().astype(np.float32), maps32, rtol=3e-7, atol=1e-30)


def test_adjoint_vs_oracle_differences(srt, oracle, setup):
    """The adjoint gradient against central differences of the host FP64 oracle."""
    tr, phi, shift, target = setup
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    _, g = loss_and_grad(maps, target)
    _, gphi, gshift = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=K)
    tgt = target.cpu().numpy()
    pix = np.arange(12 * NSIDE * NSIDE)

    def oracle_maps(p, s):
        _, _, st = oracle.trace(oracle.Fields(p, s), NSIDE, pix, STEPS, return_state=True)
        return np.stack([st[:, 16], st[:, 17], st[:, 12]])

    rng = np.random.default_rng(35)
    res = []
    for comp, (i, j, k) in [("phi", v) for v in voxels(rng, 3)] + [(1, v) for v in voxels(rng, 2)]:
        base = phi if comp == "phi" else shift[comp]
        h = FD_STEP[comp] * float(np.sqrt(np.mean(base.astype(np.float64) ** 2)))
        vals = []
        for sgn in (+1, -1):
            p = phi.copy()
            s = shift.copy()
            arr = p if comp == "phi" else s[comp]
            arr[k, j, i] = np.float32(float(base[k, j, i]) + sgn * h)
            vals.append((oracle_maps(p, s), float(arr[k, j, i])))
        mp, mm = vals[0][0], vals[1][0]
        fd = float(((mp - mm) * (mp + mm - 2 * tgt)).sum()) / (vals[0][1] - vals[1][1])
        ad = float(gphi[k, j, i] if comp == "phi" else gshift[k, j, i, comp])
        res.append((fd, ad))
    scale = max(abs(a) for a, _ in res)
    worst = max(abs(a - b) for a, b in res) / scale
    print(f"oracle central differences vs adjoint: worst {worst:.3e}")
    assert worst < 1e-6


def test_autograd_wrapper(srt, setup):
    import torch
    tr, phi, shift, target = setup
    dev = torch.device("cuda:0")
    p = torch.from_numpy(phi).to(dev).double().requires_grad_(True)
    s = torch.from_numpy(np.ascontiguousarray(np.moveaxis(shift, 0, -1))).to(dev).double().requires_grad_(True)
    maps = srt.ray_maps(p, s, NSIDE, STEPS, tracer=tr, checkpoint_every=K)
    loss = ((maps[[1, 2, 0]] - target) ** 2).sum()
    loss.backward()
    L, g = loss_and_grad(maps.detach(), target)
    _, gphi, gshift = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=K)
    np.testing.assert_allclose(p.grad.cpu().numpy(), gphi.cpu().numpy(), rtol=1e-9, atol=1e-12 * float(gphi.abs().max()))
    np.testing.assert_allclose(s.grad.cpu().numpy(), gshift.cpu().numpy(), rtol=1e-9,
                               atol=1e-12 * float(gshift.abs().max()))


def test_adjoint_pixel_lists_partial_waves_and_invalid_pixels(srt, setup):
    """Gradients of a pixel-list loss do not depend on the list's order, on
    invalid entries (which trace nothing and get NaN maps) or on a ray count
    that leaves wavefronts partly empty."""
    import torch
    print("[gpu] adjoint over pixel lists", flush=True)
    tr, phi, shift, _ = setup
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    dev = torch.device("cuda:0")
    npix = 12 * NSIDE * NSIDE
    rng = np.random.default_rng(44)
    pix = rng.choice(npix, 100, replace=False).astype(np.int64)
    g = torch.from_numpy(rng.standard_normal((6, 100))).to(dev)

    def grads(p, gm):
        pt = torch.from_numpy(p).to(dev)
        m, gp, gs = tr.trace_adjoint(NSIDE, STEPS, gm, checkpoint_every=K, pix_list=pt, count=p.size)
        return m.cpu().numpy(), gp.cpu().numpy(), gs.cpu().numpy()

    m1, gp1, gs1 = grads(pix, g)
    # reversed order
    m2, gp2, gs2 = grads(pix[::-1].copy(), g.flip(1).contiguous())
    # three invalid entries inserted
    ins = np.array([5, 40, 77])
    pix3 = np.insert(pix, ins, [-1, npix, npix + 7])
    g3 = torch.from_numpy(np.insert(g.cpu().numpy(), ins, 1e3, axis=1)).to(dev)
    m3, gp3, gs3 = grads(pix3, g3)
    valid = np.ones(pix3.size, bool)
    valid[ins + np.arange(3)] = False
    assert np.all(np.isnan(m3[:, ~valid]))
    assert np.array_equal(m3[:, valid], m1) and np.array_equal(m2[:, ::-1], m1)
    for a, b in ((gp2, gp1), (gs2, gs1), (gp3, gp1), (gs3, gs1)):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_autograd_two_snapshots_on_one_tracer(srt, fields, setup):
    """Two ray_maps calls on one tracer (two snapshots) and one backward through a loss over both:
    each output's gradient is taken at its own snapshot, equal to separate adjoint calls."""
    import torch
    tr, phi, shift, target = setup
    dev = torch.device("cuda:0")
    phi2, shift2 = fields.snapshot(N, 36)
    grids = []
    for p_np, s_np in ((phi, shift), (phi2.numpy(), shift2.numpy())):
        p = torch.from_numpy(p_np).to(dev).double().requires_grad_(True)
        s = torch.from_numpy(np.ascontiguousarray(np.moveaxis(s_np, 0, -1))).to(dev).double().requires_grad_(True)
        grids.append((p, s))
    m1 = srt.ray_maps(*grids[0], NSIDE, STEPS, tracer=tr, checkpoint_every=K)
    m2 = srt.ray_maps(*grids[1], NSIDE, STEPS, tracer=tr, checkpoint_every=K)
    loss = ((m1[[1, 2, 0]] - target) ** 2).sum() + ((m2[[1, 2, 0]] - target) ** 2).sum()
    loss.backward()
    for (p, s), m in zip(grids, (m1, m2)):
        tr.upload(p.detach().float().contiguous(), s.detach().float().contiguous())
        _, g = loss_and_grad(m.detach(), target)
        _, gphi, gshift = tr.trace_adjoint(NSIDE, STEPS, g, checkpoint_every=K)
        np.testing.assert_allclose(p.grad.cpu().numpy(), gphi.cpu().numpy(), rtol=1e-9,
                                   atol=1e-12 * float(gphi.abs().max()))
        np.testing.assert_allclose(s.grad.cpu().numpy(), gshift.cpu().numpy(), rtol=1e-9,
                                   atol=1e-12 * float(gshift.abs().max()))
    assert not torch.allclose(grids[0][0].grad, grids[1][0].grad)


def test_adjoint_rejects_launches_of_2_to_the_32_lanes(srt, setup):
    """A ray count whose single adjoint launch would hold 2^32 lanes is refused before any device
    work or allocation; the output is left untouched."""
    import ctypes
    import torch
    tr, phi, shift, _ = setup
    tr.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    L = srt.lib()
    maps64 = torch.full((16,), 7.0, dtype=torch.float64, device="cuda:0")
    nside = 1 << 15  # 12 * 2^30 pixels
    for count in (1 << 32, (1 << 32) - 63):
        rc = L.srt_trace_adjoint(tr._ctx, nside, 0, count, None, 4, None, 4, None,
                                 ctypes.c_void_p(maps64.data_ptr()), None, None, None)
        assert rc == 6, (count, rc)
        assert b"2^32" in L.srt_last_error()
    torch.cuda.synchronize()
    assert bool((maps64 == 7.0).all())


def test_adjoint_in_ray_chunks(srt, setup, monkeypatch):
    """Rays beyond one chunk's work buffers are processed chunk after chunk: a context limited to
    128-ray chunks (6 chunks of the 768 rays, and 3 of a 300-entry pixel list with invalid entries, the
    last chunk partial) gives the maps bit for bit and the gradients to the atomics' summation order."""
    import torch
    print("[gpu] adjoint in ray chunks", flush=True)
    tr, phi, shift, target = setup
    monkeypatch.setenv("SRT_ADJ_CHUNK_RAYS", "128")
    small = srt.Tracer(0)
    monkeypatch.delenv("SRT_ADJ_CHUNK_RAYS")
    dev = torch.device("cuda:0")
    for t in (tr, small):
        t.upload(phi, np.ascontiguousarray(np.moveaxis(shift, 0, -1)))
    maps, _, _ = tr.trace_adjoint(NSIDE, STEPS)
    _, g = loss_and_grad(maps, target)
    rng = np.random.default_rng(46)
    pix = rng.choice(12 * NSIDE * NSIDE, 300, replace=False).astype(np.int64)
    pix[[7, 150, 299]] = [-1, 12 * NSIDE * NSIDE, -5]
    gl = torch.from_numpy(rng.standard_normal((6, pix.size))).to(dev)
    for kw, gm in ((dict(), g), (dict(pix_list=torch.from_numpy(pix).to(dev), count=pix.size), gl)):
        a = tr.trace_adjoint(NSIDE, STEPS, gm, checkpoint_every=K, **kw)
        b = small.trace_adjoint(NSIDE, STEPS, gm, checkpoint_every=K, **kw)
        ma, mb = a[0].cpu().numpy(), b[0].cpu().numpy()
        assert np.array_equal(np.isnan(ma), np.isnan(mb))
        assert np.array_equal(ma[~np.isnan(ma)], mb[~np.isnan(mb)])
        for x, y in zip(a[1:], b[1:]):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            assert np.abs(x - y).max() <= 1e-12 * np.abs(x).max()
    small.close()
