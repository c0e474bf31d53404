"""The command-line drop-in end to end on the GPU: HDF5 snapshot in, six RING
FITS maps out (snapshot-raytracer.cu:600-1006, README.md:27-49).  Maps must
equal the library's own trace of the same pixels bit for bit, be exactly zero
outside the batch, and match the FP64 oracle within the kernel's tolerance."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "snapshot-raytracer_amd", "bin", "snapshot-raytracer")
NAMES = ("dAmap", "ellipticity1map", "ellipticity2map", "rotationmap", "deflection1map", "deflection2map")


@pytest.fixture(scope="module")
def files(tmp_path_factory, fields):
    from snapshot_raytracer_amd import snapshot_io
    if not os.path.exists(CLI):
        pytest.skip("CLI not built (no HDF5)")
    d = tmp_path_factory.mktemp("cli")
    phi, shift = fields.snapshot(32, 21)
    snapshot_io.write_potential(str(d / "phi.h5"), phi.numpy())
    snapshot_io.write_shift(str(d / "B.h5"), fields.interleave_shift(shift).numpy())
    return d, phi.numpy(), shift.numpy(), fields.interleave_shift(shift).numpy()


def run(d, args, env=None):
    print(f"[gpu] cli {args} {env or ''}", flush=True)
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([CLI, "phi.h5", "B.h5"] + args, cwd=d, capture_output=True, text=True, timeout=120,
                          env=e)


@pytest.mark.parametrize("precision,layout", [("mixed", "auto"), ("mixed", "taps"), ("fp64", "auto")])
def test_cli_full_sky_matches_library(srt, files, precision, layout):
    from snapshot_raytracer_amd import snapshot_io
    d, phi, shift, shift_il = files
    nside, steps = 4, 10
    r = run(d, [str(nside), str(steps)], {"SRT_PRECISION": precision, "SRT_LAYOUT": layout})
    assert r.returncode == 0, r.stderr
    assert "All done!" in r.stdout
    tr = srt.Tracer(0)
    if precision == "mixed" and layout == "auto":  # a 32^3 snapshot's records fit: the CLI reads them
        tr.set_layout(srt.LAYOUT_BICUBIC, 8)
    tr.upload(phi, shift_il)
    ref = tr.trace(nside, steps, precision=srt.PRECISIONS[precision])
    tr.close()
    for m, name in enumerate(NAMES):
        got, hdr = snapshot_io.read_healpix_fits(str(d / f"{name}_batch0.fits"))
        assert hdr["NSIDE"] == nside and hdr["ORDERING"] == "RING" and hdr["COORDSYS"] == "C"
        assert np.array_equal(got, ref[m]), name


def test_cli_batch_masks_and_oracle(srt, oracle, files):
    from snapshot_raytracer_amd import snapshot_io
    d, phi, shift, shift_il = files
    nside, steps, batch = 8, 12, 7
    r = run(d, [str(nside), str(steps), str(batch)], {"SRT_PRECISION": "fp64"})
    assert r.returncode == 0, r.stderr
    lo, cnt = srt.batch_range(nside, batch)
    ref, _ = oracle.trace(oracle.Fields(phi, shift), nside, np.arange(lo, lo + cnt), steps)
    for m, name in enumerate(NAMES):
        got, _ = snapshot_io.read_healpix_fits(str(d / f"{name}_batch{batch}.fits"))
        assert np.all(got[:lo] == 0) and np.all(got[lo + cnt:] == 0)
        np.testing.assert_allclose(got[lo:lo + cnt], ref[m], rtol=3e-7, atol=1e-30)


def test_cli_reference_storage_mode(srt, oracle, files):
    """SRT_PRECISION=ref with SRT_SHIFT=f16, the reference's arithmetic on its binary16 x 2^27 shift storage
    (.cu:585-597, 241-244): the FITS maps equal the library's trace in that mode bit for bit, and the oracle's REF
    form on the same binary16 codes within the REF bars (dA 1e-6 relative, other maps 2e-5 of their rms).  Unknown
    values of the three variables exit with status 255."""
    from snapshot_raytracer_amd import snapshot_io
    d, phi, shift, shift_il = files
    nside, steps = 4, 6
    r = run(d, [str(nside), str(steps)], {"SRT_PRECISION": "ref", "SRT_SHIFT": "f16"})
    assert r.returncode == 0, r.stderr
    tr = srt.Tracer(0)
    tr.upload(phi, shift_il, shift_format=srt.SHIFT_F16_SCALED)
    lib = tr.trace(nside, steps, precision=srt.PREC_REF)
    tr.close()
    F = oracle.Fields(phi, shift, b16=oracle.shift_to_f16(shift))
    ref, _ = oracle.trace(F, nside, np.arange(12 * nside * nside), steps, form=oracle.FORM_REF)
    for m, name in enumerate(NAMES):
        got, _ = snapshot_io.read_healpix_fits(str(d / f"{name}_batch0.fits"))
        assert np.array_equal(got, lib[m]), name
        g, e = got.astype(np.float64), ref[m].astype(np.float64)
        if m == 0:
            assert np.max(np.abs(g - e) / np.abs(e)) < 1e-6
        else:
            assert np.max(np.abs(g - e)) / np.sqrt(np.mean(e ** 2)) < 2e-5, name
    for var in ("SRT_PRECISION", "SRT_SHIFT", "SRT_LAYOUT"):
        r = run(d, ["4", "6"], {var: "bogus"})
        assert r.returncode == 255, var


def test_cli_records_that_do_not_fit(srt, files):
    """SRT_LAYOUT=auto falls back to the padded grids when the bicubic records do
    not fit (forced with a per-allocation limit) and says so; SRT_LAYOUT=bicubic
    fails with exit status 255."""
    from snapshot_raytracer_amd import snapshot_io
    d, phi, shift, shift_il = files
    limit = {"SRT_MAX_ALLOC_BYTES": str(4 << 20)}  # 32^3: padded grids 0.7 MB fit, records 9.2 MB do not
    r = run(d, ["4", "6"], dict(limit, SRT_LAYOUT="auto"))
    assert r.returncode == 0, r.stderr
    assert "do not fit" in r.stdout
    tr = srt.Tracer(0)
    tr.upload(phi, shift_il)
    ref = tr.trace(4, 6, precision=srt.PREC_MIXED)
    tr.close()
    got, _ = snapshot_io.read_healpix_fits(str(d / "dAmap_batch0.fits"))
    assert np.array_equal(got, ref[0])
    r = run(d, ["4", "6"], dict(limit, SRT_LAYOUT="bicubic"))
    assert r.returncode == 255


def test_cli_headline_scale(srt, fields, tmp_path):
    """The drop-in at the benchmark's scale: the 512^3 snapshot (seed 2) as the two HDF5 files, the full sky at
    Nside 1024 with 256 steps, then batch 5 alone.  Every FITS map equals the library's trace of the same
    snapshot bit for bit (mixed kernel on the bicubic records, which the CLI builds when they fit), and the
    batch's maps are exactly zero outside its RING range."""
    from snapshot_raytracer_amd import snapshot_io
    if not os.path.exists(CLI):
        pytest.skip("CLI not built (no HDF5)")
    n, nside, steps, batch = 512, 1024, 256, 5
    print(f"[gpu] cli headline scale: {n}^3 snapshot to HDF5", flush=True)
    phi, shift = fields.snapshot(n, 2, device="cuda:0")
    phi_h = phi.cpu().numpy()
    shift_h = fields.interleave_shift(shift).cpu().numpy()
    del phi, shift
    snapshot_io.write_potential(str(tmp_path / "phi.h5"), phi_h)
    snapshot_io.write_shift(str(tmp_path / "B.h5"), shift_h)
    r = run(tmp_path, [str(nside), str(steps)])
    assert r.returncode == 0, r.stderr
    assert "Timing: HDF5 read" in r.stdout and r.stdout.rstrip().endswith("All done!")
    r_b = run(tmp_path, [str(nside), str(steps), str(batch)])
    assert r_b.returncode == 0, r_b.stderr
    tr = srt.Tracer(0)
    tr.set_layout(srt.LAYOUT_BICUBIC, 16)
    tr.upload(phi_h, shift_h)
    ref = tr.trace(nside, steps, precision=srt.PREC_MIXED, order=srt.ORDER_NESTED)
    tr.close()
    lo, cnt = srt.batch_range(nside, batch)
    for m, name in enumerate(NAMES):
        got, hdr = snapshot_io.read_healpix_fits(str(tmp_path / f"{name}_batch0.fits"))
        assert hdr["NSIDE"] == nside and got.size == 12 * nside * nside
        assert np.array_equal(got, ref[m]), name
        got_b, _ = snapshot_io.read_healpix_fits(str(tmp_path / f"{name}_batch{batch}.fits"))
        assert np.all(got_b[:lo] == 0) and np.all(got_b[lo + cnt:] == 0), name
        assert np.array_equal(got_b[lo:lo + cnt], ref[m][lo:lo + cnt]), name


def test_cli_batches_add_up_to_full_sky(files):
    """README.md:41-42 of the reference: batch outputs are zero outside their base-pixel range, "so the maps can
    simply be added together".  The 12 batch runs' FITS maps, summed, equal the full-sky run's bit for bit
    (each pixel is non-zero in exactly one batch), at a Nside that is not a power of two."""
    from snapshot_raytracer_amd import snapshot_io
    d = files[0]
    nside, steps = 6, 8
    r = run(d, [str(nside), str(steps)])
    assert r.returncode == 0, r.stderr
    full = [snapshot_io.read_healpix_fits(str(d / f"{name}_batch0.fits"))[0].copy() for name in NAMES]
    total = [np.zeros_like(m) for m in full]
    for b in range(12):
        r = run(d, [str(nside), str(steps), str(b)])
        assert r.returncode == 0, r.stderr
        for m, name in enumerate(NAMES):
            got, hdr = snapshot_io.read_healpix_fits(str(d / f"{name}_batch{b}.fits"))
            assert hdr["NSIDE"] == nside
            total[m] += got
    for m, name in enumerate(NAMES):
        assert np.array_equal(total[m], full[m]), name
    r = run(d, [str(nside), str(steps), "12"])  # beyond the last batch: the reference's message and status
    assert r.returncode == 255 and "Batch exceeds total number of pixels" in r.stderr
