"""Pins of the oracle's HEALPix RING directions and binary16 conversion.

HEALPix (chealpix pix2ang_ring64, called at snapshot-raytracer.cu:865) is pinned
by the closed forms of Gorski et al. 2005, by ring structure, by the sum of all
pixel directions, and by a round trip through an independently written
ang2pix (numpy, below).  The binary16 conversion (the reference's FP16 x 2^27
shift storage, .cu:585-597) is pinned exhaustively against the x86 F16C
hardware conversion.
"""
import numpy as np
import pytest


def ang2pix_ring(nside, z, phi):
    """Independent RING ang2pix (vectorised, HEALPix conventions)."""
    n = nside
    za = np.abs(z)
    tt = np.mod(phi, 2 * np.pi) / (0.5 * np.pi)  # in [0, 4)
    ncap = 2 * n * (n - 1)
    npix = 12 * n * n
    pix = np.empty(z.shape, np.int64)
    eq = za <= 2.0 / 3.0
    # equatorial belt
    t1 = n * (0.5 + tt[eq])
    t2 = n * z[eq] * 0.75
    jp = np.floor(t1 - t2).astype(np.int64)
    jm = np.floor(t1 + t2).astype(np.int64)
    ir = n + 1 + jp - jm
    kshift = 1 - (ir & 1)
    ip = (jp + jm - n + kshift + 1) // 2
    ip = np.mod(ip, 4 * n)
    pix[eq] = ncap + (ir - 1) * 4 * n + ip
    # polar caps
    pc = ~eq
    tp = tt[pc] - np.floor(tt[pc])
    tmp = n * np.sqrt(3.0 * (1.0 - za[pc]))
    jp = np.floor(tp * tmp).astype(np.int64)
    jm = np.floor((1.0 - tp) * tmp).astype(np.int64)
    ir = jp + jm + 1
    ip = np.floor(tt[pc] * ir).astype(np.int64)
    ip = np.mod(ip, 4 * ir)
    north = z[pc] > 0
    pix[pc] = np.where(north, 2 * ir * (ir - 1) + ip, npix - 2 * ir * (ir + 1) + ip)
    return pix


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 16, 64, 1024])
def test_first_and_last_pixel(oracle, nside):
    z, s, phi = oracle.pix2zphi_ring(nside, [0, 12 * nside * nside - 1])
    assert z[0] == pytest.approx(1 - 1 / (3 * nside ** 2), abs=1e-15)
    assert phi[0] == pytest.approx(np.pi / 4, abs=1e-15)
    assert z[1] == pytest.approx(-1 + 1 / (3 * nside ** 2), abs=1e-15)
    assert phi[1] == pytest.approx(7 * np.pi / 4, abs=1e-14)
    assert s[0] == pytest.approx(np.sqrt(1 - z[0] ** 2), rel=1e-12)


@pytest.mark.parametrize("nside", [1, 2, 5, 8, 32])
def test_ring_structure_and_centroid(oracle, nside):
    npix = 12 * nside * nside
    z, s, phi = oracle.pix2zphi_ring(nside, np.arange(npix))
    # rings: 4 i pixels in the caps, 4 nside in the belt, z strictly decreasing by ring
    uz, counts = np.unique(-z, return_counts=True)
    assert uz.size == 4 * nside - 1
    i = np.arange(1, 4 * nside)
    expect = np.where(i < nside, 4 * i, np.where(i > 3 * nside, 4 * (4 * nside - i), 4 * nside))
    assert np.array_equal(counts, expect)
    # ring z values: 1 - i^2/(3n^2) in the cap, 4/3 - 2i/(3n) in the belt
    zr = np.where(i < nside, 1 - i ** 2 / (3 * nside ** 2),
                  np.where(i > 3 * nside, -1 + (4 * nside - i) ** 2 / (3 * nside ** 2), 4 / 3 - 2 * i / (3 * nside)))
    np.testing.assert_allclose(-uz, zr, atol=1e-15)
    assert np.all(np.diff(z) <= 0)
    np.testing.assert_allclose(s * s + z * z, 1.0, atol=1e-15)
    r = np.stack([s * np.cos(phi), s * np.sin(phi), z])
    assert np.abs(r.sum(axis=1)).max() < 1e-11 * npix
    assert np.all((phi >= 0) & (phi < 2 * np.pi))


@pytest.mark.parametrize("nside", [1, 2, 7, 16, 128])
def test_round_trip_with_independent_ang2pix(oracle, nside):
    npix = 12 * nside * nside
    pix = np.arange(npix)
    z, s, phi = oracle.pix2zphi_ring(nside, pix)
    assert np.array_equal(ang2pix_ring(nside, z, phi), pix)


def test_invalid_pixels_rejected(oracle):
    with pytest.raises(ValueError):
        oracle.pix2zphi_ring(4, [192])
    with pytest.raises(ValueError):
        oracle.pix2zphi_ring(4, [-1])


def test_f16_decode_all_codes(oracle):
    codes = np.arange(65536, dtype=np.uint16)
    mine = oracle.f16_to_f32(codes)
    ref = codes.view(np.float16).astype(np.float32)
    same = (mine.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(mine) & np.isnan(ref))
    assert same.all()


def test_f16_encode_matches_numpy_sample(oracle):
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, 2_000_000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[~np.isnan(x)]
    assert np.array_equal(oracle.f32_to_f16(x), x.astype(np.float16).view(np.uint16))


@pytest.mark.slow
def test_f16_encode_exhaustive_vs_f16c(oracle):
    """All 2^32 float bit patterns against the x86 F16C conversion (RNE)."""
    assert oracle.f16_check_range(0, 1 << 32) == 0


def test_shift_quantisation_scale(oracle):
    b = np.array([1e-7, -3e-7, 4.8e-4, 1e-12], np.float32)
    h = oracle.shift_to_f16(b)
    back = oracle.f16_to_f32(h) / np.float32(2 ** 27)
    np.testing.assert_allclose(back[:3], b[:3], rtol=1e-3)
    assert oracle.f16_to_f32(oracle.shift_to_f16(np.float32(5e-4)))[()] == np.inf  # overflow above ~4.88e-4
