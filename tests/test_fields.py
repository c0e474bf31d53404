"""The seeded synthetic snapshot generator (benchmark and test inputs)."""
import numpy as np
import pytest
import torch

M64 = (1 << 64) - 1


def splitmix_py(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_splitmix_matches_integer_reference(fields):
    xs = [0, 1, 2, 12345678901234, (1 << 63) - 1, (1 << 63) + 5, M64]
    t = fields.splitmix64(torch.tensor([fields._s64(x) for x in xs], dtype=torch.int64))
    assert [int(v) & M64 for v in t] == [splitmix_py(x) for x in xs]


def test_white_noise_is_keyed_and_standard(fields):
    a = fields.white_noise(32, 5)
    b = fields.white_noise(32, 5, chunk=1000)  # chunking does not change values
    c = fields.white_noise(32, 6)
    assert torch.equal(a, b)
    assert not torch.equal(a, c)
    v = a.flatten().numpy()
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1) < 0.01
    assert abs(np.mean(v ** 4) - 3) < 0.1  # Gaussian kurtosis


def test_field_rms_mean_and_spectrum(fields):
    n = 64
    f = fields.gaussian_field(n, 3, 1e-5).double().numpy()
    assert abs(np.sqrt(np.mean(f * f)) - 1e-5) < 1e-5 * 1e-5
    assert abs(f.mean()) < 1e-12
    fk = np.fft.rfftn(f)
    m = np.fft.fftfreq(n, 1.0 / n)
    mh = np.fft.rfftfreq(n, 1.0 / n)
    k = 2 * np.pi / n * np.sqrt(m[:, None, None] ** 2 + m[None, :, None] ** 2 + mh[None, None, :] ** 2)
    p = np.abs(fk) ** 2
    model = k ** -3.0 * np.exp(-(k / (np.pi / 2)) ** 2)
    # |noise|^2 is exponential with a constant mean: P(k) / model is flat; shells with
    # >1000 modes keep the sampling error of each mean near 3%
    shells = [((k > a) & (k < b)) for a, b in ((0.4, 0.8), (0.8, 1.2), (1.2, 2.0))]
    means = [float((p[s] / model[s]).mean()) for s in shells]
    assert max(means) / min(means) < 1.08, means
    assert p[0, 0, 0] < 1e-20


def test_snapshot_shapes_and_interleave(fields):
    phi, shift = fields.snapshot(16, 1)
    assert phi.shape == (16, 16, 16) and shift.shape == (3, 16, 16, 16)
    assert phi.dtype == torch.float32
    il = fields.interleave_shift(shift)
    assert il.shape == (16, 16, 16, 3)
    assert torch.equal(il[3, 4, 5], shift[:, 3, 4, 5])
    assert abs(float(shift.std()) - 1e-7) < 1e-9
