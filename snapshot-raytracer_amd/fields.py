"""Seeded synthetic snapshots: Gaussian random fields with P(k) ~ k^-3.

The spec (identical on CPU and GPU; only FFT rounding differs between them):
  * white noise: one N(0, 1) value per voxel from a counter-based generator
    keyed by (seed, C-order voxel index): splitmix64 of the key gives two
    53-bit uniforms, Box-Muller gives the normal;
  * real FFT in FP64, multiplied by sqrt(P(k)) with
    P(k) = k^-3 exp(-(k / k_c)^2), k = 2 pi |m| / N, k_c = pi / 2 (half the
    Nyquist wavenumber); the k = 0 mode is set to zero;
  * inverse FFT, mean removed, rescaled to the exact requested rms, cast to
    float32.
The potential uses seed s and rms 1e-5; the shift components use seeds s+1,
s+2, s+3 and rms 1e-7.  Arrays are (nz, ny, nx) with x fastest, the layout the
ray tracer reads (HDF5 C order).
"""
from __future__ import annotations

import math

import torch

_MASK64 = (1 << 64) - 1


def _s64(v: int) -> int:
    """Two's-complement int64 value of an unsigned 64-bit constant."""
    v &= _MASK64
    return v - (1 << 64) if v >= (1 << 63) else v


_GOLDEN = _s64(0x9E3779B97F4A7C15)
_M1 = _s64(0xBF58476D1CE4E5B9)
_M2 = _s64(0x94D049BB133111EB)


def _shr(x: torch.Tensor, k: int) -> torch.Tensor:
    """Logical right shift of int64 bit patterns."""
    return (x >> k) & ((1 << (64 - k)) - 1)


def splitmix64(x: torch.Tensor) -> torch.Tensor:
    """splitmix64 finaliser on int64 tensors (wrapping arithmetic)."""
    z = x + _GOLDEN
    z = (z ^ _shr(z, 30)) * _M1
    z = (z ^ _shr(z, 27)) * _M2
    return z ^ _shr(z, 31)


def _seed_key(seed: int) -> int:
    z = (seed + 0x9E3779B97F4A7C15) & _MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return _s64(z ^ (z >> 31))


def white_noise(n: int, seed: int, device="cpu", chunk: int = 1 << 24) -> torch.Tensor:
    """N(0, 1) float64 array of shape (n, n, n) keyed by (seed, C-order index)."""
    total = n * n * n
    out = torch.empty(total, dtype=torch.float64, device=device)
    key = _seed_key(seed)
    two_pi = 2.0 * math.pi
    for start in range(0, total, chunk):
        stop = min(total, start + chunk)
        idx = torch.arange(start, stop, dtype=torch.int64, device=device)
        h1 = splitmix64((2 * idx) ^ key)
        h2 = splitmix64((2 * idx + 1) ^ key)
        u1 = (_shr(h1, 11).to(torch.float64) + 0.5) * (2.0 ** -53)
        u2 = (_shr(h2, 11).to(torch.float64) + 0.5) * (2.0 ** -53)
        out[start:stop] = torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(two_pi * u2)
    return out.view(n, n, n)


def gaussian_field(n: int, seed: int, rms: float, device="cpu", slope: float = -3.0) -> torch.Tensor:
    """Float32 (n, n, n) periodic Gaussian random field, P(k) ~ k^slope exp(-(k/k_c)^2)."""
    noise = white_noise(n, seed, device=device)
    fk = torch.fft.rfftn(noise)
    del noise
    m_full = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    m_half = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    k2 = (m_full.view(-1, 1, 1) ** 2 + m_full.view(1, -1, 1) ** 2 + m_half.view(1, 1, -1) ** 2)
    k2 = k2 * (2.0 * math.pi / n) ** 2
    kc = 0.5 * math.pi
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0) * torch.exp(-0.5 * k2 / (kc * kc))
    amp[0, 0, 0] = 0.0
    fk *= amp
    del amp, k2
    f = torch.fft.irfftn(fk, s=(n, n, n))
    del fk
    f -= f.mean()
    f *= rms / torch.sqrt(torch.mean(f * f))
    return f.to(torch.float32)


def snapshot(n: int, seed: int, phi_rms: float = 1e-5, shift_rms: float = 1e-7, device="cpu"):
    """(phi (n,n,n) float32, shift (3,n,n,n) float32 planar components)."""
    phi = gaussian_field(n, seed, phi_rms, device=device)
    shift = torch.stack([gaussian_field(n, seed + 1 + a, shift_rms, device=device) for a in range(3)])
    return phi, shift


def periodic_field(n: int, period: int = 128, device="cpu"):
    """A field of period `period` cells along every axis, for grids too large to generate spectrally:
    phi (n, n, n) a sum of one-dimensional modes along z, y and x (amplitudes 1e-5, 0.7e-5, 0.5e-5), and the
    interleaved shift (n, n, n, 3) with B_x varying along y, B_y along z and B_z along x (amplitudes 1e-7,
    0.8e-7, 0.6e-7), float32, arrays (z, y, x) with x fastest.  An n = k * period grid is the period^3 grid
    tiled k^3 times with bit-identical values."""
    t = torch.arange(period, dtype=torch.float64, device=device) * (2 * math.pi / period)
    ax = (1e-5 * torch.sin(t)).float()
    ay = (0.7e-5 * torch.cos(2 * t + 0.4)).float()
    az = (0.5e-5 * torch.sin(3 * t + 1.1)).float()
    bx = (1e-7 * torch.cos(t + 0.2)).float()
    by = (0.8e-7 * torch.sin(2 * t)).float()
    bz = (0.6e-7 * torch.cos(t + 2.0)).float()
    idx = torch.arange(n, device=device) % period
    phi = az[idx].view(n, 1, 1) + ay[idx].view(1, n, 1) + ax[idx].view(1, 1, n)
    shift = torch.empty((n, n, n, 3), dtype=torch.float32, device=device)
    shift[..., 0] = by[idx].view(1, n, 1)
    shift[..., 1] = bz[idx].view(n, 1, 1)
    shift[..., 2] = bx[idx].view(1, 1, n)
    return phi, shift


def interleave_shift(shift: torch.Tensor) -> torch.Tensor:
    """(3, n, n, n) planar -> (n, n, n, 3) interleaved, the HDF5 H5T_ARRAY{3} layout."""
    return shift.permute(1, 2, 3, 0).contiguous()


def uneven_batch_params(count: int = 48, rng_seed: int = 7):
    """Parameters of the batched uneven workload: `count` snapshot/observer instances.

    Instance i (numpy default_rng(rng_seed), drawn in this order): grid edge n
    from {128, 192, 256, 384}, Nside from {128, 256, 512}, one RING batch b in
    [0, 12), n_steps uniform in [64, n / 2], observer uniform in [0, n)^3;
    snapshot seed 1000 + 4 i.  Returns a list of dicts with keys n, nside,
    batch, n_steps, observer (3 floats), seed.
    """
    import numpy as np

    rng = np.random.default_rng(rng_seed)
    out = []
    for i in range(count):
        n = int(rng.choice([128, 192, 256, 384]))
        nside = int(rng.choice([128, 256, 512]))
        b = int(rng.integers(0, 12))
        n_steps = int(rng.integers(64, n // 2 + 1))
        obs = rng.uniform(0, n, 3)
        out.append(dict(n=n, nside=nside, batch=b, n_steps=n_steps, observer=obs, seed=1000 + 4 * i))
    return out
