"""Pins of the oracle's shift-coupling terms (snapshot-raytracer.cu:370-467)
by a pure-gradient shift, where first-order physics fixes the answer.

With phi = 0 and B = grad chi (a static chi: a pure coordinate mode at first
order), the right-hand side's shift terms must cancel in a known way:

* Deflection.  N_j = sum_a n_a d_j B_a = (H_chi n)_j, so n' = -P N = (H_chi n)_perp
  = -d/dlambda [grad_perp chi] along the ray (x' = -n), and
  n_perp(lambda) = -[grad_perp chi(x) - grad_perp chi(x_obs)].  The drift
  x' = -n - B then gives x_perp' = -grad_perp chi(x_obs): the deflection maps
  equal the observer-only dipole, -grad chi(x_obs) . theta-hat and
  -grad chi(x_obs) . (sin phi, -cos phi, 0), whatever chi does along the ray.
  This fixes N's sign and coefficient against the -B drift (.cu:465-471) and
  the half-cell staggering of B_a (.cu:377, 406, 432).
* Shear.  T_ij = sum_a n_a d_i d_j B_a - 1/2 [(n H^i)_j + (n H^j)_i] vanishes
  identically for B = grad chi, and the potential Hessian is zero, so sigma
  (and with it epsilon) stays zero at first order (.cu:389-452, 484-495).
  The same amplitude of a curl shift B = curl(chi z-hat) shears at first order.
* Frame.  d(n . e1)/dlambda = N . e1 - E1 at first order, and for a
  symmetric shift gradient E1 = 1/2 sum_a [e1_a (n . grad B_a) + n_a (e1 . grad B_a)]
  equals N . e1, so the screen vectors stay orthogonal to n (.cu:383-387,
  473-479).  This pins the symmetric part of E1, E2 (their factor 1/2 and
  the sum of the two terms); their antisymmetric part, L (nLaplaceB, which
  enters only D'') and T beyond its gradient-shift cancellation are pinned by
  a divergence-free shift in test_oracle_shift_gr.py.

The shift samples are exact: B_a at grid node g is d_a chi(g + e_a / 2), the
point the oracle samples when it evaluates B_a at x - e_a / 2.  chi is a sum
of five long-wave periodic modes, so the Catmull-Rom interpolation error is
~2e-4 of the dipole.

test_planted_mistakes_fail compiles variants of the oracle with one plausible
mistake each (signs, factors, staggering, transposed combinations) and shows
that every one of them fails at least one of these checks.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

N, NSIDE, STEPS = 96, 4, 32
OBS = np.array([41.3, 52.1, 45.6])
MODES = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, -1)]

# bars: deflection vs the dipole (measured 1.8e-4; the planted mistakes that
# move it: >= 3.4e-2), gradient-shift shear vs curl-shift shear (measured
# 0.46 %; mistakes in T or the staggering: >= 4.0 %), frame orthogonality
# |n . e| vs the dipole (measured 1.5e-4; mistakes in E1 or N: >= 0.76)
DEFLECTION_BAR = 1e-3
SHEAR_RATIO_BAR = 0.03
FRAME_BAR = 1e-2


class Chi:
    """chi(x) = sum_m c_m cos(2 pi m . x / N + a_m) and its gradient."""

    def __init__(self, seed=123):
        rng = np.random.default_rng(seed)
        self.c = rng.standard_normal(len(MODES))
        self.a = rng.uniform(0, 2 * np.pi, len(MODES))
        self.k = [2 * np.pi * np.array(m, dtype=np.float64) / N for m in MODES]

    def grad(self, pts):
        g = np.zeros(pts.shape)
        for c, a, k in zip(self.c, self.a, self.k):
            g += (-c * np.sin(pts @ k + a))[..., None] * k
        return g


def shift_grids(chi, kind, amp):
    """(3, N, N, N) float32 staggered samples of B = grad chi ('grad') or curl(chi z-hat) ('curl')."""
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    nodes = np.stack([x, y, z], -1).astype(np.float64)
    out = np.zeros((3, N, N, N))
    for a in range(3):
        p = nodes.copy()
        p[..., a] += 0.5  # B_a at node g is sampled at g + e_a / 2
        g = chi.grad(p)
        if kind == "grad":
            out[a] = g[..., a]
        elif a == 0:
            out[a] = g[..., 1]
        elif a == 1:
            out[a] = -g[..., 0]
    return (out * amp).astype(np.float32)


def frame_overlap(st):
    """max over rays of |n . e1|, |n . e2| of the final (unit-normalised) frame."""
    n = st[:, 3:6] / np.linalg.norm(st[:, 3:6], axis=1, keepdims=True)
    e1 = st[:, 6:9] / np.linalg.norm(st[:, 6:9], axis=1, keepdims=True)
    e2 = st[:, 9:12] / np.linalg.norm(st[:, 9:12], axis=1, keepdims=True)
    return max(np.abs((n * e1).sum(1)).max(), np.abs((n * e2).sum(1)).max())


@pytest.fixture(scope="module")
def setup(oracle):
    chi = Chi()
    amp = 1e-7 / np.sqrt(np.mean(shift_grids(chi, "grad", 1.0).astype(np.float64) ** 2))
    pix = np.arange(12 * NSIDE * NSIDE)
    z, s, ph = oracle.pix2zphi_ring(NSIDE, pix)
    th = np.stack([z * np.cos(ph), z * np.sin(ph), -s], 1)
    p2 = np.stack([np.sin(ph), -np.cos(ph), 0 * ph], 1)
    g_obs = chi.grad(OBS[None])[0] * amp
    dipole = (-(th @ g_obs), -(p2 @ g_obs))
    grids = {(kind, f): shift_grids(chi, kind, amp * f) for kind in ("grad", "curl") for f in (1.0, 2.0)}
    return dict(pix=pix, dipole=dipole, grids=grids, phi=np.zeros((N, N, N), np.float32))


def run(oracle, setup, kind, f=1.0, library=None):
    F = oracle.Fields(setup["phi"], setup["grids"][(kind, f)])
    maps, _, st = oracle.trace(F, NSIDE, setup["pix"], STEPS, observer=OBS, return_state=True, library=library)
    return maps, st


def checks(oracle, setup, library=None):
    """(deflection error / dipole rms, grad-shift eps rms / curl-shift eps rms, max |n.e| / dipole rms)."""
    maps, st = run(oracle, setup, "grad", library=library)
    curl, _ = run(oracle, setup, "curl", library=library)
    d1, d2 = setup["dipole"]
    rms = np.sqrt(np.mean(d1 ** 2 + d2 ** 2))
    defl = max(np.abs(maps[4] - d1).max(), np.abs(maps[5] - d2).max()) / rms

    def eps_rms(m):
        return np.sqrt(np.mean(m[1].astype(np.float64) ** 2 + m[2].astype(np.float64) ** 2))

    return defl, eps_rms(maps) / eps_rms(curl), frame_overlap(st) / rms


def test_gradient_shift_deflection_is_observer_dipole(oracle, setup):
    defl, _, _ = checks(oracle, setup)
    d1, d2 = setup["dipole"]
    assert np.sqrt(np.mean(d1 ** 2 + d2 ** 2)) > 1e-8  # the dipole itself is not trivially small
    print(f"deflection - dipole: {defl:.2e} of the dipole rms")
    assert defl < DEFLECTION_BAR


def test_gradient_shift_leaves_no_first_order_shear(oracle, setup):
    _, ratio, _ = checks(oracle, setup)
    curl, _ = run(oracle, setup, "curl")
    assert np.abs(curl[1]).max() > 1e-8  # the curl shift does shear
    print(f"eps rms, gradient / curl shift: {ratio:.3e}")
    assert ratio < SHEAR_RATIO_BAR


def test_gradient_shift_keeps_screen_frame_orthogonal(oracle, setup):
    _, _, overlap = checks(oracle, setup)
    _, st_curl = run(oracle, setup, "curl")
    d1, d2 = setup["dipole"]
    rms = np.sqrt(np.mean(d1 ** 2 + d2 ** 2))
    print(f"max |n.e| / dipole rms: gradient {overlap:.2e}, curl {frame_overlap(st_curl) / rms:.2e}")
    assert overlap < FRAME_BAR
    assert frame_overlap(st_curl) / rms > 0.1  # the check can see an E1/E2 mismatch


@pytest.mark.parametrize("kind", ["grad", "curl"])
def test_shift_maps_linear_in_amplitude(oracle, setup, kind):
    m1, s1 = run(oracle, setup, kind, 1.0)
    m2, s2 = run(oracle, setup, kind, 2.0)
    for m in (4, 5):  # deflection
        r = np.linalg.norm(m2[m].astype(np.float64)) / np.linalg.norm(m1[m].astype(np.float64))
        assert abs(r - 2) < 1e-4, (kind, m, r)
    # dA - n_steps from the FP64 state (the float32 map holds dA ~ 32 to 4e-6)
    r = np.linalg.norm(s2[:, 12] - STEPS) / np.linalg.norm(s1[:, 12] - STEPS)
    assert abs(r - 2) < 1e-3, (kind, r)
    if kind == "curl":  # the gradient shift's shear is second order plus interpolation error
        for m in (1, 2):
            r = np.linalg.norm(m2[m].astype(np.float64)) / np.linalg.norm(m1[m].astype(np.float64))
            assert abs(r - 2) < 1e-4, (kind, m, r)


# one plausible mistake each in the oracle's shift coupling (.cu:370-471)
MUTANTS = {
    "N sign": ("N[j] = I(S(N[j]) + g[j] * n[a])", "N[j] = I(S(N[j]) - g[j] * n[a])"),
    "N halved": ("N[j] = I(S(N[j]) + g[j] * n[a])", "N[j] = I(S(N[j]) + S(0.5) * g[j] * n[a])"),
    "B unstaggered": ("p[a] = p[a] - S(0.5);", "p[a] = p[a];"),
    "B staggered the wrong way": ("p[a] = p[a] - S(0.5);", "p[a] = p[a] + S(0.5);"),
    "drift sign": ("- S(Bv[i]);", "+ S(Bv[i]);"),
    "E1 without 1/2": ("E1 = I(S(E1) + S(0.5) * e1[a] * ng + S(0.5) * n[a] * e1g);",
                       "E1 = I(S(E1) + e1[a] * ng + n[a] * e1g);"),
    "E1 antisymmetrised": ("E1 = I(S(E1) + S(0.5) * e1[a] * ng + S(0.5) * n[a] * e1g);",
                           "E1 = I(S(E1) + S(0.5) * e1[a] * ng - S(0.5) * n[a] * e1g);"),
    "T diagonal": ("c = c - nH[a];", "c = c - S(0.5) * nH[a];"),
    "T off-diagonal": ("c = c - S(0.5) * nH[j];", "c = c - nH[j];"),
    "n' with 2 N": ("(S(2) * g[j] + S(N[j]))", "(S(2) * g[j] + S(2) * S(N[j]))"),
}


def test_planted_mistakes_fail(oracle, setup, tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src_dir = os.path.dirname(oracle.__file__)
    src = open(os.path.join(src_dir, "srt_oracle.hpp")).read()
    good = checks(oracle, setup)
    assert good[0] < DEFLECTION_BAR and good[1] < SHEAR_RATIO_BAR and good[2] < FRAME_BAR
    for name, (before, after) in MUTANTS.items():
        assert src.count(before) == 1, f"mutation site of '{name}' not found once in srt_oracle.hpp"
        d = tmp_path / name.replace(" ", "_").replace("'", "").replace("/", "")
        d.mkdir()
        (d / "srt_oracle.hpp").write_text(src.replace(before, after))
        shutil.copy(os.path.join(src_dir, "oracle_capi.cpp"), d)
        so = d / "liboracle_mutant.so"
        subprocess.run(["g++", "-O2", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o",
                        str(so), str(d / "oracle_capi.cpp")], check=True)
        defl, ratio, overlap = checks(oracle, setup, library=oracle.bind(str(so)))
        print(f"{name:28s} deflection {defl:.2e}  shear ratio {ratio:.2e}  frame {overlap:.2e}")
        assert defl > 10 * DEFLECTION_BAR or ratio > SHEAR_RATIO_BAR or overlap > 10 * FRAME_BAR, name
