"""Pins of the oracle's right-hand side (snapshot-raytracer.cu:315-510) by
linearised general relativity: the first-order maps of a metric perturbation,
computed here from the metric alone (Christoffel symbols, Riemann tensor,
geodesic and geodesic-deviation equations), without the oracle's equations.

Metric (static, grid units):
    ds^2 = -(1 + 2 phi) dt^2 - 2 B_i dx^i dt + (1 - 2 phi) dx^2,
i.e. h_00 = -2 phi, h_0i = -B_i, h_ij = -2 phi delta_ij.  The oracle's state
follows one past-directed null geodesic k^mu from the observer:

* n is the direction of the covariant spatial momentum, n = -k_i / |k_i|, and
  the ray parameter s has ds/dv = |k_i| (v affine): then dx/ds = k^i / |k_i|
  = -n (1 + 2 phi) - B at first order, the reference's position equation
  (.cu:465-467), which fixes s.  The initial momentum is k_i = -n0 with k_0
  from the null condition.
* Deflection: the geodesic d(dk^mu)/dv = -Gamma^mu_ab k^a k^b integrated
  twice gives the transverse displacement at s = n_steps.  A divergence-free
  ("curl") shift deflects by the gravitomagnetic force only, through the full
  (not only the symmetric) shift gradient: this pins N_j = sum_a n_a d_j B_a
  and the -B drift beyond what a gradient shift can see.
* dA: the Sachs focusing equation D'' = -R_kk D / 2 (R_kk = Ricci(k, k), here
  from the Riemann tensor traced over the screen) in the affine parameter,
  read at v(s = n_steps) (the n . N rescaling of the reference's D'' comes
  from ds/dv), and in grid units: the screen lengths of the spatial metric
  (1 - 2 phi) are (1 - phi) x the coordinate lengths, so the grid-unit D is
  (1 + phi(x) - phi(x_obs)) x the physical one (normalised at the observer).
  For the curl shift, R_kk = n . laplacian(B): this pins L (nLaplaceB, .cu:396,
  422, 452, 482) and its coefficient 1/2.
* Ellipticity: the Jacobi map J'' = -R J, J(0) = 0, J'(0) = 1, with the
  optical tidal matrix R_ab = Riemann(e_a, k, e_b, k) in the screen basis
  (e1, e2) of the initial conditions (.cu:560).  The reference's complex
  ellipticity (README: arXiv:2002.04024) is a/b - b/a of the image axes, 4x
  the shear of J at first order: eps1 = 2 (A11 - A22), eps2 = 2 (A12 + A21)
  with J(S) = S (1 + A).  For the curl shift this pins T (tidalBterm,
  .cu:389-452) in full: a gradient shift has T = 0 whatever the coefficients
  that cancel there, so only a curl shift shows T itself.  For phi alone it
  pins the potential's tidal coupling.
* Frame: parallel transport of the screen vector e (e^0 = 0 at the observer)
  keeps e . k = 0; e^0 grows by Gamma^0_jl e^j k^l, and removing it by a
  multiple of k adds -n E1 with E1 = 1/2 e1 . (grad B + grad B^T) . n,
  symmetric in the shift gradient because Gamma^0_jl is symmetric in j, l.
  The reference's e' (.cu:473-479) has that term but not the rotation of e
  by Gamma^i_j0 e^j k^0 (the gravitomagnetic precession 1/2 curl B x e, which
  turns e within the screen and moves no map at first order), so its n . e
  drifts by exactly the n-component of that rotation.  This pins the
  antisymmetric part of E1, E2.

The curvature and Christoffel symbols are evaluated from the oracle's own
interpolation of the grids along the unperturbed ray (the interpolation is
pinned separately against the Keys kernel), so the checks test the
right-hand side and the integrator, not the interpolant.  The remaining
differences (~1e-2 of rms in the maps driven by second derivatives) are the
RK4 step h = 1 across the Catmull-Rom Hessian's jumps at cell faces; the
finely integrated reference equations agree with these GR integrals to 2e-4.

test_planted_mistakes_fail compiles the oracle with one plausible mistake in
these terms each and shows that every one fails a check by > 10x its bar.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

N, NSIDE, STEPS = 96, 4, 32
OBS = np.array([41.3, 52.1, 45.6])
MODES = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, -1), (1, 0, 1)]
ETA = np.array([-1.0, 1.0, 1.0, 1.0])
QPS = 64  # quadrature points per unit of the ray parameter
HIDX = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]

# bars, as max |oracle - GR| / rms(GR) over the 192 rays (measured: deflection
# <= 8.1e-5, frame <= 8.2e-5, dA <= 1.3e-2, ellipticity <= 2.8e-2; the planted
# mistakes move the check they fail by >= 0.79)
BAR_DEFLECTION = 1e-3
BAR_FRAME = 1e-3
BAR_DA = 5e-2
BAR_EPS = 8e-2


class Modes:
    """f(x) = sum_m c_m cos(2 pi m . x / N + a_m), periodic on the grid, and its derivatives."""

    def __init__(self, seed):
        rng = np.random.default_rng(seed)
        self.c = rng.standard_normal(len(MODES))
        self.a = rng.uniform(0, 2 * np.pi, len(MODES))
        self.k = [2 * np.pi * np.array(m, dtype=np.float64) / N for m in MODES]

    def value(self, pts):
        return sum(c * np.cos(pts @ k + a) for c, a, k in zip(self.c, self.a, self.k))

    def grad(self, pts):
        return sum((-c * np.sin(pts @ k + a))[..., None] * k for c, a, k in zip(self.c, self.a, self.k))


def nodes():
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return np.stack([x, y, z], -1).astype(np.float64)


def curl_shift(vec_pot, amp):
    """(3, N, N, N) float32 B = curl A, component a sampled at node + e_a / 2 (where the oracle reads B_a)."""
    out = np.zeros((3, N, N, N))
    for a in range(3):
        p = nodes()
        p[..., a] += 0.5
        g = [m.grad(p) for m in vec_pot]  # g[i][..., j] = d_j A_i
        b, c = (a + 1) % 3, (a + 2) % 3
        out[a] = g[c][..., b] - g[b][..., c]
    return (out * amp / np.sqrt(np.mean(out ** 2))).astype(np.float32)


def potential(modes, amp):
    f = modes.value(nodes())
    return (f * amp / np.sqrt(np.mean(f ** 2))).astype(np.float32)


def metric_along(oracle, phi, shift, xs):
    """h_mn, d_l h_mn, d_l d_p h_mn (index 0 = time; static) along points xs, from the interpolated grids."""
    vals, grads, hess = [], [], []
    for f, off in [(phi, None)] + [(shift[a], a) for a in range(3)]:
        p = xs.copy()
        if off is not None:
            p[:, off] -= 0.5
        it = oracle.tricubic(f, p)
        vals.append(it[:, 0])
        grads.append(it[:, 1:4])
        hess.append(it[:, 4:][:, HIDX])
    m = xs.shape[0]
    h, dh, ddh = np.zeros((m, 4, 4)), np.zeros((m, 4, 4, 4)), np.zeros((m, 4, 4, 4, 4))

    def add(i, j, coef, f):
        h[:, i, j] += coef * vals[f]
        dh[:, 1:, i, j] += coef * grads[f]
        ddh[:, 1:, 1:, i, j] += coef * hess[f]

    add(0, 0, -2.0, 0)
    for a in range(3):
        add(0, a + 1, -1.0, a + 1)
        add(a + 1, 0, -1.0, a + 1)
        add(a + 1, a + 1, -2.0, 0)
    return h, dh, ddh


def christoffel(dh):
    """Gamma^m_ab = 1/2 eta^mm (d_a h_mb + d_b h_ma - d_m h_ab), first order; dh[..., l, m, n] = d_l h_mn."""
    return 0.5 * ETA[:, None, None] * (np.einsum("...amb->...mab", dh) + np.einsum("...bma->...mab", dh) - dh)


def riemann(ddh):
    """R_abcd = 1/2 (d_b d_c h_ad + d_a d_d h_bc - d_b d_d h_ac - d_a d_c h_bd), first order (MTW signs)."""
    return 0.5 * (np.einsum("...bcad->...abcd", ddh) + np.einsum("...adbc->...abcd", ddh)
                  - np.einsum("...bdac->...abcd", ddh) - np.einsum("...acbd->...abcd", ddh))


def gr_ray(oracle, phi, shift, n, e1, e2):
    """First-order GR values of (deflection vector, dA - S, eps1, eps2, n.e1, n.e2) for one ray."""
    S = STEPS
    v = np.linspace(0.0, S, S * QPS + 1)
    dv = v[1] - v[0]
    w = np.full(v.size, dv)
    w[0] = w[-1] = 0.5 * dv

    def cumint(f):
        out = np.zeros_like(f)
        out[1:] = np.cumsum(0.5 * (f[1:] + f[:-1]) * dv, axis=0)
        return out

    h, dh, ddh = metric_along(oracle, phi, shift, OBS[None] - np.outer(v, n))
    gam, riem = christoffel(dh), riemann(ddh)
    k = np.concatenate([[-1.0], -n])  # zeroth-order k^mu, past-directed
    k_low = ETA * k
    # observer: k_i = -n exactly, k_0 from g^mn k_m k_n = 0
    dk_low0 = np.zeros(4)
    dk_low0[0] = (k @ h[0] @ k) / (2 * ETA[0] * k_low[0])
    dk0 = ETA * (dk_low0 - h[0] @ k)
    dk = dk0[None] + cumint(-np.einsum("vmab,a,b->vm", gam, k, k))
    dx = cumint(dk)
    dk_low = ETA[None] * dk + np.einsum("vmb,b->vm", h, k)
    v_minus_s = np.sum(w * (dk_low[:, 1:] @ n))  # v(S) - S, from dv/ds = 1 / |k_i|
    defl = (dx[-1, 1:] - n * (dx[-1, 1:] @ n)) / S
    screen = [np.concatenate([[0.0], e]) for e in (e1, e2)]
    tidal = np.empty((v.size, 2, 2))
    for a in range(2):
        for b in range(2):
            tidal[:, a, b] = np.einsum("vabcd,a,b,c,d->v", riem, screen[a], k, screen[b], k)
    A = -np.einsum("v,vab->ab", w * (S - v) * v, tidal) / S
    phi_end, phi_obs = -0.5 * h[-1, 0, 0], -0.5 * h[0, 0, 0]
    dA = (-0.5 * np.sum(w * (S - v) * v * (tidal[:, 0, 0] + tidal[:, 1, 1])) + v_minus_s
          + S * (phi_end - phi_obs))
    # n . e: minus the n-component of the rotation -Gamma^i_j0 e^j k^0 that the frame equation leaves out
    ne = [-np.sum(w * ((-np.einsum("vij,j->vi", gam[:, 1:, 1:, 0], e) * k[0]) @ n)) for e in (e1, e2)]
    return defl, dA, 2 * (A[0, 0] - A[1, 1]), 2 * (A[0, 1] + A[1, 0]), ne[0], ne[1]


COLUMNS = ["deflection1", "deflection2", "dA - n_steps", "ellipticity1", "ellipticity2", "n.e1", "n.e2"]


def oracle_columns(oracle, F, library=None):
    pix = np.arange(12 * NSIDE * NSIDE)
    maps, _, st = oracle.trace(F, NSIDE, pix, STEPS, observer=OBS, return_state=True, library=library)
    unit = lambda a: a / np.linalg.norm(a, axis=1, keepdims=True)  # noqa: E731
    nn, f1, f2 = unit(st[:, 3:6]), unit(st[:, 6:9]), unit(st[:, 9:12])
    return np.stack([maps[4], maps[5], st[:, 12] - STEPS, st[:, 16], st[:, 17],
                     (nn * f1).sum(1), (nn * f2).sum(1)], 1)


@pytest.fixture(scope="module")
def cases(oracle):
    """Fields and their GR predictions: a divergence-free shift (phi = 0) and a potential (B = 0)."""
    out = {}
    zero_phi = np.zeros((N, N, N), np.float32)
    zero_b = np.zeros((3, N, N, N), np.float32)
    for kind, phi, shift in [("curl", zero_phi, curl_shift([Modes(11), Modes(12), Modes(13)], 1e-7)),
                             ("phi", potential(Modes(21), 1e-5), zero_b)]:
        pix = np.arange(12 * NSIDE * NSIDE)
        z, s, ph = oracle.pix2zphi_ring(NSIDE, pix)
        r = np.stack([s * np.cos(ph), s * np.sin(ph), z], 1)
        th = np.stack([z * np.cos(ph), z * np.sin(ph), -s], 1)  # e1 at the observer, .cu:560
        p2 = np.stack([np.sin(ph), -np.cos(ph), 0 * ph], 1)     # e2
        want = []
        for i in range(pix.size):
            defl, dA, eps1, eps2, ne1, ne2 = gr_ray(oracle, phi, shift, -r[i], th[i], p2[i])
            want.append([defl @ th[i], defl @ p2[i], dA, eps1, eps2, ne1, ne2])
        out[kind] = (oracle.Fields(phi, shift), np.array(want))
    return out


def errors(got, want):
    """max |got - want| / rms(want) per column (nan where GR predicts zero)."""
    rms = np.sqrt(np.mean(want ** 2, axis=0))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.abs(got - want).max(axis=0) / np.where(rms > 0, rms, np.nan)


def check_errors(oracle, cases, library=None):
    """{kind: per-column errors} for the default oracle build or a variant."""
    return {kind: errors(oracle_columns(oracle, F, library), want) for kind, (F, want) in cases.items()}


def report(errs):
    for kind, e in errs.items():
        print(kind, "  ".join(f"{c} {x:.2e}" for c, x in zip(COLUMNS, e)))


def test_curl_shift_deflection_follows_the_geodesic(oracle, cases):
    F, want = cases["curl"]
    e = errors(oracle_columns(oracle, F), want)
    report({"curl": e})
    assert np.sqrt(np.mean(want[:, :2] ** 2)) > 1e-8  # the gravitomagnetic deflection is not trivially small
    assert e[0] < BAR_DEFLECTION and e[1] < BAR_DEFLECTION


def test_curl_shift_dA_follows_sachs_focusing(oracle, cases):
    F, want = cases["curl"]
    e = errors(oracle_columns(oracle, F), want)
    assert e[2] < BAR_DA


def test_curl_shift_ellipticity_follows_the_tidal_matrix(oracle, cases):
    F, want = cases["curl"]
    e = errors(oracle_columns(oracle, F), want)
    assert e[3] < BAR_EPS and e[4] < BAR_EPS


def test_curl_shift_frame_drift_is_the_left_out_precession(oracle, cases):
    F, want = cases["curl"]
    e = errors(oracle_columns(oracle, F), want)
    assert np.sqrt(np.mean(want[:, 5:] ** 2)) > 1e-8  # the drift is visible
    assert e[5] < BAR_FRAME and e[6] < BAR_FRAME


def test_potential_maps_follow_gr(oracle, cases):
    F, want = cases["phi"]
    got = oracle_columns(oracle, F)
    e = errors(got, want)
    report({"phi": e})
    assert e[0] < BAR_DEFLECTION and e[1] < BAR_DEFLECTION
    assert e[2] < BAR_DA
    assert e[3] < BAR_EPS and e[4] < BAR_EPS
    assert np.abs(got[:, 5:]).max() < 1e-12  # a scalar potential keeps the frame orthogonal


# one plausible mistake each in the terms these checks pin (.cu:370-509)
MUTANTS = {
    "L halved": ("L = I(S(L) + n[a] * (H[0][0] + H[1][1] + H[2][2]));",
                 "L = I(S(L) + S(0.5) * n[a] * (H[0][0] + H[1][1] + H[2][2]));"),
    "L sign": ("- PH * D + S(0.5) * S(L) * D;", "- PH * D - S(0.5) * S(L) * D;"),
    "N transposed": ("for (int j = 0; j < 3; ++j) N[j] = I(S(N[j]) + g[j] * n[a]);", "N[a] = I(S(N[a]) + ng);"),
    "E1 first term only": ("E1 = I(S(E1) + S(0.5) * e1[a] * ng + S(0.5) * n[a] * e1g);",
                           "E1 = I(S(E1) + e1[a] * ng);"),
    "E1 second term only": ("E1 = I(S(E1) + S(0.5) * e1[a] * ng + S(0.5) * n[a] * e1g);",
                            "E1 = I(S(E1) + n[a] * e1g);"),
    "T dropped": ("Hm[i][j] = S(I(H[i][j] + S(0.5) * S(T[hidx(i, j)])));", "Hm[i][j] = S(I(H[i][j]));"),
    "T doubled": ("Hm[i][j] = S(I(H[i][j] + S(0.5) * S(T[hidx(i, j)])));",
                  "Hm[i][j] = S(I(H[i][j] + S(T[hidx(i, j)])));"),
    "n.N rescaling dropped": ("S(0.25) * (n[0] * S(N[0])", "S(0) * (n[0] * S(N[0])"),
    "shear coupling doubled": ("dy[14] = S(4) * s1 * gam + S1H * D * D;",
                               "dy[14] = S(4) * s1 * gam + S(2) * S1H * D * D;"),
    "P:H sign": ("dy[13] = S(-4) * Dp * gam - PH * D", "dy[13] = S(-4) * Dp * gam + PH * D"),
}
BARS = np.array([BAR_DEFLECTION, BAR_DEFLECTION, BAR_DA, BAR_EPS, BAR_EPS, BAR_FRAME, BAR_FRAME])


def test_planted_mistakes_fail(oracle, cases, tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src_dir = os.path.dirname(oracle.__file__)
    src = open(os.path.join(src_dir, "srt_oracle.hpp")).read()
    good = check_errors(oracle, cases)
    report(good)
    for e in good.values():
        assert np.all(np.nan_to_num(e) < BARS)
    for name, (before, after) in MUTANTS.items():
        assert src.count(before) == 1, f"mutation site of '{name}' not found once in srt_oracle.hpp"
        d = tmp_path / name.replace(" ", "_").replace(".", "").replace(":", "")
        d.mkdir()
        (d / "srt_oracle.hpp").write_text(src.replace(before, after))
        shutil.copy(os.path.join(src_dir, "oracle_capi.cpp"), d)
        so = d / "liboracle_mutant.so"
        subprocess.run(["g++", "-O2", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o",
                        str(so), str(d / "oracle_capi.cpp")], check=True)
        errs = check_errors(oracle, cases, library=oracle.bind(str(so)))
        worst = max(np.nanmax(e / BARS) for e in errs.values())
        print(f"{name:24s} worst check / bar = {worst:.1f}")
        assert worst > 10, name
