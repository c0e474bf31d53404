"""Pins of the oracle's right-hand side, RK4 integrator and ray maps
(snapshot-raytracer.cu:315-583).

Right-hand side (.cu:315-510): fields that tricubic Catmull-Rom reproduces
exactly (quadratic polynomials near the grid centre) give closed-form values,
gradients and Hessians, so the 19 derivatives can be computed independently in
tensor form (numpy einsum) from the equations of the geodesic and Jacobi map:
    x' = -n (1 + 2 phi) - B,  n' = -P (2 grad phi + N),  e' = -n (2 e.grad phi + E),
    D'' = -4 D' G - (P:H) D + L D / 2 [+ (s1^2 + s2^2) / D^3 if D > 1/2],
    s' = 4 s G + (S:(H + T/2)) D^2, eps' = -2 s sqrt(4 + |eps|^2) / D^2,
    omega' = -(s1 eps2 - s2 eps1) / (D^2 (2 + sqrt(4 + |eps|^2))),
with P = n n^T - 1, S1 = e1 e1^T - e2 e2^T, S2 = e1 e2^T + e2 e1^T,
G = n.grad phi + n.N / 4, N_j = sum_a n_a dB_a/dx_j,
E = 1/2 sum_a [e_a (n . grad B_a) + n_a (e . grad B_a)],
T_ij = sum_a n_a H^a_ij - 1/2 [(n H^i)_j + (n H^j)_i], L = sum_a n_a tr H^a,
and B_a sampled at x - e_a / 2.  Non-symmetric shift gradients make any
transposed index visible.

Integrator and maps (.cu:513-583): zero field (dA = n_steps exactly), constant
potential (path length (1 + 2 phi0) n_steps), constant shift (closed-form
deflection), amplitude scaling (deflection linear, rotation quadratic), a
90-degree rotation of the whole configuration about the z axis (maps move a
quarter ring), and the first-order (Born) line-of-sight integrals for a single
Fourier mode.
"""
import numpy as np
import pytest

N = 32
C = N / 2.0


def quad_grid(c0, g, H, center=C, n=N):
    """Samples of c0 + g.(p - center) + 1/2 (p - center)^T H (p - center) on the grid nodes, (z, y, x)."""
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = np.stack([x, y, z]).astype(np.float64) - center
    f = c0 + np.einsum("i,ixyz->xyz", g, d) + 0.5 * np.einsum("ixyz,ij,jxyz->xyz", d, H, d)
    return f


def quad_eval(c0, g, H, p, center=C):
    d = p - center
    return c0 + g @ d + 0.5 * d @ H @ d, g + H @ d, H


def sym(rng, scale):
    a = rng.standard_normal((3, 3)) * scale
    return 0.5 * (a + a.T)


def orthonormal_frame(rng):
    n = rng.standard_normal(3)
    n /= np.linalg.norm(n)
    e1 = np.cross(n, rng.standard_normal(3))
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    return n, e1, e2


def expected_rhs(y, phi_q, b_q):
    """Independent tensor-form evaluation of the 19 derivatives."""
    x = y[0:3]
    n = y[3:6] / np.linalg.norm(y[3:6])
    e1 = y[6:9] / np.linalg.norm(y[6:9])
    e2 = y[9:12] / np.linalg.norm(y[9:12])
    D, Dp, s1, s2, eps1, eps2 = y[12:18]
    phi, g, H = quad_eval(*phi_q, x)
    B = np.empty(3)
    gB = np.empty((3, 3))  # gB[a, j] = dB_a/dx_j
    HB = np.empty((3, 3, 3))
    for a in range(3):
        pa = x.copy()
        pa[a] -= 0.5
        B[a], gB[a], HB[a] = quad_eval(*b_q[a], pa)
    Nv = np.einsum("a,aj->j", n, gB)
    E1 = 0.5 * (np.einsum("a,j,aj->", e1, n, gB) + np.einsum("a,j,aj->", n, e1, gB))
    E2 = 0.5 * (np.einsum("a,j,aj->", e2, n, gB) + np.einsum("a,j,aj->", n, e2, gB))
    nH = np.einsum("k,ikj->ij", n, HB)  # nH[i, j] = (n H^i)_j
    T = np.einsum("a,aij->ij", n, HB) - 0.5 * (nH + nH.T)
    L = np.einsum("a,aii->", n, HB)
    P = np.outer(n, n) - np.eye(3)
    S1 = np.outer(e1, e1) - np.outer(e2, e2)
    S2 = np.outer(e1, e2) + np.outer(e2, e1)
    G = n @ g + 0.25 * n @ Nv
    d = np.zeros(19)
    d[0:3] = -n * (1 + 2 * phi) - B
    d[3:6] = -P @ (2 * g + Nv)
    d[6:9] = -n * (2 * (e1 @ g) + E1)
    d[9:12] = -n * (2 * (e2 @ g) + E2)
    d[12] = -Dp
    d[13] = -4 * Dp * G - np.sum(P * H) * D + 0.5 * L * D
    Hm = H + 0.5 * T
    d[14] = 4 * s1 * G + np.sum(S1 * Hm) * D * D
    d[15] = 4 * s2 * G + np.sum(S2 * Hm) * D * D
    if D > 0.5:
        q = np.sqrt(4 + eps1 ** 2 + eps2 ** 2)
        d[13] += (s1 ** 2 + s2 ** 2) / D ** 3
        d[16] = -2 * s1 * q / D ** 2
        d[17] = -2 * s2 * q / D ** 2
        d[18] = -(s1 * eps2 - s2 * eps1) / (D ** 2 * (2 + q))
    return d


@pytest.fixture(scope="module")
def quad_fields(oracle):
    rng = np.random.default_rng(42)
    phi_q = (3e-5, rng.standard_normal(3) * 2e-6, sym(rng, 3e-7))
    b_q = []
    for a in range(3):
        # independent, non-symmetric gradients and Hessians per component
        b_q.append((rng.standard_normal() * 1e-6, rng.standard_normal(3) * 5e-7, sym(rng, 1e-7)))
    phi = quad_grid(*phi_q).astype(np.float32)
    shift = np.stack([quad_grid(*b_q[a]).astype(np.float32) for a in range(3)])
    # compare against the polynomials of the *stored* float32 samples: refit is not
    # needed because the float32 rounding is ~1e-7 relative and checked below
    return oracle.Fields(phi, shift), phi_q, b_q


def random_states(rng, count, D_range=(0.6, 20.0)):
    ys = []
    for _ in range(count):
        n, e1, e2 = orthonormal_frame(rng)
        y = np.zeros(19)
        y[0:3] = C + rng.uniform(-3, 3, 3)
        y[3:6] = n * rng.uniform(0.9, 1.1)  # unnormalised on purpose: the RHS normalises copies
        y[6:9] = e1 * rng.uniform(0.9, 1.1)
        y[9:12] = e2
        y[12] = rng.uniform(*D_range)
        y[13] = -1 + rng.standard_normal() * 1e-3
        y[14:18] = rng.standard_normal(4) * 1e-3
        y[18] = rng.standard_normal() * 1e-6
        ys.append(y)
    return np.array(ys)


def test_rhs_matches_tensor_form(oracle, quad_fields):
    F, phi_q, b_q = quad_fields
    rng = np.random.default_rng(7)
    ys = random_states(rng, 40)
    got = oracle.rhs(F, ys, form=oracle.FORM_FP64)
    exp = np.array([expected_rhs(y, phi_q, b_q) for y in ys])
    scale = np.abs(exp).max(axis=0)
    scale[scale == 0] = 1
    err = (np.abs(got - exp) / scale).max(axis=0)
    # float32 samples of the polynomials limit agreement to ~1e-6 of each component's scale
    assert err.max() < 2e-5, err


def test_rhs_shift_terms_are_visible(oracle, quad_fields):
    """The comparison above would not notice a wrong shift term if its effect were tiny."""
    F, phi_q, b_q = quad_fields
    rng = np.random.default_rng(8)
    ys = random_states(rng, 20)
    got = oracle.rhs(F, ys)
    for a in range(3):  # transpose one component's gradient: the derivatives must move
        bad = list(b_q)
        c0, g, H = b_q[a]
        bad[a] = (c0, g[::-1].copy(), H)
        alt = np.array([expected_rhs(y, phi_q, bad) for y in ys])
        diff = np.abs(alt - got).max(axis=0) / np.abs(got).max(axis=0).clip(1e-300)
        assert diff[3:6].max() > 1e-3


def test_rhs_branch_and_jacobi_algebra(oracle):
    z = np.zeros((8, 8, 8), np.float32)
    F = oracle.Fields(z, np.zeros((3, 8, 8, 8), np.float32))
    rng = np.random.default_rng(9)
    for D in (0.3, 0.5, 0.7, 5.0):
        ys = random_states(rng, 5, D_range=(D, D))
        got = oracle.rhs(F, ys)
        for y, d in zip(ys, got):
            n = y[3:6] / np.linalg.norm(y[3:6])
            np.testing.assert_allclose(d[0:3], -n, rtol=0, atol=1e-15)
            assert np.all(d[3:12] == 0)
            assert d[12] == -y[13]
            s1, s2, e1, e2 = y[14:18]
            if D > 0.5:
                q = np.sqrt(4 + e1 * e1 + e2 * e2)
                np.testing.assert_allclose(d[13], (s1 ** 2 + s2 ** 2) / D ** 3, rtol=1e-13)
                np.testing.assert_allclose(d[16:18], -2 * np.array([s1, s2]) * q / D ** 2, rtol=1e-13)
                np.testing.assert_allclose(d[18], -(s1 * e2 - s2 * e1) / (D ** 2 * (2 + q)), rtol=1e-12)
            else:
                assert d[13] == 0 and np.all(d[16:19] == 0)
            assert np.all(d[14:16] == 0)


# ---------------------------------------------------------------- traces ----
def test_zero_field_trace(oracle):
    z = np.zeros((16, 16, 16), np.float32)
    F = oracle.Fields(z, np.zeros((3, 16, 16, 16), np.float32))
    maps, _, st = oracle.trace(F, 4, np.arange(192), 12, return_state=True)
    assert np.all(maps[0] == 12.0)
    assert np.all(maps[1:4] == 0.0)
    assert np.abs(maps[4:]).max() < 1e-14
    zz, s, ph = oracle.pix2zphi_ring(4, np.arange(192))
    r = np.stack([s * np.cos(ph), s * np.sin(ph), zz], axis=1)
    np.testing.assert_allclose(st[:, 0:3] - 8.0, 12 * r, atol=1e-12)


def test_constant_potential_and_shift(oracle):
    n = 16
    phi0 = 2e-3
    B = np.array([3e-4, -2e-4, 1e-4])
    F = oracle.Fields(np.full((n, n, n), phi0, np.float32),
                      np.stack([np.full((n, n, n), b, np.float32) for b in B]))
    steps = 10
    pix = np.arange(192)
    maps, _, st = oracle.trace(F, 4, pix, steps, return_state=True)
    phi0f = float(np.float32(phi0))
    Bf = B.astype(np.float32).astype(np.float64)
    assert np.all(maps[0] == steps)  # no gradients: D' stays -1
    zz, s, ph = oracle.pix2zphi_ring(4, pix)
    r = np.stack([s * np.cos(ph), s * np.sin(ph), zz], axis=1)
    disp = steps * ((1 + 2 * phi0f) * r - Bf)  # x' = -n (1 + 2 phi0) - B with n = -r
    np.testing.assert_allclose(st[:, 0:3] - n / 2, disp, atol=1e-11)
    u = disp / np.linalg.norm(disp, axis=1, keepdims=True)
    dlt = u - r
    th = np.stack([zz * np.cos(ph), zz * np.sin(ph), -s], axis=1)
    ph2 = np.stack([np.sin(ph), -np.cos(ph), 0 * ph], axis=1)
    np.testing.assert_allclose(maps[4], np.sum(dlt * th, axis=1), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(maps[5], np.sum(dlt * ph2, axis=1), rtol=1e-6, atol=1e-12)


@pytest.fixture(scope="module")
def field64(fields):
    phi, shift = fields.snapshot(64, 1)
    return phi.numpy(), shift.numpy()


def test_amplitude_scaling(oracle, field64):
    phi, shift = field64
    pix = np.arange(0, 3072, 7)
    m1, _ = oracle.trace(oracle.Fields(phi, shift), 16, pix, 24)
    m2, _ = oracle.trace(oracle.Fields(2 * phi, 2 * shift), 16, pix, 24)
    for m in (1, 2, 4, 5):  # first order in the fields
        ratio = np.linalg.norm(m2[m]) / np.linalg.norm(m1[m])
        assert abs(ratio - 2) < 2e-3, (m, ratio)
    ratio = np.linalg.norm(m2[0] - 24) / np.linalg.norm(m1[0] - 24)
    assert abs(ratio - 2) < 2e-3
    ratio = np.linalg.norm(m2[3]) / np.linalg.norm(m1[3])  # rotation is second order
    assert abs(ratio - 4) < 2e-2, ratio


def rotate_config(phi, shift, n):
    """90-degree rotation about the z axis through (n/2 + 1/2, n/2 + 1/2):
    (x, y) -> (n + 1 - y, x), fields and the staggered shift components rotated."""
    phi_r = np.empty_like(phi)
    bx_r = np.empty_like(phi)
    by_r = np.empty_like(phi)
    bz_r = np.empty_like(phi)
    for ii in range(n):
        # phi' (i, j) = phi(j, n + 1 - i);  Bz' likewise
        phi_r[:, :, ii] = phi[:, (n + 1 - ii) % n, :]
        bz_r[:, :, ii] = shift[2][:, (n + 1 - ii) % n, :]
        # Bx' node (i, j) at (i + 1/2, j): -By at (j, n + 1/2 - i), node (j, n - i)
        bx_r[:, :, ii] = -shift[1][:, (n - ii) % n, :]
        # By' node (i, j) at (i, j + 1/2): Bx at (j + 1/2, n + 1 - i), node (j, n + 1 - i)
        by_r[:, :, ii] = shift[0][:, (n + 1 - ii) % n, :]
    return phi_r, np.stack([bx_r, by_r, bz_r])


def quarter_ring_shift(nside):
    """perm[p] = RING pixel at the same ring with phi + pi/2."""
    npix = 12 * nside * nside
    perm = np.empty(npix, np.int64)
    start = 0
    for i in range(1, 4 * nside):
        L = 4 * i if i < nside else (4 * nside if i <= 3 * nside else 4 * (4 * nside - i))
        idx = np.arange(L)
        perm[start + idx] = start + (idx + L // 4) % L
        start += L
    return perm


def test_rotation_covariance(oracle, field64):
    phi, shift = field64
    n = 64
    phi_r, shift_r = rotate_config(phi, shift, n)
    # sanity of the field rotation: the node (i, j) of the rotated potential holds the
    # old value at (j, n + 1 - i)
    assert phi_r[3, 5, 7] == phi[3, (n + 1 - 7) % n, 5]
    obs = (n / 2 + 0.5, n / 2 + 0.5, n / 2 + 0.25)  # on the rotation axis
    nside, steps = 8, 20
    pix = np.arange(12 * nside * nside)
    m, _ = oracle.trace(oracle.Fields(phi, shift), nside, pix, steps, observer=obs)
    mr, _ = oracle.trace(oracle.Fields(phi_r, shift_r), nside, pix, steps, observer=obs)
    perm = quarter_ring_shift(nside)
    for k in range(6):
        ref = m[k].astype(np.float64)
        scale = np.sqrt(np.mean((ref - ref.mean()) ** 2)) if k == 0 else np.sqrt(np.mean(ref ** 2))
        d = np.abs(mr[k][perm] - ref).max() / scale
        assert d < 1e-4, (k, d)  # float32 map outputs; the maps themselves are O(1) of scale


def test_born_single_mode(oracle):
    """First-order line-of-sight integrals for phi = A cos(k.x + a), B = 0."""
    n, steps, nside = 64, 24, 4
    # a long wave (not periodic on the grid; the rays and their stencils stay within
    # [6, 58]) keeps the RK4 / piecewise-cubic discretisation error near 1%
    kv = 2 * np.pi * np.array([1, 2, 1]) / (4 * n)
    A, alpha = 1e-6, 0.3
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    phi = (A * np.cos(kv[0] * x + kv[1] * y + kv[2] * z + alpha)).astype(np.float32)
    F = oracle.Fields(phi, np.zeros((3, n, n, n), np.float32))
    pix = np.arange(12 * nside * nside)
    maps, _, state = oracle.trace(F, nside, pix, steps, return_state=True)
    zz, s, ph = oracle.pix2zphi_ring(nside, pix)
    r = np.stack([s * np.cos(ph), s * np.sin(ph), zz], axis=1)
    th = np.stack([zz * np.cos(ph), zz * np.sin(ph), -s], axis=1)
    p2 = np.stack([np.sin(ph), -np.cos(ph), 0 * ph], axis=1)
    lam = np.linspace(0, steps, 4001)
    x0 = np.array([n / 2] * 3)
    w = np.full(lam.size, 1.0)
    w[0] = w[-1] = 0.5
    w *= lam[1] - lam[0]
    hidx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    got, want = [], []
    for p in range(pix.size):
        xs = x0 + np.outer(lam, r[p])
        # gradient and Hessian of the interpolated field along the unperturbed ray (the
        # interpolation itself is pinned separately against the Keys kernel)
        it = oracle.tricubic(phi, xs)
        grad = it[:, 1:4]
        hess = it[:, 4:][:, hidx]
        nvec = -r[p]
        gperp = grad - np.outer(grad @ nvec, nvec)
        # transverse displacement: dx_perp(N) = -int (N - l) 2 grad_perp phi dl
        dx = -np.sum(w[:, None] * (steps - lam)[:, None] * 2 * gperp, axis=0)
        defl = dx / steps
        # dA: D(N) = N - int (N - l) [4 n.grad phi - (P:H) l] dl, P:H = n^T H n - tr H
        PH = np.einsum("i,lij,j->l", nvec, hess, nvec) - np.trace(hess, axis1=1, axis2=2)
        dD = -np.sum(w * (steps - lam) * (4 * (grad @ nvec) - PH * lam))
        got.append((maps[4][p], maps[5][p], state[p, 12] - steps))
        want.append((defl @ th[p], defl @ p2[p], dD))
    got, want = np.array(got), np.array(want)
    err = np.abs(got - want).max(axis=0) / np.abs(want).max(axis=0)
    assert err.max() < 2e-2, err
