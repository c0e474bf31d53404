"""Pins of the oracle's Catmull-Rom / tricubic interpolation (snapshot-raytracer.cu:195-312).

* brute force against an independently written Keys (a = -1/2) cubic
  convolution kernel W(s) and its analytic derivatives, summed over the
  4 x 4 x 4 taps with periodic wrap;
* exact reproduction of quadratic polynomials (value, gradient, Hessian),
  and the failure to reproduce cubics (Catmull-Rom is exact to degree 2 only);
* periodicity under integer shifts, across the wrap;
* binary16 grids read exactly as their decoded values / 2^27;
* the REF form equals the FP64 form evaluated at the float-rounded position.
"""
import numpy as np
import pytest

A = -0.5


def _inner(a):  # |s| in [0, 1]
    return (A + 2) * a ** 3 - (A + 3) * a ** 2 + 1, 3 * (A + 2) * a ** 2 - 2 * (A + 3) * a, 6 * (A + 2) * a - 2 * (A + 3)


def _outer(a):  # |s| in [1, 2]
    return (A * a ** 3 - 5 * A * a ** 2 + 8 * A * a - 4 * A, 3 * A * a ** 2 - 10 * A * a + 8 * A, 6 * A * a - 10 * A)


def keys_taps(s):
    """Kernel W(s), W'(s), W''(s) at the 4 signed tap distances s = x - x_i (taps -1, 0, 1, 2).

    The kernel is piecewise cubic; on a cell the two inner taps use the |s| <= 1
    piece and the two outer taps the 1 <= |s| <= 2 piece (this fixes the side of
    the second-derivative jump when x sits exactly on a node).
    """
    w = np.empty((3, 4))
    for m in range(4):
        a = abs(s[m])
        sg = 1.0 if s[m] >= 0 else -1.0
        v, d1, d2 = _inner(a) if m in (1, 2) else _outer(a)
        w[0, m], w[1, m], w[2, m] = v, sg * d1, d2
    return w


def brute_tricubic(grid, p):
    """value, grad, Hessian (xx, xy, xz, yy, yz, zz) by explicit tap sums."""
    nz, ny, nx = grid.shape
    base = np.floor(p).astype(np.int64)
    offs = np.arange(-1, 3)
    taps = [base[d] + offs for d in range(3)]
    s = [p[d] - taps[d] for d in range(3)]  # signed distances x - x_i, in (-2, 2]
    W = [keys_taps(s[d]) for d in range(3)]
    sub = grid[np.ix_(taps[2] % nz, taps[1] % ny, taps[0] % nx)].astype(np.float64)  # [z, y, x]

    def contract(ox, oy, oz):
        return np.einsum("kji,i,j,k->", sub, W[0][ox], W[1][oy], W[2][oz])

    v = contract(0, 0, 0)
    g = [contract(1, 0, 0), contract(0, 1, 0), contract(0, 0, 1)]
    h = [contract(2, 0, 0), contract(1, 1, 0), contract(1, 0, 1), contract(0, 2, 0), contract(0, 1, 1),
         contract(0, 0, 2)]
    return np.array([v] + g + h)


@pytest.fixture(scope="module")
def rgrid():
    rng = np.random.default_rng(5)
    return rng.standard_normal((12, 10, 9)).astype(np.float32)


def test_brute_force_keys_kernel(oracle, rgrid):
    rng = np.random.default_rng(6)
    nz, ny, nx = rgrid.shape
    pos = rng.uniform(-2 * nx, 3 * nx, size=(200, 3)) * np.array([1, ny / nx, nz / nx])
    pos[:10] = np.round(pos[:10])  # exactly on nodes
    got = oracle.tricubic(rgrid, pos)
    ref = np.array([brute_tricubic(rgrid, p) for p in pos])
    scale = np.abs(ref).max(axis=0)
    assert np.abs(got - ref).max(axis=0).max() < 1e-12 * scale.max()
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12)


def poly_grid(n, coef, shape=None):
    """Grid of a quadratic polynomial f(x,y,z) = sum c_ijk x^i y^j z^k (i+j+k <= 2 or general)."""
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = np.zeros((n, n, n))
    for (i, j, k), c in coef.items():
        f += c * x ** i * y ** j * z ** k
    return f


def poly_eval(coef, p):
    x, y, z = p
    v = sum(c * x ** i * y ** j * z ** k for (i, j, k), c in coef.items())

    def d(e, a, q):
        return e * a ** (e - 1) if e >= 1 else 0.0

    gx = sum(c * (i * x ** (i - 1) if i else 0) * y ** j * z ** k for (i, j, k), c in coef.items())
    gy = sum(c * x ** i * (j * y ** (j - 1) if j else 0) * z ** k for (i, j, k), c in coef.items())
    gz = sum(c * x ** i * y ** j * (k * z ** (k - 1) if k else 0) for (i, j, k), c in coef.items())

    def dd(e, a):
        return e * (e - 1) * a ** (e - 2) if e >= 2 else 0.0

    def d1(e, a):
        return e * a ** (e - 1) if e >= 1 else 0.0

    hxx = sum(c * dd(i, x) * y ** j * z ** k for (i, j, k), c in coef.items())
    hyy = sum(c * x ** i * dd(j, y) * z ** k for (i, j, k), c in coef.items())
    hzz = sum(c * x ** i * y ** j * dd(k, z) for (i, j, k), c in coef.items())
    hxy = sum(c * d1(i, x) * d1(j, y) * z ** k for (i, j, k), c in coef.items())
    hxz = sum(c * d1(i, x) * y ** j * d1(k, z) for (i, j, k), c in coef.items())
    hyz = sum(c * x ** i * d1(j, y) * d1(k, z) for (i, j, k), c in coef.items())
    return np.array([v, gx, gy, gz, hxx, hxy, hxz, hyy, hyz, hzz])


QUAD = {(0, 0, 0): 0.3, (1, 0, 0): -0.02, (0, 1, 0): 0.05, (0, 0, 1): 0.01, (2, 0, 0): 1e-3, (0, 2, 0): -2e-3,
        (0, 0, 2): 5e-4, (1, 1, 0): 7e-4, (1, 0, 1): -3e-4, (0, 1, 1): 9e-4, (1, 1, 1): 2e-5, (2, 1, 0): 1e-5,
        (2, 2, 2): 1e-9}


def test_quadratic_products_reproduced_exactly(oracle):
    n = 24
    grid = poly_grid(n, QUAD).astype(np.float64)
    rng = np.random.default_rng(1)
    pos = rng.uniform(2.0, n - 3.0, size=(100, 3))
    # float32 grid storage rounds the samples; compare against the polynomial of
    # the stored values by using a grid whose values are exactly representable
    g32 = grid.astype(np.float32)
    got = oracle.tricubic(g32, pos)
    exact = np.array([poly_eval(QUAD, p) for p in pos])
    err = np.abs(got - exact).max(axis=0)
    # the float32 rounding of the samples (<= 6e-8 relative of |f| <= ~2) bounds the error
    assert err.max() < 5e-7, err


def test_cubic_not_reproduced(oracle):
    n = 24
    cub = {(3, 0, 0): 1e-4}
    grid = poly_grid(n, cub).astype(np.float32)
    p = np.array([[10.37, 8.0, 8.0]])
    got = oracle.tricubic(grid, p)[0]
    exact = poly_eval(cub, p[0])
    assert abs(got[0] - exact[0]) > 1e-7  # Catmull-Rom is exact to degree 2 only


def test_constant_grid(oracle):
    grid = np.full((8, 8, 8), 2.5, np.float32)
    rng = np.random.default_rng(2)
    out = oracle.tricubic(grid, rng.uniform(-20, 20, (50, 3)))
    assert np.all(out[:, 0] == 2.5)
    assert np.all(out[:, 1:] == 0.0)


def test_periodic_integer_shift(oracle, rgrid):
    nz, ny, nx = rgrid.shape
    rng = np.random.default_rng(3)
    pos = np.round(rng.uniform(0, 9, (40, 3)) * 1024) / 1024  # few mantissa bits: shifts are exact
    a = oracle.tricubic(rgrid, pos)
    for sx, sy, sz in [(1, 0, 0), (0, 3, 0), (0, 0, 5), (-2, 7, -11)]:
        rolled = np.roll(rgrid, shift=(sz, sy, sx), axis=(0, 1, 2))
        b = oracle.tricubic(rolled, pos + np.array([sx, sy, sz]))
        assert np.array_equal(a, b)
        c = oracle.tricubic(rgrid, pos + np.array([sx * nx, sy * ny, sz * nz]))
        assert np.array_equal(a, c)


def test_f16_grid_reads_decoded_values(oracle, rgrid):
    b = (rgrid * 1e-7).astype(np.float32)
    b16 = oracle.shift_to_f16(b)
    decoded = oracle.f16_to_f32(b16) / np.float32(2 ** 27)
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 9, (30, 3))
    np.testing.assert_array_equal(oracle.tricubic(None, pos, grid16=b16), oracle.tricubic(decoded, pos))


def test_ref_form_offsets(oracle, rgrid):
    rng = np.random.default_rng(7)
    pos = rng.uniform(200, 300, (50, 3))
    ref = oracle.tricubic(rgrid, pos, form=oracle.FORM_REF)
    exact = oracle.tricubic(rgrid, pos, form=oracle.FORM_FP64)
    at_f32 = oracle.tricubic(rgrid, pos.astype(np.float32).astype(np.float64), form=oracle.FORM_FP64)
    scale = np.abs(exact).max(axis=0)
    # the REF form sees the float-rounded position (offset error up to ~1e-5 cell here) ...
    assert (np.abs(ref - at_f32) / scale).max() < 2e-6
    # ... which differs from the exact FP64 position measurably
    assert (np.abs(ref - exact) / scale).max() > 1e-7


def test_weights_form_equals_restatement(oracle, fields):
    """The work-efficient weights form (weight vectors per axis, one x -> y -> z contraction) whose
    operations give the roofline's algorithmic flop count evaluates the same interpolant as the
    restatement of .cu:215-312, including across the periodic wrap."""
    phi, _ = fields.snapshot(32, 21)
    g = phi.numpy()
    pos = np.random.default_rng(22).uniform(-40, 75, (3000, 3))
    pos[:10] = np.floor(pos[:10])  # exact nodes
    a = oracle.tricubic(g, pos)
    b = oracle.tricubic_weights(g, pos)
    scale = np.abs(a).max(axis=0)
    assert (np.abs(a - b) / scale).max() < 1e-13


def test_algorithmic_flop_constant(oracle, fields):
    """The roofline's algorithmic flops per ray-step (bench.ALGO_FLOPS_PER_RAY_STEP) is the operation
    count of the work-efficient form with shared weight sets; the reference's expanded-polynomial form
    counts about 2.4x as many."""
    import bench
    phi, shift = fields.snapshot(64, 1)
    F = oracle.Fields(phi.numpy(), shift.numpy())
    counts = [oracle.count_ops(F, 16, 1000, 32, variant=v)["total"] for v in (0, 1, 2)]
    print("flops per ray-step (reference form, weights form, shared weight sets):", counts)
    assert counts[2] == bench.ALGO_FLOPS_PER_RAY_STEP
    assert counts[0] == bench.REF_FORM_FLOPS_PER_RAY_STEP
    assert counts[2] < counts[1] < counts[0]
