"""Host checks of the bicubic cell-record arithmetic the mixed kernel uses
(record construction and Horner evaluation), in FP64, against the oracle's
Catmull-Rom tricubic (snapshot-raytracer.cu:195-312).

The records hold, per cell column and plane, the monomial coefficients
c = M^T V M of the Catmull-Rom bicubic (M = Catmull-Rom weights by power of
t); the kernel evaluates value, gradient and Hessian from four planes of
records by Horner's rule in x and y and a Catmull-Rom pass in z.  Written
here from that description (independently of the device code) and compared
with the oracle's tap-weight evaluation: agreement to rounding means the
monomial matrix, the derivative factors (2t, 3t^2 for d/dt, 3t for half the
second derivative) and the plane/column indexing are right.  A wrong entry
of M, a dropped factor 2 on the xx/yy terms or a swapped x/y order fails.
"""
import numpy as np
import pytest

# w_a(t) = sum_p M[a, p] t^p: the Catmull-Rom (Keys a = -1/2) weights of taps -1, 0, 1, 2
M = np.array([[0.0, -0.5, 1.0, -0.5],
              [1.0, 0.0, -2.5, 1.5],
              [0.0, 0.5, 2.0, -1.5],
              [0.0, 0.0, -0.5, 0.5]])


def cr_weights(t):
    """(w, w', w'') of the four taps by direct differentiation of the polynomials above."""
    p = np.array([1.0, t, t * t, t ** 3])
    dp = np.array([0.0, 1.0, 2 * t, 3 * t * t])
    ddp = np.array([0.0, 0.0, 2.0, 6 * t])
    return M @ p, M @ dp, M @ ddp


def record(grid, i, j, kp):
    """16 coefficients c[p][q] of cell column (i, j) on grid plane kp (periodic), float64."""
    n = grid.shape[0]
    V = np.empty((4, 4))  # V[b, a]: tap (i - 1 + a, j - 1 + b)
    for b in range(4):
        for a in range(4):
            V[b, a] = grid[kp % n, (j - 1 + b) % n, (i - 1 + a) % n]
    X = V @ M           # X[b, p] = sum_a M[a, p] V[b, a]
    return (X.T @ M)    # c[p, q] = sum_b M[b, q] X[b, p]


def eval_column(grid, pos):
    """value, gradient, Hessian (xx, xy, xz, yy, yz, zz) from the records of the stencil's four planes."""
    n = grid.shape[0]
    f = np.floor(pos)
    tx, ty, tz = pos - f
    i, j, k = (int(v) % n for v in f)
    P00, P10, P20h, P01, P11, P02h = (np.empty(4) for _ in range(6))
    for r in range(4):
        c = record(grid, i, j, k - 1 + r)
        hv = np.empty(4)
        hd = np.empty(4)
        hh = np.empty(4)
        for p in range(4):
            hv[p] = c[p, 0] + ty * (c[p, 1] + ty * (c[p, 2] + ty * c[p, 3]))
            hd[p] = c[p, 1] + 2 * ty * c[p, 2] + 3 * ty * ty * c[p, 3]
            hh[p] = c[p, 2] + 3 * ty * c[p, 3]
        P00[r] = hv[0] + tx * (hv[1] + tx * (hv[2] + tx * hv[3]))
        P10[r] = hv[1] + 2 * tx * hv[2] + 3 * tx * tx * hv[3]
        P20h[r] = hv[2] + 3 * tx * hv[3]
        P01[r] = hd[0] + tx * (hd[1] + tx * (hd[2] + tx * hd[3]))
        P11[r] = hd[1] + 2 * tx * hd[2] + 3 * tx * tx * hd[3]
        P02h[r] = hh[0] + tx * (hh[1] + tx * (hh[2] + tx * hh[3]))
    w, d, dd = cr_weights(tz)
    return np.array([w @ P00, w @ P10, w @ P01, d @ P00, 2 * (w @ P20h), w @ P11, d @ P10, 2 * (w @ P02h),
                     d @ P01, dd @ P00])


@pytest.fixture(scope="module")
def grid(fields):
    phi, _ = fields.snapshot(12, 5)
    return phi.numpy().astype(np.float64) * 1e5  # O(1) values: compare at FP64 rounding


def test_weights_partition_of_unity():
    for t in np.linspace(0, 1, 7):
        w, d, dd = cr_weights(t)
        assert abs(w.sum() - 1) < 1e-15 and abs(d.sum()) < 1e-15 and abs(dd.sum()) < 1e-14
    w, _, _ = cr_weights(0.0)
    assert np.array_equal(w, [0, 1, 0, 0])


def test_records_reproduce_oracle_tricubic(oracle, grid):
    n = grid.shape[0]
    rng = np.random.default_rng(3)
    pos = rng.uniform(-2 * n, 3 * n, (200, 3))
    pos[:5] = np.floor(pos[:5])  # on grid nodes
    pos[5:10, 0] = np.floor(pos[5:10, 0]) + 0.5
    ref = oracle.tricubic(grid.astype(np.float32), pos)
    g32 = grid.astype(np.float32).astype(np.float64)
    got = np.array([eval_column(g32, p) for p in pos])
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0) / scale
    assert err.max() < 1e-12, err


def test_planted_mistakes_are_caught(oracle, grid):
    """A restatement with one plausible mistake must fail the comparison above."""
    global M
    n = grid.shape[0]
    pos = np.random.default_rng(4).uniform(0, n, (40, 3))
    g32 = grid.astype(np.float32).astype(np.float64)
    ref = oracle.tricubic(grid.astype(np.float32), pos)
    good = M.copy()
    try:
        M = good.copy()
        M[2, 3] = -1.0  # wrong cubic coefficient of tap 1
        bad = np.array([eval_column(g32, p) for p in pos])
        assert np.abs(bad - ref).max() > 1e-3 * np.abs(ref).max()
    finally:
        M = good
    # the x and y axes of the taps exchanged when building the records
    swapped = np.array([eval_column(np.swapaxes(g32, 1, 2), p) for p in pos])
    assert np.abs(swapped[:, 1] - ref[:, 1]).max() > 1e-3 * np.abs(ref[:, 1]).max()
