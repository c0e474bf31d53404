// Reverse-mode (adjoint) building blocks of the FP64 ray integrator.
//
// For a loss L(maps) the adjoint propagates dL/dy backwards through the RK4
// steps (.cu:513-550), the right-hand side (.cu:315-510) and the tricubic
// interpolation (.cu:215-312), and scatters dL/d(grid value) into the grids.
// Everything here is FP64.  The work is split so that each part runs with the
// parallelism it has:
//   * the interpolations of a ray's four fields (potential, B_x, B_y, B_z) are
//     independent: field_stencil / field_interp evaluate one field, so
//     several lanes can serve one ray (value, gradient, Hessian and, for the
//     reverse sweep, the third derivatives);
//   * the right-hand side is restated with its intermediates kept
//     (rhs_forward_from) from the four fields' interpolants, and its
//     vector-Jacobian product (rhs_vjp_bars) is pure algebra: it takes the
//     stored third derivatives for the position adjoint and returns, per field,
//     the adjoints of the interpolated value, gradient and Hessian;
//   * the grid gradient of those adjoints is the transpose of the 64-tap
//     contraction (scatter_taps), independent for every (ray, stage, field).
#pragma once

#include "srt_device.hpp"

namespace srt {
namespace adj {

// Global-memory pointers: loads and atomics through them are global (not
// flat) instructions even where the compiler cannot see the pointer's origin.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

// Catmull-Rom weights of orders 0..3 along one axis (third-derivative weights
// are constant: d3/dt3 of the value weights).
struct Axis4 {
    double w[4][4];  // w[order][tap]
};

__device__ __forceinline__ Axis4 cr_axis4(double t) {
    const Axis<double> a = cr_axis<double>(t);
    Axis4 r;
    for (int i = 0; i < 4; ++i) {
        r.w[0][i] = a.w[i];
        r.w[1][i] = a.d[i];
        r.w[2][i] = a.dd[i];
    }
    r.w[3][0] = -3.0;
    r.w[3][1] = 9.0;
    r.w[3][2] = -9.0;
    r.w[3][3] = 3.0;
    return r;
}

// derivative orders (x, y, z) of the 20 interpolation outputs: value, gradient,
// Hessian (xx xy xz yy yz zz), third derivatives (xxx xxy xxz xyy xyz xzz yyy
// yyz yzz zzz)
constexpr signed char kOrders[20][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2},
    {3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1}, {0, 1, 2}, {0, 0, 3}};

// index of the third derivative d3/(dx_i dx_j dx_k) among the ten third derivatives (kOrders 10..19)
constexpr int t3idx(int i, int j, int k) {
    // sort i <= j <= k, then xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    const int a = i < j ? i : j, b = i < j ? j : i;  // a <= b
    const int lo = a < k ? a : k, hi = b > k ? b : k, mid = i + j + k - lo - hi;
    constexpr signed char tab[3][3][3] = {
        {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}, {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}}, {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
    return tab[lo][mid][hi];
}

__device__ __forceinline__ int h6(int i, int j) {
    return (i == j) ? (i == 0 ? 0 : (i == 1 ? 3 : 5)) : ((i + j == 1) ? 1 : ((i + j == 2) ? 2 : 4));
}

// Adds the gradient of  sum_o bar[o] * interp_o  with respect to the 64 tap
// values (o = value, gradient, Hessian) into the padded gradient grid.  The
// workgroup is one wavefront.  Its lanes' stencils usually overlap (rays of
// one sky patch), so the 64 x 64 tap contributions are first added in LDS over
// the bounding box of the wavefront's stencils, and each touched cell then
// takes one global FP64 atomic; a box larger than the tile falls back to one
// global atomic per lane and tap.  Cell (x, y, z) of the padded grid is
// x + py y + pz z; (cx, cy, cz) is the stencil's tap (-1, -1, -1).
constexpr int kGradTileCells = 1024;
__shared__ double grad_tile[kGradTileCells];
__shared__ int grad_box[6];  // min x, y, z; max x, y, z of the stencil origins

// w[i] for a runtime i in [0, 4), by selects (a runtime index into registers
// would put the array in scratch)
__device__ __forceinline__ double pick4(const double w[4], int i) {
    return i == 0 ? w[0] : (i == 1 ? w[1] : (i == 2 ? w[2] : w[3]));
}

// Calls f(a, b, c, dL/d tap) for the 64 taps (a, b, c) of a stencil whose
// interpolation outputs have adjoints bar[10]; with c_lo <= c_hi only the z
// slices c in [c_lo, c_hi] (others are skipped).  Lane l visits the slices
// from (l mod 4) on, so lanes whose stencils coincide (rays near the
// observer) add to different cells in the same instruction: 128^3 / Nside 64
// / 64-step gradient 18.6 -> 15.5 ms (rotating the rows and columns as well
// gained nothing more).
template <class F>
__device__ __forceinline__ void for_each_tap(const Axis4& ax, const Axis4& ay, const Axis4& az, const double bar[10],
                                             F&& f, int c_lo = 0, int c_hi = 3) {
    // z-pass transpose: per slice c, adjoints of the six y-pass sums; the slices
    // are a loop (not unrolled), which keeps the code small
    const int c0 = max(c_lo, 0), nc = min(c_hi, 3) - c0 + 1, rot = (int)(threadIdx.x & 3);
#pragma unroll 1
    for (int k = 0; k < nc; ++k) {
        int c = k + rot;
        c = c0 + (c >= nc ? (c >= 2 * nc ? (c >= 3 * nc ? c - 3 * nc : c - 2 * nc) : c - nc) : c);
        const double w0 = pick4(az.w[0], c), w1 = pick4(az.w[1], c), w2 = pick4(az.w[2], c);
        const double yv = w0 * bar[0] + w1 * bar[3] + w2 * bar[9];
        const double yd = w0 * bar[2] + w1 * bar[8];
        const double ydd = w0 * bar[7];
        const double xd = w0 * bar[1] + w1 * bar[6];
        const double xdyd = w0 * bar[5];
        const double xdd = w0 * bar[4];
        for (int b = 0; b < 4; ++b) {
            const double rv = ay.w[0][b] * yv + ay.w[1][b] * yd + ay.w[2][b] * ydd;
            const double rd = ay.w[0][b] * xd + ay.w[1][b] * xdyd;
            const double rdd = ay.w[0][b] * xdd;
            for (int a = 0; a < 4; ++a) f(a, b, c, ax.w[0][a] * rv + ax.w[1][a] * rd + ax.w[2][a] * rdd);
        }
    }
}

// Box of the wavefront's stencils: origin (tap (-1,-1,-1)) and extent in
// cells, reduced over the active lanes from each lane's range of stencil
// origins [lo, hi].  The same in every lane.
struct TileBox {
    int x0, y0, z0, nx, ny, nz;
    __device__ int cells() const { return nx * ny * nz; }
};

__device__ __forceinline__ TileBox wave_box(const int lo[3], const int hi[3]) {
    const int lane = threadIdx.x & 63;
    const uint64_t active = __ballot(1);
    if (__popcll(active & ((uint64_t(1) << lane) - 1)) == 0) {
        grad_box[0] = grad_box[1] = grad_box[2] = 0x7fffffff;
        grad_box[3] = grad_box[4] = grad_box[5] = -1;
    }
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        (void)__hip_atomic_fetch_min(&grad_box[k], lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        (void)__hip_atomic_fetch_max(&grad_box[3 + k], hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    TileBox B;
    B.x0 = grad_box[0];
    B.y0 = grad_box[1];
    B.z0 = grad_box[2];
    B.nx = grad_box[3] - B.x0 + 4;
    B.ny = grad_box[4] - B.y0 + 4;
    B.nz = grad_box[5] - B.z0 + 4;
    __syncthreads();  // the box is read before the next call resets it
    return B;
}

// Clears the first `cells` entries of the LDS tile (the active lanes share the work).
__device__ __forceinline__ void tile_clear(int cells) {
    const int lane = threadIdx.x & 63;
    const uint64_t active = __ballot(1);
    const int n_active = __popcll(active), rank = __popcll(active & ((uint64_t(1) << lane) - 1));
    for (int q = rank; q < cells; q += n_active) grad_tile[q] = 0.0;
    __syncthreads();
}

// Adds the taps of one stencil (origin cx, cy, cz) that lie inside box B
// (its z range; x and y are whole) into the tile of B.
__device__ __forceinline__ void tile_add(const TileBox& B, int cx, int cy, int cz, const Axis4& ax, const Axis4& ay,
                                         const Axis4& az, const double bar[10]) {
    const int t0 = (cx - B.x0) + B.nx * ((cy - B.y0) + B.ny * (cz - B.z0));
    for_each_tap(
        ax, ay, az, bar,
        [&](int a, int b, int c, double pb) {
            (void)__hip_atomic_fetch_add(&grad_tile[t0 + a + B.nx * (b + B.ny * c)], pb, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
        },
        B.z0 - cz, B.z0 + B.nz - 1 - cz);
}

// Adds every non-zero tile entry to the padded gradient grid (one global FP64
// atomic per touched cell) after the tile is complete.
__device__ __forceinline__ void tile_flush(double* __restrict__ gbase_dev, const TileBox& B, size_t py, size_t pz) {
    const int lane = threadIdx.x & 63;
    const uint64_t active = __ballot(1);
    const int n_active = __popcll(active), rank = __popcll(active & ((uint64_t(1) << lane) - 1));
    __syncthreads();
    const int cells = B.cells(), nxy = B.nx * B.ny;
    for (int q = rank; q < cells; q += n_active) {
        const double v = grad_tile[q];
        if (v != 0.0) {
            const int qz = q / nxy, r = q - qz * nxy, qy = r / B.nx, qx = r - qy * B.nx;
            const gptr<double> dst = (gptr<double>)(gbase_dev + (size_t)(B.x0 + qx) + (size_t)(B.y0 + qy) * py +
                                                    (size_t)(B.z0 + qz) * pz);
            (void)__builtin_amdgcn_global_atomic_fadd_f64(dst, v);
        }
    }
    __syncthreads();  // the tile is read before the next call clears it
}

// z slabs of box B that fit the tile: slab k covers z planes
// [z0 + k dz, z0 + min((k + 1) dz, nz)); 0 when one z plane is larger than the
// tile (a box across the periodic wrap), and then the caller falls back to
// global atomics.
__device__ __forceinline__ int slab_planes(const TileBox& B) {
    const int nxy = B.nx * B.ny;
    return nxy > kGradTileCells ? 0 : kGradTileCells / nxy;
}

__device__ __forceinline__ TileBox slab(const TileBox& B, int zlo, int dz) {
    TileBox P = B;
    P.z0 = B.z0 + zlo;
    P.nz = min(dz, B.nz - zlo);
    return P;
}

// One stencil per lane: the tap gradients through the tile, in as many z
// slabs of the wavefront's box as the tile needs, else (a box wider than the
// tile in one plane) one global atomic per lane and tap.
__device__ __forceinline__ void scatter_taps(double* __restrict__ gbase_dev, int cx, int cy, int cz, size_t py,
                                             size_t pz, const Axis4& ax, const Axis4& ay, const Axis4& az,
                                             const double bar[10]) {
    const int c[3] = {cx, cy, cz};
    const TileBox B = wave_box(c, c);
    const int dz = slab_planes(B);
    if (dz > 0) {
        for (int zlo = 0; zlo < B.nz; zlo += dz) {
            const TileBox P = slab(B, zlo, dz);
            tile_clear(P.cells());
            tile_add(P, cx, cy, cz, ax, ay, az, bar);
            tile_flush(gbase_dev, P, py, pz);
        }
    } else {
        const size_t off = (size_t)cx + (size_t)cy * py + (size_t)cz * pz;
        for_each_tap(ax, ay, az, bar, [&](int a, int b, int c, double pb) {
            (void)__builtin_amdgcn_global_atomic_fadd_f64((gptr<double>)(gbase_dev + off + a + b * py + c * pz), pb);
        });
    }
}

// ------------------------------------------------------------ RHS forward ----
struct RhsTape {
    // stage inputs
    double nd[3], e1d[3], e2d[3], D, Dp, s1, s2, eps1, eps2;
    // normalised basis
    double rn, r1, r2, n[3], e1[3], e2[3];
    // interpolations: potential and the three shift components
    double phi, g[3], H[3][3];
    double Bv[3], gb[3][3], Hb[3][3][3];
    // intermediates
    double N[3], E1, E2, T[3][3], L, gam, lapse, v[3], nv, c1, c2, PH, Hm[3][3], S1H, S2H, q;
    bool br;
};

// One field's stencil: offset of tap (-1,-1,-1) in the padded grid, its
// padded index along x, y, z, and the cell offsets along x, y, z.
struct Stencil {
    size_t off;
    int c[3];
    double t[3];
};

// Stencil of field f at position x: f = 0 is the potential, sampled at x;
// f = 1 + a is shift component a, sampled at x - e_a / 2 (.cu:377, 406, 432).
// The staggered cell along a is the unstaggered one or the one below it, and
// its offset t -/+ 1/2 is exact.
__device__ __forceinline__ Stencil field_stencil(const Grids& G, const double x[3], int f) {
    Stencil st;
    for (int a = 0; a < 3; ++a) {
        double t;
        const int i0 = locate(x[a], G.n, G.inv_n, t);
        if (f == a + 1) {
            const bool upper = t >= 0.5;
            const int i = upper ? i0 : i0 - 1;
            st.c[a] = i < 0 ? i + G.n : i;
            st.t[a] = upper ? t - 0.5 : t + 0.5;
        } else {
            st.c[a] = i0;
            st.t[a] = t;
        }
    }
    st.off = (size_t)st.c[0] + (size_t)st.c[1] * (size_t)G.pad + (size_t)st.c[2] * (size_t)G.pad * (size_t)G.pad;
    return st;
}

__device__ __forceinline__ const double* field_grid64(const Grids& G, int f) {
    return f == 0 ? G.phi64 : (f == 1 ? G.b64[0] : (f == 2 ? G.b64[1] : G.b64[2]));
}

// Interpolation outputs of one field at its stencil, in kOrders order: the
// first 10 (value, gradient, Hessian), and with kThird the 10 third
// derivatives.  FP64 grid rows; the z slices are a loop, which keeps the code
// and the live values small.
template <bool kThird>
__device__ __forceinline__ void field_interp(const double* __restrict__ grid, size_t py, size_t pz,
                                             const Stencil& st, double o[20]) {
    constexpr int kOut = kThird ? 20 : 10;
    constexpr int kOrd = kThird ? 4 : 3;
    const Axis4 ax = cr_axis4(st.t[0]), ay = cr_axis4(st.t[1]), az = cr_axis4(st.t[2]);
    for (int k = 0; k < kOut; ++k) o[k] = 0.0;
    const double* base = grid + st.off;
#pragma unroll 1
    for (int c = 0; c < 4; ++c) {
        double Y[4][4];  // y-pass sums [ox][oy], ox + oy < kOrd
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) Y[i][j] = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const Row4d r = load_row(base + b * py + c * pz);
            double X[4];
#pragma unroll
            for (int ox = 0; ox < kOrd; ++ox)
                X[ox] = ax.w[ox][0] * r.x + ax.w[ox][1] * r.y + ax.w[ox][2] * r.z + ax.w[ox][3] * r.w;
#pragma unroll
            for (int ox = 0; ox < kOrd; ++ox)
#pragma unroll
                for (int oy = 0; oy < kOrd - ox; ++oy) Y[ox][oy] += ay.w[oy][b] * X[ox];
        }
        // the z weights of this slice, selected by constant indices
        const double wz[4] = {az.w[0][0] * (c == 0) + az.w[0][1] * (c == 1) + az.w[0][2] * (c == 2) + az.w[0][3] * (c == 3),
                              az.w[1][0] * (c == 0) + az.w[1][1] * (c == 1) + az.w[1][2] * (c == 2) + az.w[1][3] * (c == 3),
                              az.w[2][0] * (c == 0) + az.w[2][1] * (c == 1) + az.w[2][2] * (c == 2) + az.w[2][3] * (c == 3),
                              -3.0 * (c == 0) + 9.0 * (c == 1) - 9.0 * (c == 2) + 3.0 * (c == 3)};
#pragma unroll
        for (int k = 0; k < kOut; ++k) o[k] += wz[kOrders[k][2]] * Y[kOrders[k][0]][kOrders[k][1]];
    }
}

// Right-hand side at stage state s (19, FP64), keeping the intermediates, from
// the four fields' interpolants at s: q[10 f + o], o = value, gradient (3),
// Hessian (xx xy xz yy yz zz), read with stride `qs`.
__device__ __forceinline__ void rhs_forward_from(const double* s, const double* q, int qs, RhsTape& T,
                                                 double dy[19]);

__device__ __forceinline__ void rhs_forward_from(const double* s, const double* q, int qs, RhsTape& T,
                                                 double dy[19]) {
    for (int i = 0; i < 3; ++i) {
        T.nd[i] = s[3 + i];
        T.e1d[i] = s[6 + i];
        T.e2d[i] = s[9 + i];
    }
    T.D = s[12];
    T.Dp = s[13];
    T.s1 = s[14];
    T.s2 = s[15];
    T.eps1 = s[16];
    T.eps2 = s[17];
    T.rn = 1.0 / sqrt(T.nd[0] * T.nd[0] + T.nd[1] * T.nd[1] + T.nd[2] * T.nd[2]);
    T.r1 = 1.0 / sqrt(T.e1d[0] * T.e1d[0] + T.e1d[1] * T.e1d[1] + T.e1d[2] * T.e1d[2]);
    T.r2 = 1.0 / sqrt(T.e2d[0] * T.e2d[0] + T.e2d[1] * T.e2d[1] + T.e2d[2] * T.e2d[2]);
    for (int i = 0; i < 3; ++i) {
        T.n[i] = T.nd[i] * T.rn;
        T.e1[i] = T.e1d[i] * T.r1;
        T.e2[i] = T.e2d[i] * T.r2;
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double* qf = q + 10 * f * qs;
        const struct {
            double v, g[3];
        } r = {qf[0], {qf[qs], qf[2 * qs], qf[3 * qs]}};
        const double hxx = qf[4 * qs], hxy = qf[5 * qs], hxz = qf[6 * qs], hyy = qf[7 * qs], hyz = qf[8 * qs],
                     hzz = qf[9 * qs];
        const double H[3][3] = {{hxx, hxy, hxz}, {hxy, hyy, hyz}, {hxz, hyz, hzz}};
        if (f == 0) {
            T.phi = r.v;
            for (int i = 0; i < 3; ++i) {
                T.g[i] = r.g[i];
                for (int j = 0; j < 3; ++j) T.H[i][j] = H[i][j];
            }
        } else {
            T.Bv[f - 1] = r.v;
            for (int i = 0; i < 3; ++i) {
                T.gb[f - 1][i] = r.g[i];
                for (int j = 0; j < 3; ++j) T.Hb[f - 1][i][j] = H[i][j];
            }
        }
    }
    const double* n = T.n;
    const double* e1 = T.e1;
    const double* e2 = T.e2;
    T.E1 = 0.0;
    T.E2 = 0.0;
    T.L = 0.0;
    for (int j = 0; j < 3; ++j) T.N[j] = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T.T[i][j] = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double* gb = T.gb[a];
        double ng = 0, e1g = 0, e2g = 0;
        for (int j = 0; j < 3; ++j) {
            ng += n[j] * gb[j];
            e1g += e1[j] * gb[j];
            e2g += e2[j] * gb[j];
            T.N[j] += n[a] * gb[j];
        }
        T.E1 += 0.5 * (e1[a] * ng + n[a] * e1g);
        T.E2 += 0.5 * (e2[a] * ng + n[a] * e2g);
        T.L += n[a] * (T.Hb[a][0][0] + T.Hb[a][1][1] + T.Hb[a][2][2]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) T.T[i][j] += n[a] * T.Hb[a][i][j];
    }
    // - 1/2 [(n H^i)_j + (n H^j)_i]
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double nhij = 0, nhji = 0;
            for (int k = 0; k < 3; ++k) {
                nhij += n[k] * T.Hb[i][k][j];
                nhji += n[k] * T.Hb[j][k][i];
            }
            T.T[i][j] -= 0.5 * (nhij + nhji);
        }
    T.gam = 0.0;
    for (int i = 0; i < 3; ++i) T.gam += n[i] * T.g[i] + 0.25 * n[i] * T.N[i];
    T.lapse = 1.0 + 2.0 * T.phi;
    for (int i = 0; i < 3; ++i) dy[i] = -n[i] * T.lapse - T.Bv[i];
    T.nv = 0.0;
    for (int i = 0; i < 3; ++i) {
        T.v[i] = 2.0 * T.g[i] + T.N[i];
        T.nv += n[i] * T.v[i];
    }
    for (int i = 0; i < 3; ++i) dy[3 + i] = T.v[i] - n[i] * T.nv;
    T.c1 = T.E1;
    T.c2 = T.E2;
    for (int i = 0; i < 3; ++i) {
        T.c1 += 2.0 * e1[i] * T.g[i];
        T.c2 += 2.0 * e2[i] * T.g[i];
    }
    for (int i = 0; i < 3; ++i) {
        dy[6 + i] = -n[i] * T.c1;
        dy[9 + i] = -n[i] * T.c2;
    }
    dy[12] = -T.Dp;
    T.PH = -(T.H[0][0] + T.H[1][1] + T.H[2][2]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T.PH += n[i] * T.H[i][j] * n[j];
    dy[13] = -4.0 * T.Dp * T.gam - T.PH * T.D + 0.5 * T.L * T.D;
    T.S1H = 0.0;
    T.S2H = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T.Hm[i][j] = T.H[i][j] + 0.5 * T.T[i][j];
            T.S1H += (e1[i] * e1[j] - e2[i] * e2[j]) * T.Hm[i][j];
            T.S2H += (e1[i] * e2[j] + e2[i] * e1[j]) * T.Hm[i][j];
        }
    const double D2 = T.D * T.D;
    dy[14] = 4.0 * T.s1 * T.gam + T.S1H * D2;
    dy[15] = 4.0 * T.s2 * T.gam + T.S2H * D2;
    T.br = T.D > 0.5;
    T.q = sqrt(4.0 + T.eps1 * T.eps1 + T.eps2 * T.eps2);
    if (T.br) {
        dy[13] += (T.s1 * T.s1 + T.s2 * T.s2) / (D2 * T.D);
        dy[16] = -2.0 * T.s1 * T.q / D2;
        dy[17] = -2.0 * T.s2 * T.q / D2;
        dy[18] = -(T.s1 * T.eps2 - T.s2 * T.eps1) / (D2 * (2.0 + T.q));
    } else {
        dy[16] = dy[17] = dy[18] = 0.0;
    }
}

// Vector-Jacobian product of the right-hand side without its interpolations:
// given the tape T and a = dL/d(dy) (19), returns sbar = dL/ds (19).  `fr`
// points at the stage's per-field slots (field f, output o at
// fr[(20 f + o) * stride], kOrders order): the position adjoint reads the
// third derivatives there, and the adjoints of each field's interpolated
// value, gradient and Hessian (value, gx, gy, gz, xx, xy + yx, xz + zx, yy,
// yz + zy, zz: what scatter_taps takes) are written over its first 10 slots.
__device__ __forceinline__ void rhs_vjp_bars(const RhsTape& T, const double a[19], gptr<double> fr, int64_t stride,
                                             double sbar[19]) {
    const double* n = T.n;
    const double* e1 = T.e1;
    const double* e2 = T.e2;
    double nb[3] = {0, 0, 0}, e1b[3] = {0, 0, 0}, e2b[3] = {0, 0, 0};
    double Db = 0, Dpb = 0, s1b = 0, s2b = 0, eps1b = 0, eps2b = 0;
    double phib = 0, gb_[3] = {0, 0, 0}, Hbar[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Bvb[3] = {0, 0, 0}, gbb[3][3] = {{0}}, Hbb[3][3][3] = {{{0}}};
    double Nb[3] = {0, 0, 0}, E1b = 0, E2b = 0, Tb[3][3] = {{0}}, Lb = 0, gamb = 0;
    const double D = T.D, D2 = D * D;

    // omega', eps', and the sigma^2 / D^3 term (branch D > 1/2)
    if (T.br) {
        double qb = 0;
        s1b += a[16] * (-2.0 * T.q / D2);
        qb += a[16] * (-2.0 * T.s1 / D2);
        Db += a[16] * (4.0 * T.s1 * T.q / (D2 * D));
        s2b += a[17] * (-2.0 * T.q / D2);
        qb += a[17] * (-2.0 * T.s2 / D2);
        Db += a[17] * (4.0 * T.s2 * T.q / (D2 * D));
        const double num = T.s1 * T.eps2 - T.s2 * T.eps1;
        const double den = D2 * (2.0 + T.q);
        const double numb = -a[18] / den;
        const double denb = a[18] * num / (den * den);
        s1b += numb * T.eps2;
        eps2b += numb * T.s1;
        s2b -= numb * T.eps1;
        eps1b -= numb * T.s2;
        Db += denb * 2.0 * D * (2.0 + T.q);
        qb += denb * D2;
        const double ss = T.s1 * T.s1 + T.s2 * T.s2;
        s1b += a[13] * 2.0 * T.s1 / (D2 * D);
        s2b += a[13] * 2.0 * T.s2 / (D2 * D);
        Db += a[13] * (-3.0 * ss / (D2 * D2));
        eps1b += qb * T.eps1 / T.q;
        eps2b += qb * T.eps2 / T.q;
    }
    // sigma' = 4 sigma Gamma + S:Hm D^2
    s1b += a[14] * 4.0 * T.gam;
    s2b += a[15] * 4.0 * T.gam;
    gamb += a[14] * 4.0 * T.s1 + a[15] * 4.0 * T.s2;
    const double S1Hb = a[14] * D2, S2Hb = a[15] * D2;
    Db += a[14] * 2.0 * T.S1H * D + a[15] * 2.0 * T.S2H * D;
    double Hmb[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Hmb[i][j] = S1Hb * (e1[i] * e1[j] - e2[i] * e2[j]) + S2Hb * (e1[i] * e2[j] + e2[i] * e1[j]);
    for (int i = 0; i < 3; ++i) {
        double hm1 = 0, hm2 = 0;
        for (int j = 0; j < 3; ++j) {
            hm1 += T.Hm[i][j] * e1[j];
            hm2 += T.Hm[i][j] * e2[j];
        }
        e1b[i] += S1Hb * 2.0 * hm1 + S2Hb * 2.0 * hm2;
        e2b[i] += -S1Hb * 2.0 * hm2 + S2Hb * 2.0 * hm1;
    }
    // Hm = H + T/2
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Hbar[i][j] += Hmb[i][j];
            Tb[i][j] += 0.5 * Hmb[i][j];
        }
    // D'' = -4 D' Gamma - PH D + L D / 2
    Dpb += a[13] * (-4.0 * T.gam);
    gamb += a[13] * (-4.0 * T.Dp);
    const double PHb = a[13] * (-D);
    Db += a[13] * (-T.PH + 0.5 * T.L);
    Lb += a[13] * 0.5 * D;
    // PH = n^T H n - tr H
    for (int i = 0; i < 3; ++i) {
        double hn = 0;
        for (int j = 0; j < 3; ++j) {
            Hbar[i][j] += PHb * (n[i] * n[j] - (i == j ? 1.0 : 0.0));
            hn += T.H[i][j] * n[j];
        }
        nb[i] += PHb * 2.0 * hn;
    }
    // D' component
    Dpb -= a[12];
    // e' = -n c
    double c1b = 0, c2b = 0;
    for (int i = 0; i < 3; ++i) {
        nb[i] -= a[6 + i] * T.c1 + a[9 + i] * T.c2;
        c1b -= a[6 + i] * n[i];
        c2b -= a[9 + i] * n[i];
    }
    for (int i = 0; i < 3; ++i) {
        e1b[i] += c1b * 2.0 * T.g[i];
        e2b[i] += c2b * 2.0 * T.g[i];
        gb_[i] += c1b * 2.0 * e1[i] + c2b * 2.0 * e2[i];
    }
    E1b += c1b;
    E2b += c2b;
    // n' = v - n (n . v), v = 2 g + N
    double vb[3], nvb = 0;
    for (int i = 0; i < 3; ++i) {
        vb[i] = a[3 + i];
        nb[i] -= a[3 + i] * T.nv;
        nvb -= a[3 + i] * n[i];
    }
    for (int i = 0; i < 3; ++i) {
        nb[i] += nvb * T.v[i];
        vb[i] += nvb * n[i];
        gb_[i] += 2.0 * vb[i];
        Nb[i] += vb[i];
    }
    // x' = -n lapse - B
    double lapseb = 0;
    for (int i = 0; i < 3; ++i) {
        nb[i] -= a[i] * T.lapse;
        lapseb -= a[i] * n[i];
        Bvb[i] -= a[i];
    }
    phib += 2.0 * lapseb;
    // Gamma = n.g + n.N / 4
    for (int i = 0; i < 3; ++i) {
        nb[i] += gamb * (T.g[i] + 0.25 * T.N[i]);
        gb_[i] += gamb * n[i];
        Nb[i] += 0.25 * gamb * n[i];
    }
    // shift couplings N, E1, E2, T, L
#pragma unroll
    for (int aa = 0; aa < 3; ++aa) {
        const double* gb = T.gb[aa];
        double ng = 0, e1g = 0, e2g = 0;
        for (int j = 0; j < 3; ++j) {
            ng += n[j] * gb[j];
            e1g += e1[j] * gb[j];
            e2g += e2[j] * gb[j];
        }
        // N_j = sum_a n_a gb[a][j]
        for (int j = 0; j < 3; ++j) {
            nb[aa] += Nb[j] * gb[j];
            gbb[aa][j] += Nb[j] * n[aa];
        }
        // E1 = 1/2 sum_a [e1_a (n.gb_a) + n_a (e1.gb_a)]
        e1b[aa] += 0.5 * E1b * ng;
        nb[aa] += 0.5 * E1b * e1g;
        e2b[aa] += 0.5 * E2b * ng;
        nb[aa] += 0.5 * E2b * e2g;
        for (int j = 0; j < 3; ++j) {
            nb[j] += 0.5 * E1b * e1[aa] * gb[j] + 0.5 * E2b * e2[aa] * gb[j];
            e1b[j] += 0.5 * E1b * n[aa] * gb[j];
            e2b[j] += 0.5 * E2b * n[aa] * gb[j];
            gbb[aa][j] += 0.5 * E1b * (e1[aa] * n[j] + n[aa] * e1[j]) + 0.5 * E2b * (e2[aa] * n[j] + n[aa] * e2[j]);
        }
        // L = sum_a n_a tr Hb[a]
        nb[aa] += Lb * (T.Hb[aa][0][0] + T.Hb[aa][1][1] + T.Hb[aa][2][2]);
        for (int i = 0; i < 3; ++i) Hbb[aa][i][i] += Lb * n[aa];
        // T_ij += n_a Hb[a]_ij
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                nb[aa] += Tb[i][j] * T.Hb[aa][i][j];
                Hbb[aa][i][j] += Tb[i][j] * n[aa];
            }
    }
    // T_ij -= 1/2 [sum_k n_k Hb[i]_kj + sum_k n_k Hb[j]_ki]
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                nb[k] -= 0.5 * Tb[i][j] * (T.Hb[i][k][j] + T.Hb[j][k][i]);
                Hbb[i][k][j] -= 0.5 * Tb[i][j] * n[k];
                Hbb[j][k][i] -= 0.5 * Tb[i][j] * n[k];
            }

    // interpolation adjoints: field f has (value, gradient, Hessian) adjoints
    double xb[3] = {0, 0, 0};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double bar[10];
        double gcur[3];
        double Hcur[3][3];
        if (f == 0) {
            bar[0] = phib;
            for (int i = 0; i < 3; ++i) {
                bar[1 + i] = gb_[i];
                gcur[i] = T.g[i];
                for (int j = 0; j < 3; ++j) Hcur[i][j] = T.H[i][j];
            }
            bar[4] = Hbar[0][0];
            bar[5] = Hbar[0][1] + Hbar[1][0];
            bar[6] = Hbar[0][2] + Hbar[2][0];
            bar[7] = Hbar[1][1];
            bar[8] = Hbar[1][2] + Hbar[2][1];
            bar[9] = Hbar[2][2];
        } else {
            const int aa = f - 1;
            bar[0] = Bvb[aa];
            for (int i = 0; i < 3; ++i) {
                bar[1 + i] = gbb[aa][i];
                gcur[i] = T.gb[aa][i];
                for (int j = 0; j < 3; ++j) Hcur[i][j] = T.Hb[aa][i][j];
            }
            bar[4] = Hbb[aa][0][0];
            bar[5] = Hbb[aa][0][1] + Hbb[aa][1][0];
            bar[6] = Hbb[aa][0][2] + Hbb[aa][2][0];
            bar[7] = Hbb[aa][1][1];
            bar[8] = Hbb[aa][1][2] + Hbb[aa][2][1];
            bar[9] = Hbb[aa][2][2];
        }
        // position: d(value)/dx = g, d(g_i)/dx_k = H_ik, d(h_u)/dx_k = third derivative
        const gptr<double> ff = fr + (int64_t)(20 * f) * stride;
        double t3[10];
        for (int u = 0; u < 10; ++u) t3[u] = ff[(10 + u) * stride];
        constexpr int hu[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double acc = bar[0] * gcur[k];
#pragma unroll
            for (int i = 0; i < 3; ++i) acc += bar[1 + i] * Hcur[i][k];
#pragma unroll
            for (int u = 0; u < 6; ++u) acc += bar[4 + u] * t3[t3idx(hu[u][0], hu[u][1], k)];
            xb[k] += acc;
        }
        for (int o = 0; o < 10; ++o) ff[o * stride] = bar[o];
    }

    // normalisations: n = nd / |nd| etc.
    double nn = 0, n1 = 0, n2 = 0;
    for (int i = 0; i < 3; ++i) {
        nn += n[i] * nb[i];
        n1 += e1[i] * e1b[i];
        n2 += e2[i] * e2b[i];
    }
    for (int i = 0; i < 3; ++i) {
        sbar[i] = xb[i];
        sbar[3 + i] = T.rn * (nb[i] - n[i] * nn);
        sbar[6 + i] = T.r1 * (e1b[i] - e1[i] * n1);
        sbar[9 + i] = T.r2 * (e2b[i] - e2[i] * n2);
    }
    sbar[12] = Db;
    sbar[13] = Dpb;
    sbar[14] = s1b;
    sbar[15] = s2b;
    sbar[16] = eps1b;
    sbar[17] = eps2b;
    sbar[18] = 0.0;
}

}  // namespace adj
}  // namespace srt
