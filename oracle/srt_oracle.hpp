// Host restatement of the snapshot ray tracer's per-ray computation.
//
// TEST INFRASTRUCTURE ONLY.  This file gives the reference results: it is compiled into
// oracle/liboracle.so, which only the tests, the smoke check and the CPU
// baseline leg of the benchmark load.  The product path (the HIP library) never
// includes or calls it.
//
// Reference: snapshot-raytracer.cu (JulianAdamek/snapshot-raytracer).
//   cr_value/cr_d1/cr_d2 ........ .cu:195-212   (1-D Catmull-Rom, Keys a=-1/2)
//   tricubic ..................... .cu:215-312   (periodic value/gradient/Hessian)
//   rhs .......................... .cu:315-510   (19-component geodesic + Jacobi map)
//   rk4_integrate ................ .cu:513-550   (classical RK4, h = 1)
//   trace_ray .................... .cu:552-583   (initial conditions, deflection, 6 maps)
//   f32_to_f16_rne / shift scale . .cu:585-597, .cu:241-244 (FP16 x 2^27 shift storage)
//   pix2zphi_ring ................ chealpix pix2ang_ring64 (called at .cu:865),
//                                  restated from Gorski et al. 2005, ApJ 622, 759, sec. 4
//
// Two instantiations are used:
//   FP64 form: S = I = double.  Offsets x - floor(x) are exact, every
//              intermediate is double.  This is the accuracy oracle.
//   REF form:  S = double, I = float.  Interpolation runs in float with the
//              offset taken as float(x) - float(floor(x)) (.cu:228-230), and
//              the shift-coupling intermediates (N, E1, E2, T, L), the
//              interpolated value/gradient/Hessian and the modified Hessian are
//              rounded to float (.cu:366-396, .cu:484-492), as the reference
//              GPU code does.  State and RK4 stay double (.cu:513-549).
// A third instantiation with an operation-counting scalar type counts flops.
//
// Pins (tests/): every function here is checked against closed forms, brute
// force or first-order physics.  Every term of rhs is pinned by linearised
// general relativity (test_oracle_shift_gr.py: geodesic, Jacobi map and
// parallel-transported frame of the metric -(1+2 phi) dt^2 - 2 B.dx dt +
// (1-2 phi) dx^2 for a divergence-free shift and for a potential) and, for
// the shift terms, by a pure-gradient shift (test_oracle_shift_gauge.py).
// The REF form's float rounding points are checked only to tolerances: the
// reference is built with --use_fast_math, so its float results are not
// reproducible bit for bit on any host.
//
// Periodic wrap: the reference forms (i + o + 20 n) % n (.cu:241-256), valid
// while a ray stays within 20 boxes on the negative side.  This restatement
// uses the true modulo, which agrees with the reference wherever the
// reference's expression is in range.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace srt_oracle {

// ---------------------------------------------------------------- binary16 --
// IEEE-754 binary16 with round-to-nearest-even, written from the format
// definition.  Used for the reference's FP16 shift storage (B * 2^27).
constexpr float kShiftScale = 134217728.0f;  // 2^27, .cu:589 and .cu:241-244

inline uint16_t f32_to_f16_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {  // inf / nan (nan keeps its top payload bits, made quiet)
        if (ax == 0x7f800000u) return uint16_t(sign | 0x7c00u);
        return uint16_t(sign | 0x7c00u | 0x200u | ((ax >> 13) & 0x3ffu));
    }
    if (ax >= 0x477ff000u) return uint16_t(sign | 0x7c00u);  // >= 65520 rounds to inf
    if (ax >= 0x38800000u) {                                  // normal binary16 result
        const uint32_t mant = ax & 0x7fffffu;
        const uint32_t e = (ax >> 23) - 127u + 15u;
        uint32_t h = (e << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // carry may bump the exponent
        return uint16_t(sign | h);
    }
    if (ax < 0x33000000u) return uint16_t(sign);  // below 2^-25 (2^-25 itself ties to 0 below)
    // subnormal binary16: value / 2^-24 = m * 2^(e - 126), e in [102, 112]
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t e = ax >> 23;
    const uint32_t shift = 126u - e;  // 14 .. 24
    uint32_t h = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return uint16_t(sign | h);
}

inline float f16_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {  // subnormal: m * 2^-24, exact in float
            float v = float(m) * 5.9604644775390625e-08f;
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e - 15u + 127u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// --------------------------------------------------------------- HEALPix ----
inline int64_t isqrt64(int64_t v) {
    int64_t r = int64_t(std::sqrt(double(v)));
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// RING pixel -> (z = cos theta, s = sin theta, phi).  Gorski et al. 2005, sec. 4.
inline void pix2zphi_ring(int64_t n, int64_t p, double& z, double& s, double& phi) {
    const double kPi = 3.14159265358979323846;
    const int64_t ncap = 2 * n * (n - 1);
    const int64_t npix = 12 * n * n;
    if (p < ncap) {  // north polar cap: ring i has 4 i pixels
        const int64_t i = (1 + isqrt64(1 + 2 * p)) / 2;
        const int64_t j = p + 1 - 2 * i * (i - 1);
        const double t = double(i) * double(i) / (3.0 * double(n) * double(n));  // 1 - z
        z = 1.0 - t;
        s = std::sqrt(t * (2.0 - t));
        phi = (double(j) - 0.5) * kPi / (2.0 * double(i));
    } else if (p < npix - ncap) {  // equatorial belt: 4 n pixels per ring
        const int64_t q = p - ncap;
        const int64_t i = q / (4 * n) + n;
        const int64_t j = q % (4 * n) + 1;
        const double sh = ((i + n) & 1) ? 1.0 : 0.5;
        z = double(2 * n - i) * 2.0 / (3.0 * double(n));
        s = std::sqrt((1.0 - z) * (1.0 + z));
        phi = (double(j) - sh) * kPi / (2.0 * double(n));
    } else {  // south polar cap, mirrored
        const int64_t q = npix - p;
        const int64_t i = (1 + isqrt64(2 * q - 1)) / 2;
        const int64_t j = 4 * i + 1 - (q - 2 * i * (i - 1));
        const double t = double(i) * double(i) / (3.0 * double(n) * double(n));  // 1 + z
        z = t - 1.0;
        s = std::sqrt(t * (2.0 - t));
        phi = (double(j) - 0.5) * kPi / (2.0 * double(i));
    }
}

// ----------------------------------------------------------------- grids ----
// A periodic scalar grid, x fastest: value(ix, iy, iz) at ix + nx (iy + ny iz)
// (.cu:220, 253-256).  Exactly one of f32 / f16 is set; f16 holds B * 2^27.
struct Grid {
    const float* f32 = nullptr;
    const uint16_t* f16 = nullptr;
    int64_t nx = 0, ny = 0, nz = 0;

    static int64_t wrap(int64_t i, int64_t n) {
        int64_t r = i % n;
        return r < 0 ? r + n : r;
    }
    float at(int64_t ix, int64_t iy, int64_t iz) const {
        const int64_t o = wrap(ix, nx) + nx * (wrap(iy, ny) + ny * wrap(iz, nz));
        if (f16) return f16_to_f32(f16[o]) / kShiftScale;
        return f32[o];
    }
};

struct Fields {
    Grid phi;   // Weyl potential, sampled at x
    Grid b[3];  // shift components; b[a] is sampled at x - 1/2 e_a (.cu:377, 406, 432)
};

// ----------------------------------------------------------- Catmull-Rom ----
// Taps p0..p3 at offsets -1, 0, 1, 2.  With
//   a0 = 2 p1, a1 = p2 - p0, a2 = 2 p0 - 5 p1 + 4 p2 - p3, a3 = -p0 + 3 p1 - 3 p2 + p3
// value = (a0 + a1 t + a2 t^2 + a3 t^3) / 2, d/dt = (a1 + 2 a2 t + 3 a3 t^2) / 2,
// d2/dt2 = a2 + 3 a3 t  (.cu:195-212).
template <class T>
struct CrCoef {
    T a0, a1, a2, a3;
    explicit CrCoef(const T p[4])
        : a0(T(2) * p[1]),
          a1(p[2] - p[0]),
          a2(T(2) * p[0] - T(5) * p[1] + T(4) * p[2] - p[3]),
          a3(T(3) * p[1] - T(3) * p[2] + p[3] - p[0]) {}
};
template <class T>
inline T cr_value(const T p[4], T t) {
    CrCoef<T> c(p);
    return T(0.5) * (c.a0 + c.a1 * t + c.a2 * t * t + c.a3 * t * t * t);
}
template <class T>
inline T cr_d1(const T p[4], T t) {
    CrCoef<T> c(p);
    return T(0.5) * (c.a1 + T(2) * c.a2 * t + T(3) * c.a3 * t * t);
}
template <class T>
inline T cr_d2(const T p[4], T t) {
    CrCoef<T> c(p);
    return c.a2 + T(3) * c.a3 * t;
}

// ------------------------------------------------------------- tricubic ----
// Hessian order: xx, xy, xz, yy, yz, zz.
template <class I>
struct Interp {
    I v;
    I g[3];
    I h[6];
};

template <class I, class S>
inline Interp<I> tricubic(const Grid& G, S x, S y, S z) {
    using std::floor;
    const int64_t ix = int64_t(double(floor(x))), iy = int64_t(double(floor(y))), iz = int64_t(double(floor(z)));
    // offset within the cell: float(x) - float(ix) in the REF form (.cu:228-230),
    // exact x - ix when I is double
    const I tx = I(x) - I(double(ix));
    const I ty = I(y) - I(double(iy));
    const I tz = I(z) - I(double(iz));

    // pass 1, along x: value, d/dx, d2/dx2 of each of the 16 (y, z) rows
    I rv[4][4], rd[4][4], rdd[4][4];
    for (int b = 0; b < 4; ++b)
        for (int c = 0; c < 4; ++c) {
            I p[4];
            for (int a = 0; a < 4; ++a) p[a] = I(G.at(ix - 1 + a, iy - 1 + b, iz - 1 + c));
            rv[b][c] = cr_value(p, tx);
            rd[b][c] = cr_d1(p, tx);
            rdd[b][c] = cr_d2(p, tx);
        }
    // pass 2, along y, per z slice
    I zv[4], zdx[4], zdy[4], zdxx[4], zdyy[4], zdxy[4];
    for (int c = 0; c < 4; ++c) {
        I p[4];
        for (int b = 0; b < 4; ++b) p[b] = rv[b][c];
        zv[c] = cr_value(p, ty);
        zdy[c] = cr_d1(p, ty);
        zdyy[c] = cr_d2(p, ty);
        for (int b = 0; b < 4; ++b) p[b] = rd[b][c];
        zdx[c] = cr_value(p, ty);
        zdxy[c] = cr_d1(p, ty);
        for (int b = 0; b < 4; ++b) p[b] = rdd[b][c];
        zdxx[c] = cr_value(p, ty);
    }
    // pass 3, along z
    Interp<I> r;
    r.v = cr_value(zv, tz);
    r.g[0] = cr_value(zdx, tz);
    r.g[1] = cr_value(zdy, tz);
    r.g[2] = cr_d1(zv, tz);
    r.h[0] = cr_value(zdxx, tz);
    r.h[1] = cr_value(zdxy, tz);
    r.h[2] = cr_d1(zdx, tz);
    r.h[3] = cr_value(zdyy, tz);
    r.h[4] = cr_d1(zdy, tz);
    r.h[5] = cr_d2(zv, tz);
    return r;
}

// ------------------------------------------------------------------ RHS ----
// State y = (x[3], n[3], e1[3], e2[3], D, D', sigma1, sigma2, eps1, eps2, omega).
template <class S>
inline void unit3(const S* v, S* u) {
    using std::sqrt;
    const S norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) u[i] = v[i] / norm;
}

inline int hidx(int i, int j) {
    static const int k[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return k[i][j];
}

// The interpolation the right-hand side calls: the restatement above.  (The
// operation count in opcount.cpp substitutes a weights-form evaluation of the
// same interpolant to count a work-efficient algorithm.)
struct OracleTricubic {
    template <class I, class S>
    static Interp<I> eval(const Grid& G, S x, S y, S z) {
        return tricubic<I, S>(G, x, y, z);
    }
};

template <class S, class I, class Tri = OracleTricubic>
inline void rhs(const Fields& F, const S* y, S* dy) {
    using std::sqrt;
    S n[3], e1[3], e2[3];
    unit3(y + 3, n);
    unit3(y + 6, e1);
    unit3(y + 9, e2);
    const S D = y[12], Dp = y[13], s1 = y[14], s2 = y[15], eps1 = y[16], eps2 = y[17];

    // projectors: P = n n^T - 1, S1 = e1 e1^T - e2 e2^T, S2 = e1 e2^T + e2 e1^T (.cu:351-367)
    S P[3][3], S1[3][3], S2[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            P[i][j] = n[i] * n[j] - (i == j ? S(1) : S(0));
            S1[i][j] = e1[i] * e1[j] - e2[i] * e2[j];
            S2[i][j] = e1[i] * e2[j] + e1[j] * e2[i];
        }

    // shift: B_a sampled at x - 1/2 e_a; accumulate (.cu:370-460)
    //   N_j = sum_a n_a dB_a/dx_j
    //   E1  = 1/2 sum_a [e1_a (n . grad B_a) + n_a (e1 . grad B_a)]   (E2 likewise)
    //   T_ij = sum_a n_a H^a_ij - 1/2 [(n H^i)_j + (n H^j)_i]        (symmetric)
    //   L   = sum_a n_a tr H^a
    I Bv[3];
    I N[3] = {I(0), I(0), I(0)};
    I T[6] = {I(0), I(0), I(0), I(0), I(0), I(0)};
    I L = I(0), E1 = I(0), E2 = I(0);
    for (int a = 0; a < 3; ++a) {
        S p[3] = {y[0], y[1], y[2]};
        p[a] = p[a] - S(0.5);
        const Interp<I> r = Tri::template eval<I, S>(F.b[a], p[0], p[1], p[2]);
        Bv[a] = r.v;
        S g[3] = {S(r.g[0]), S(r.g[1]), S(r.g[2])};
        S H[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) H[i][j] = S(r.h[hidx(i, j)]);
        const S ng = n[0] * g[0] + n[1] * g[1] + n[2] * g[2];
        const S e1g = e1[0] * g[0] + e1[1] * g[1] + e1[2] * g[2];
        const S e2g = e2[0] * g[0] + e2[1] * g[1] + e2[2] * g[2];
        for (int j = 0; j < 3; ++j) N[j] = I(S(N[j]) + g[j] * n[a]);
        E1 = I(S(E1) + S(0.5) * e1[a] * ng + S(0.5) * n[a] * e1g);
        E2 = I(S(E2) + S(0.5) * e2[a] * ng + S(0.5) * n[a] * e2g);
        S nH[3];
        for (int j = 0; j < 3; ++j) nH[j] = n[0] * H[0][j] + n[1] * H[1][j] + n[2] * H[2][j];
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) {
                S c = H[i][j] * n[a];
                if (i == a && j == a)
                    c = c - nH[a];
                else if (i == a)
                    c = c - S(0.5) * nH[j];
                else if (j == a)
                    c = c - S(0.5) * nH[i];
                T[hidx(i, j)] = I(S(T[hidx(i, j)]) + c);
            }
        L = I(S(L) + n[a] * (H[0][0] + H[1][1] + H[2][2]));
    }

    // potential at x (.cu:462)
    const Interp<I> f = Tri::template eval<I, S>(F.phi, y[0], y[1], y[2]);
    const S phi = S(f.v);
    const S g[3] = {S(f.g[0]), S(f.g[1]), S(f.g[2])};
    S H[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H[i][j] = S(f.h[hidx(i, j)]);

    // x' = -n (1 + 2 phi) - B   (.cu:465-467)
    for (int i = 0; i < 3; ++i) dy[i] = -n[i] * (S(1) + S(2) * phi) - S(Bv[i]);
    // n' = -P (2 grad phi + N)   (.cu:469-471)
    for (int i = 0; i < 3; ++i) {
        S acc = S(0);
        for (int j = 0; j < 3; ++j) acc = acc + P[i][j] * (S(2) * g[j] + S(N[j]));
        dy[3 + i] = -acc;
    }
    // e1' = -n [2 (e1 . grad phi) + E1], e2' likewise   (.cu:473-479)
    const S e1g = e1[0] * g[0] + e1[1] * g[1] + e1[2] * g[2];
    const S e2g = e2[0] * g[0] + e2[1] * g[1] + e2[2] * g[2];
    for (int i = 0; i < 3; ++i) {
        dy[6 + i] = S(-2) * n[i] * e1g - n[i] * S(E1);
        dy[9 + i] = S(-2) * n[i] * e2g - n[i] * S(E2);
    }
    // Jacobi map (.cu:481-509)
    dy[12] = -Dp;
    const S gam = n[0] * g[0] + n[1] * g[1] + n[2] * g[2] +
                  S(0.25) * (n[0] * S(N[0]) + n[1] * S(N[1]) + n[2] * S(N[2]));
    S PH = S(0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) PH = PH + P[i][j] * H[i][j];
    dy[13] = S(-4) * Dp * gam - PH * D + S(0.5) * S(L) * D;
    // H <- H + T/2 after D'' has used the unmodified H (.cu:484-492)
    S Hm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Hm[i][j] = S(I(H[i][j] + S(0.5) * S(T[hidx(i, j)])));
    S S1H = S(0), S2H = S(0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            S1H = S1H + S1[i][j] * Hm[i][j];
            S2H = S2H + S2[i][j] * Hm[i][j];
        }
    dy[14] = S(4) * s1 * gam + S1H * D * D;
    dy[15] = S(4) * s2 * gam + S2H * D * D;
    if (D > S(0.5)) {
        const S D2 = D * D;
        const S q = sqrt(S(4) + eps1 * eps1 + eps2 * eps2);
        dy[13] = dy[13] + (s1 * s1 + s2 * s2) / (D2 * D);
        dy[16] = S(-2) * s1 * q / D2;
        dy[17] = S(-2) * s2 * q / D2;
        dy[18] = -(s1 * eps2 - s2 * eps1) / (D2 * (S(2) + q));
    } else {
        dy[16] = S(0);
        dy[17] = S(0);
        dy[18] = S(0);
    }
}

// ------------------------------------------------------------------ RK4 ----
// k1 = f(y), k2 = f(y + k1/2), k3 = f(y + k2/2), k4 = f(y + k3),
// y += ((k1 + 2 k2 + 2 k3) + k4) / 6, accumulated in that order (.cu:513-550).
template <class S, class I, class Tri = OracleTricubic>
inline void rk4_integrate(const Fields& F, S* y, int n_steps) {
    S tmp[19], k[19], acc[19];
    for (int step = 0; step < n_steps; ++step) {
        rhs<S, I, Tri>(F, y, acc);
        for (int j = 0; j < 19; ++j) tmp[j] = y[j] + S(0.5) * acc[j];
        rhs<S, I, Tri>(F, tmp, k);
        for (int j = 0; j < 19; ++j) {
            tmp[j] = y[j] + S(0.5) * k[j];
            acc[j] = acc[j] + S(2) * k[j];
        }
        rhs<S, I, Tri>(F, tmp, k);
        for (int j = 0; j < 19; ++j) {
            tmp[j] = y[j] + k[j];
            acc[j] = acc[j] + S(2) * k[j];
        }
        rhs<S, I, Tri>(F, tmp, k);
        for (int j = 0; j < 19; ++j) y[j] = y[j] + (acc[j] + k[j]) / S(6);
    }
}

// ------------------------------------------------------------ one ray ------
// Initial conditions for the ray of direction r = (s cos phi, s sin phi, z):
// x = observer, n = -r, e1 = theta-hat, e2 = (sin phi, -cos phi, 0), D = 0,
// D' = -1, rest 0 (.cu:560).  Outputs (.cu:566-582):
//   dA, eps1, eps2, omega, Delta . theta-hat, Delta . (sin phi, -cos phi, 0)
// with Delta = unit(x_end - observer) - r.
template <class S>
inline void initial_state(const double obs[3], double z, double s, double phi, S* y) {
    const double cp = std::cos(phi), sp = std::sin(phi);
    const double v[19] = {obs[0], obs[1], obs[2],             // x
                          -s * cp, -s * sp, -z,               // n
                          z * cp, z * sp, -s,                 // e1
                          sp, -cp, 0.0,                       // e2
                          0.0, -1.0,                          // D, D'
                          0.0, 0.0, 0.0, 0.0, 0.0};           // sigma1, sigma2, eps1, eps2, omega
    for (int j = 0; j < 19; ++j) y[j] = S(v[j]);
}

inline void final_maps(const double obs[3], double z, double s, double phi, const double* y, float out[6]) {
    const double cp = std::cos(phi), sp = std::sin(phi);
    double d[3] = {y[0] - obs[0], y[1] - obs[1], y[2] - obs[2]};
    const double norm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double r[3] = {s * cp, s * sp, z};
    for (int i = 0; i < 3; ++i) d[i] = d[i] / norm - r[i];
    out[0] = float(y[12]);
    out[1] = float(y[16]);
    out[2] = float(y[17]);
    out[3] = float(y[18]);
    out[4] = float(d[0] * z * cp + d[1] * z * sp - d[2] * s);
    out[5] = float(d[0] * sp - d[1] * cp);
}

template <class I>
inline void trace_ray(const Fields& F, const double obs[3], int64_t nside, int64_t pix, int n_steps,
                      float out[6], double* state_out = nullptr) {
    double z, s, phi;
    pix2zphi_ring(nside, pix, z, s, phi);
    double y[19];
    initial_state<double>(obs, z, s, phi, y);
    rk4_integrate<double, I>(F, y, n_steps);
    final_maps(obs, z, s, phi, y, out);
    if (state_out)
        for (int j = 0; j < 19; ++j) state_out[j] = y[j];
}

}  // namespace srt_oracle
