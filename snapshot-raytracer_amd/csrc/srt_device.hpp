// Device-side building blocks of the ray kernel (gfx950).
//
//   healpix ring pixel -> direction ..... chealpix pix2ang_ring64 as called at
//                                         snapshot-raytracer.cu:865 (moved on device)
//   nested -> ring pixel index .......... used only to order lanes (NESTED tiles)
//   Catmull-Rom weights ................. .cu:195-212, as weights w(t) on the 4 taps
//   tricubic value/gradient/Hessian ..... .cu:215-312, fused x -> y -> z contraction
//   right-hand side ..................... .cu:315-510
//
// Grid layout in HBM: every field (potential, B_x, B_y, B_z) is a padded cube
// of edge P = n + 3, x fastest, holding value(i, j, k) at
// (i + 1) + P * ((j + 1) + P * (k + 1)) for i, j, k in [-1, n + 1], with the
// periodic images in the halo.  Every 4-tap row of the stencil is then 16
// contiguous bytes and no index wraps inside the stencil: a position is wrapped
// once per interpolation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srt {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------- HEALPix ----
__device__ __forceinline__ int64_t isqrt64(int64_t v) {
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// RING pixel p of resolution n -> z = cos(theta), s = sin(theta), phi
// (Gorski et al. 2005, ApJ 622, 759, sec. 4).
__device__ __forceinline__ void ring_pixel_direction(int64_t n, int64_t p, double& z, double& s, double& phi) {
    const int64_t ncap = 2 * n * (n - 1);
    const int64_t npix = 12 * n * n;
    const double dn = (double)n;
    if (p < ncap) {
        const int64_t i = (1 + isqrt64(1 + 2 * p)) / 2;
        const int64_t j = p + 1 - 2 * i * (i - 1);
        const double t = (double)i * (double)i / (3.0 * dn * dn);
        z = 1.0 - t;
        s = sqrt(t * (2.0 - t));
        phi = ((double)j - 0.5) * kPi / (2.0 * (double)i);
    } else if (p < npix - ncap) {
        const int64_t q = p - ncap;
        const int64_t i = q / (4 * n) + n;
        const int64_t j = q - (i - n) * (4 * n) + 1;
        const double sh = ((i + n) & 1) ? 1.0 : 0.5;
        z = (double)(2 * n - i) * 2.0 / (3.0 * dn);
        s = sqrt((1.0 - z) * (1.0 + z));
        phi = ((double)j - sh) * kPi / (2.0 * dn);
    } else {
        const int64_t q = npix - p;
        const int64_t i = (1 + isqrt64(2 * q - 1)) / 2;
        const int64_t j = 4 * i + 1 - (q - 2 * i * (i - 1));
        const double t = (double)i * (double)i / (3.0 * dn * dn);
        z = t - 1.0;
        s = sqrt(t * (2.0 - t));
        phi = ((double)j - 0.5) * kPi / (2.0 * (double)i);
    }
}

// NESTED pixel -> RING pixel (face / in-face coordinates, Gorski et al. 2005),
// for power-of-two Nside (the NESTED scheme's domain).  Used to hand each
// wavefront a compact 8 x 8 patch of the sky in full-sky traces.
__device__ __forceinline__ uint32_t compact_even_bits(uint64_t v) {
    v &= 0x5555555555555555ull;
    v = (v | (v >> 1)) & 0x3333333333333333ull;
    v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
    v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
    v = (v | (v >> 16)) & 0x00000000ffffffffull;
    return (uint32_t)v;
}

__device__ __forceinline__ int64_t nest_to_ring(int64_t n, int64_t p) {
    const int64_t npface = n * n;
    const int face = (int)(p / npface);
    const uint64_t ipf = (uint64_t)(p - (int64_t)face * npface);
    const int64_t ix = compact_even_bits(ipf);
    const int64_t iy = compact_even_bits(ipf >> 1);
    const int jrll = (face < 4) ? 2 : (face < 8 ? 3 : 4);
    const int jpll = (face < 4) ? (2 * face + 1) : (face < 8 ? 2 * (face - 4) : 2 * (face - 8) + 1);
    const int64_t jr = (int64_t)jrll * n - ix - iy - 1;
    const int64_t npix = 12 * n * n, ncap = 2 * n * (n - 1);
    int64_t nr, before, kshift;
    if (jr < n) {
        nr = jr;
        before = 2 * nr * (nr - 1);
        kshift = 0;
    } else if (jr > 3 * n) {
        nr = 4 * n - jr;
        before = npix - 2 * (nr + 1) * nr;
        kshift = 0;
    } else {
        nr = n;
        before = ncap + (jr - n) * 4 * n;
        kshift = (jr - n) & 1;
    }
    int64_t jp = ((int64_t)jpll * nr + ix - iy + 1 + kshift) / 2;
    if (jp > 4 * n)
        jp -= 4 * n;
    else if (jp < 1)
        jp += 4 * n;
    return before + jp - 1;
}

// RING pixel -> lane-order key face * 4^L + interleave(ix, iy), L = the bit
// length of n - 1, from the pixel's base face and in-face coordinates
// (Gorski et al. 2005).  The key is unique for every Nside; for power-of-two
// Nside it is the NESTED pixel index.  Sorting rays by it gives each
// wavefront a compact patch of sky.
__device__ __forceinline__ uint64_t spread_bits(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__device__ __forceinline__ int64_t ring_order_key(int64_t n, int64_t p) {
    const int64_t ncap = 2 * n * (n - 1), npix = 12 * n * n;
    int64_t iring, iphi, kshift, nr;
    int face;
    if (p < ncap) {  // north polar cap: ring iring has 4 iring pixels
        iring = (1 + isqrt64(1 + 2 * p)) / 2;
        iphi = p + 1 - 2 * iring * (iring - 1);
        kshift = 0;
        nr = iring;
        face = (int)((iphi - 1) / nr);
    } else if (p < npix - ncap) {  // equatorial belt
        const int64_t q = p - ncap, t = q / (4 * n);
        iring = t + n;
        iphi = q - t * 4 * n + 1;
        kshift = (iring + n) & 1;
        nr = n;
        const int64_t ire = t + 1, irm = 2 * n + 2 - ire;
        const int64_t ifm = (iphi - ire / 2 + n - 1) / n, ifp = (iphi - irm / 2 + n - 1) / n;
        if (ifp == ifm)
            face = (int)(ifp == 4 ? 4 : ifp + 4);
        else if (ifp < ifm)
            face = (int)ifp;
        else
            face = (int)(ifm + 8);
    } else {  // south polar cap
        const int64_t q = npix - p;
        const int64_t r = (1 + isqrt64(2 * q - 1)) / 2;
        iphi = 4 * r + 1 - (q - 2 * r * (r - 1));
        kshift = 0;
        nr = r;
        iring = 4 * n - r;
        face = (int)(8 + (iphi - 1) / nr);
    }
    const int jrll = (face < 4) ? 2 : (face < 8 ? 3 : 4);
    const int jpll = (face < 4) ? (2 * face + 1) : (face < 8 ? 2 * (face - 4) : 2 * (face - 8) + 1);
    const int64_t irt = iring - (int64_t)jrll * n + 1;
    int64_t ipt = 2 * iphi - (int64_t)jpll * nr - kshift - 1;
    if (ipt >= 2 * n) ipt -= 8 * n;
    const int64_t ix = (ipt - irt) >> 1, iy = (-ipt - irt) >> 1;
    int L = 0;
    while ((int64_t(1) << L) < n) ++L;
    return ((int64_t)face << (2 * L)) | (int64_t)(spread_bits((uint32_t)ix) | (spread_bits((uint32_t)iy) << 1));
}

// ---------------------------------------------------------- Catmull-Rom ----
// For taps p0..p3 at offsets -1, 0, 1, 2 the Keys (a = -1/2) cubic is
// value = sum w_i p_i, d/dt = sum d_i p_i, d2/dt2 = sum dd_i p_i with
//   w  = 1/2 (-t + 2t^2 - t^3,  2 - 5t^2 + 3t^3,  t + 4t^2 - 3t^3,  -t^2 + t^3)
//   d  = 1/2 (-1 + 4t - 3t^2,  -10t + 9t^2,       1 + 8t - 9t^2,   -2t + 3t^2)
//   dd =     (2 - 3t,          -5 + 9t,           4 - 9t,          -1 + 3t)
// (the same polynomials as .cu:195-212, regrouped by tap).
template <class R>
struct Axis {
    R w[4], d[4], dd[4];
};

template <class R>
__device__ __forceinline__ Axis<R> cr_axis(R t) {
    const R t2 = t * t, t3 = t2 * t;
    const R h = R(0.5);
    Axis<R> a;
    a.w[0] = h * (R(2) * t2 - t - t3);
    a.w[1] = h * (R(2) - R(5) * t2 + R(3) * t3);
    a.w[2] = h * (t + R(4) * t2 - R(3) * t3);
    a.w[3] = h * (t3 - t2);
    a.d[0] = h * (R(4) * t - R(1) - R(3) * t2);
    a.d[1] = h * (R(9) * t2 - R(10) * t);
    a.d[2] = h * (R(1) + R(8) * t - R(9) * t2);
    a.d[3] = h * (R(3) * t2 - R(2) * t);
    a.dd[0] = R(2) - R(3) * t;
    a.dd[1] = R(9) * t - R(5);
    a.dd[2] = R(4) - R(9) * t;
    a.dd[3] = R(3) * t - R(1);
    return a;
}

// ---------------------------------------------------------------- grids ----
struct Grids {
    const float* phi;   // potential, sampled at x
    const float* b[3];  // shift components; b[a] sampled at x - e_a / 2 (.cu:377, 406, 432)
    int n;              // cube edge
    int pad;            // padded edge n + 3
    double inv_n;
    // FP64 copies of the same padded cubes (the FP64 kernel reads these, so it
    // converts no taps); null when not built
    const double* phi64;
    const double* b64[3];
    // bicubic cell records (potential, B_x, B_y, B_z), see record_offset;
    // null when the context holds only the padded cubes
    const float* rec[4];
    int rec_shift;  // log2 of the column brick edge
    int rec_nb;     // column bricks per row: ceil(n / brick)
};

// ------------------------------------------------------ bicubic records ----
// Bicubic cell records.  For every cell column (i, j), i, j in [0, n), and
// every padded plane p in [0, n + 3) (grid plane p - 1 mod n) a record holds
// the 16 monomial coefficients of the Catmull-Rom bicubic of that cell on that
// plane,
//     f(tx, ty) = sum_{a,b} v(i - 1 + a, j - 1 + b) w_a(tx) w_b(ty)
//               = sum_{p,q} c[p][q] tx^p ty^q,
// stored as c[p][0..3] for p = 0..3 (64 bytes).  The records of one column
// are consecutive in the plane index, so the 4 x 4 x 4 stencil of cell
// (i, j, k) is the 256 contiguous bytes of planes k .. k + 3.  Columns are
// tiled in bricks of 2^s x 2^s (s = rec_shift): column (i, j) is number
//     ((j >> s) * nb + (i >> s)) * 4^s + (j & m) * 2^s + (i & m),  m = 2^s - 1,
// so rays that cross neighbouring columns stay within a few pages.
__device__ __forceinline__ size_t record_offset(const Grids& G, int i, int j, int k) {
    const int s = G.rec_shift;
    const int m = (1 << s) - 1;
    const uint32_t col = ((uint32_t)((j >> s) * G.rec_nb + (i >> s)) << (2 * s)) + ((uint32_t)(j & m) << s) +
                         (uint32_t)(i & m);
    return ((size_t)col * (size_t)G.pad + (size_t)k) * 16;
}

// Catmull-Rom weights as polynomials: w_a(t) = sum_p kCrMonomial[a][p] t^p
// (the polynomials of cr_axis below, by power of t).
constexpr double kCrMonomial[4][4] = {{0.0, -0.5, 1.0, -0.5},
                                      {1.0, 0.0, -2.5, 1.5},
                                      {0.0, 0.5, 2.0, -1.5},
                                      {0.0, 0.0, -0.5, 0.5}};

// t, 2t, 3t and 3t^2: the factors of Horner evaluation of a cubic and of its
// first and (halved) second derivative.
template <class R>
struct Powers {
    R t, t2, t3, t3sq;
};

template <class R>
__device__ __forceinline__ Powers<R> powers(R t) {
    Powers<R> p;
    p.t = t;
    p.t2 = t + t;
    p.t3 = R(3) * t;
    p.t3sq = p.t3 * t;
    return p;
}

// Wraps floor(x) into [0, n) and returns the offset x - floor(x) (exact in FP64).
// kPow2: n is a power of two, and the wrap is the low bits of floor(x) (two's
// complement, so also for negative x).
template <bool kPow2 = false>
__device__ __forceinline__ int locate(double x, int n, double inv_n, double& t) {
    const double fl = floor(x);
    t = x - fl;
    if constexpr (kPow2) return (int)fl & (n - 1);
    const double q = floor(fl * inv_n);
    // fl and q * n are integers below 2^53, so r is exact; one correction
    // covers the rounding of fl * inv_n
    int r = (int)(fl - q * (double)n);
    r += (r < 0) ? n : 0;
    r -= (r >= n) ? n : 0;
    return r;
}

// Grid rows are always in global memory; the explicit address space keeps the
// loads global (not flat) also in functions that are not inlined into a
// kernel, where the compiler cannot see where a pointer comes from.
struct alignas(4) Row4 {
    float x, y, z, w;
};

__device__ __forceinline__ Row4 load_row(const float* p) {
    const __attribute__((address_space(1))) float* g = (const __attribute__((address_space(1))) float*)p;
    return Row4{g[0], g[1], g[2], g[3]};
}

struct alignas(8) Row4d {
    double x, y, z, w;
};

__device__ __forceinline__ Row4d load_row(const double* p) {
    const __attribute__((address_space(1))) double* g = (const __attribute__((address_space(1))) double*)p;
    return Row4d{g[0], g[1], g[2], g[3]};
}

// value, gradient (x, y, z), Hessian (xx, xy, xz, yy, yz, zz)
template <class R>
struct Interp {
    R v, gx, gy, gz, hxx, hxy, hxz, hyy, hyz, hzz;
};

// Tricubic interpolation as one fused contraction: 16 rows x 3 (x-pass),
// 16 x 6 (y-pass), 4 x 10 (z-pass) = 328 multiply-adds.  `base` points at
// tap (-1, -1, -1) of the stencil in the padded grid.
// kGroupedLoads (the kernels with one wavefront per SIMD, FP64 and REF):
// the 16 row loads are issued together, ahead of a scheduling fence, and each
// z-slice's contraction stays behind a fence of its own.  With no second
// wavefront to run meanwhile, the stencil's load latencies then overlap
// instead of following each other: at 512^3 / Nside 1024 / 256 steps the
// FP64 kernel runs 1158 instead of 1618 ms (loads grouped per z-slice: 1253
// ms; 1134 ms with a fence before each field as well) and the REF kernel 899
// instead of 927 ms.  The mixed padded-grid kernel
// (two wavefronts per SIMD) gains nothing from either.
template <class R, bool kGroupedLoads = false, class GT>
__device__ __forceinline__ Interp<R> tricubic(const GT* __restrict__ base, size_t pitch_y, size_t pitch_z,
                                              const Axis<R>& ax, const Axis<R>& ay, const Axis<R>& az) {
    Interp<R> o;
    using Row = decltype(load_row(base));
    Row rows[kGroupedLoads ? 16 : 1];
    if constexpr (kGroupedLoads) {
#pragma unroll
        for (int q = 0; q < 16; ++q) rows[q] = load_row(base + (q & 3) * pitch_y + (q >> 2) * pitch_z);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if constexpr (kGroupedLoads) __builtin_amdgcn_sched_barrier(0);
        R yv, yd, ydd, xd, xdyd, xdd;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            Row q;
            if constexpr (kGroupedLoads)
                q = rows[4 * c + b];
            else
                q = load_row(base + b * pitch_y + c * pitch_z);
            const R p0 = R(q.x), p1 = R(q.y), p2 = R(q.z), p3 = R(q.w);
            const R v = ax.w[0] * p0 + ax.w[1] * p1 + ax.w[2] * p2 + ax.w[3] * p3;
            const R d = ax.d[0] * p0 + ax.d[1] * p1 + ax.d[2] * p2 + ax.d[3] * p3;
            const R dd = ax.dd[0] * p0 + ax.dd[1] * p1 + ax.dd[2] * p2 + ax.dd[3] * p3;
            if (b == 0) {
                yv = ay.w[0] * v;
                yd = ay.d[0] * v;
                ydd = ay.dd[0] * v;
                xd = ay.w[0] * d;
                xdyd = ay.d[0] * d;
                xdd = ay.w[0] * dd;
            } else {
                yv += ay.w[b] * v;
                yd += ay.d[b] * v;
                ydd += ay.dd[b] * v;
                xd += ay.w[b] * d;
                xdyd += ay.d[b] * d;
                xdd += ay.w[b] * dd;
            }
        }
        if (c == 0) {
            o.v = az.w[0] * yv;
            o.gz = az.d[0] * yv;
            o.hzz = az.dd[0] * yv;
            o.gy = az.w[0] * yd;
            o.hyz = az.d[0] * yd;
            o.hyy = az.w[0] * ydd;
            o.gx = az.w[0] * xd;
            o.hxz = az.d[0] * xd;
            o.hxy = az.w[0] * xdyd;
            o.hxx = az.w[0] * xdd;
        } else {
            o.v += az.w[c] * yv;
            o.gz += az.d[c] * yv;
            o.hzz += az.dd[c] * yv;
            o.gy += az.w[c] * yd;
            o.hyz += az.d[c] * yd;
            o.hyy += az.w[c] * ydd;
            o.gx += az.w[c] * xd;
            o.hxz += az.d[c] * xd;
            o.hxy += az.w[c] * xdyd;
            o.hxx += az.w[c] * xdd;
        }
    }
    return o;
}

// Tricubic value/gradient/Hessian from the four bicubic records of planes
// k .. k + 3 at `rec`: per plane, the bicubic and its x, y, xx, xy, yy
// derivatives by Horner's rule (38 multiply-adds), then the Catmull-Rom
// z-pass over the four planes (40) and two doublings: 194 operations, against
// 328 multiply-adds for tricubic() on the raw taps.  The same interpolant as tricubic(): only the
// order of operations differs.
// The 16 record loads are issued together, at raised wavefront priority,
// ahead of a scheduling fence: at 512^3 / Nside 1024 / 256 steps the record
// kernel runs 439-441 instead of 450 ms (the priority accounts for ~1 ms).
template <class R>
__device__ __forceinline__ Interp<R> bicubic_column(const R* __restrict__ rec, const Powers<R>& px,
                                                    const Powers<R>& py, const Axis<R>& az) {
    R P00[4], P10[4], P20h[4], P01[4], P11[4], P02h[4];
    decltype(load_row(rec)) rows[16];
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int q = 0; q < 16; ++q) rows[q] = load_row(rec + 4 * q);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        R hv[4], hd[4], hh[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const auto c = rows[4 * r + p];
            hv[p] = c.x + py.t * (c.y + py.t * (c.z + py.t * c.w));  // sum_q c[p][q] ty^q
            hd[p] = c.y + py.t2 * c.z + py.t3sq * c.w;                 // d/dty
            hh[p] = c.z + py.t3 * c.w;                                 // (d2/dty2) / 2
        }
        P00[r] = hv[0] + px.t * (hv[1] + px.t * (hv[2] + px.t * hv[3]));
        P10[r] = hv[1] + px.t2 * hv[2] + px.t3sq * hv[3];
        P20h[r] = hv[2] + px.t3 * hv[3];
        P01[r] = hd[0] + px.t * (hd[1] + px.t * (hd[2] + px.t * hd[3]));
        P11[r] = hd[1] + px.t2 * hd[2] + px.t3sq * hd[3];
        P02h[r] = hh[0] + px.t * (hh[1] + px.t * (hh[2] + px.t * hh[3]));
    }
    Interp<R> o;
    o.v = az.w[0] * P00[0];
    o.gz = az.d[0] * P00[0];
    o.hzz = az.dd[0] * P00[0];
    o.gx = az.w[0] * P10[0];
    o.hxz = az.d[0] * P10[0];
    o.gy = az.w[0] * P01[0];
    o.hyz = az.d[0] * P01[0];
    o.hxx = az.w[0] * P20h[0];
    o.hxy = az.w[0] * P11[0];
    o.hyy = az.w[0] * P02h[0];
#pragma unroll
    for (int r = 1; r < 4; ++r) {
        o.v += az.w[r] * P00[r];
        o.gz += az.d[r] * P00[r];
        o.hzz += az.dd[r] * P00[r];
        o.gx += az.w[r] * P10[r];
        o.hxz += az.d[r] * P10[r];
        o.gy += az.w[r] * P01[r];
        o.hyz += az.d[r] * P01[r];
        o.hxx += az.w[r] * P20h[r];
        o.hxy += az.w[r] * P11[r];
        o.hyy += az.w[r] * P02h[r];
    }
    o.hxx += o.hxx;
    o.hyy += o.hyy;
    return o;
}

// ------------------------------------------------------------------ RHS ----
// Precision policy of one kernel form.
//   RI: interpolation arithmetic and the type the shift-coupling
//       intermediates are rounded to; RA: right-hand-side assembly and the
//       derivative type; kRefOffset: cell offset taken as float(x) - float(i)
//       (.cu:228-230) instead of the exact x - i.
//   kStateInLds: the kernel keeps the RK4 state and stage sum in LDS;
//   kGrid64: interpolate from the FP64 copies of the grids;
//   kBicubic: interpolate from the bicubic cell records (FP32 forms);
//   kWavesPerSimd: wavefronts per SIMD the register allocation must allow.
template <class RI_, class RA_, bool kRefOffset_, bool kStateInLds_ = false, bool kGrid64_ = false,
          bool kBicubic_ = false, int kWavesPerSimd_ = 1>
struct Form {
    using RI = RI_;
    using RA = RA_;
    static constexpr bool kRefOffset = kRefOffset_;
    static constexpr bool kStateInLds = kStateInLds_;
    static constexpr bool kGrid64 = kGrid64_;
    static constexpr bool kBicubic = kBicubic_;
    static constexpr int kWavesPerSimd = kWavesPerSimd_;
    static constexpr bool kPow2 = false;  // grid edge known to be a power of two
};
// A form on a grid whose edge is a power of two (the periodic wrap of a cell
// index is a mask: ~2 % of the mixed kernels' time at 512^3).
template <class F>
struct Pow2Grid : F {
    static constexpr bool kPow2 = true;
};
using FormFP64 = Form<double, double, false, true, true>;
using FormMixed = Form<float, float, false, true, false, false, 2>;
using FormMixedBicubic = Form<float, float, false, true, false, true, 2>;
using FormRef = Form<float, double, true, false, false>;

template <class F>
__device__ __forceinline__ typename F::RI cell_offset(double x, double t) {
    using RI = typename F::RI;
    if constexpr (F::kRefOffset) {
        return (RI)((float)x - (float)(x - t));  // float(x) - float(floor(x))
    } else {
        return (RI)t;
    }
}

template <class R>
__device__ __forceinline__ R rnorm3(R a, R b, R c) {
    if constexpr (sizeof(R) == 4) {
        return __frsqrt_rn(a * a + b * b + c * c);
    } else {
        return R(1) / sqrt(a * a + b * b + c * c);
    }
}

template <class R>
__device__ __forceinline__ R sqrt_r(R a) {
    return sqrt(a);
}

// 1 / a: the hardware reciprocal (1 ulp) in FP32, IEEE division in FP64
template <class R>
__device__ __forceinline__ R rcp_r(R a) {
    if constexpr (sizeof(R) == 4) {
        return __builtin_amdgcn_rcpf(a);
    } else {
        return R(1) / a;
    }
}

// Stage state in FP64: position x[3], direction nd[3] (unnormalised), D, D'.
// ys = (e1[3], e2[3], sigma1, sigma2, eps1, eps2) in RA (the rotation angle is
// never read).  Outputs: kx = x' in FP64, assembled as -nd * (|nd|^-1 (1 + 2 phi))
// - B so the transverse direction keeps FP64 accuracy; dy[3..18] the other
// derivatives in RA (dy[0..2] repeat x', dy[12] = -D').  The dA > 1/2 branch
// (.cu:497-509) is decided on the FP64 D.
template <class F>
__device__ __forceinline__ void rhs(const Grids& G, const double* __restrict__ x, const double* __restrict__ nd,
                                    double Dd, double Dpd, const typename F::RA* __restrict__ ys,
                                    double* __restrict__ kx, typename F::RA* __restrict__ dy) {
    using RI = typename F::RI;
    using RA = typename F::RA;
    // unit direction and screen basis (the state itself is not renormalised)
    RA n[3], e1[3], e2[3];
    {
        const RA nr[3] = {(RA)nd[0], (RA)nd[1], (RA)nd[2]};
        const RA rn = rnorm3<RA>(nr[0], nr[1], nr[2]);
        const RA r1 = rnorm3<RA>(ys[0], ys[1], ys[2]);
        const RA r2 = rnorm3<RA>(ys[3], ys[4], ys[5]);
        for (int i = 0; i < 3; ++i) {
            n[i] = nr[i] * rn;
            e1[i] = ys[i] * r1;
            e2[i] = ys[3 + i] * r2;
        }
    }
    const RA D = (RA)Dd, Dp = (RA)Dpd, s1 = ys[6], s2 = ys[7], eps1 = ys[8], eps2 = ys[9];

    // cells: c0 at x, c1 at x - 1/2 (staggered shift component along its own axis)
    int i0[3], i1[3];
    RI t0[3], t1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t;
        i0[a] = locate<F::kPow2>(x[a], G.n, G.inv_n, t);
        t0[a] = cell_offset<F>(x[a], t);
        if constexpr (sizeof(RI) == 8) {
            // the staggered cell follows from the unstaggered one: floor(x - 1/2)
            // is floor(x) if frac(x) >= 1/2, else floor(x) - 1; the offsets
            // t -/+ 1/2 are exact for t in [0, 1)
            const bool upper = t >= 0.5;
            const int i = upper ? i0[a] : i0[a] - 1;
            i1[a] = i < 0 ? i + G.n : i;
            t1[a] = (RI)(upper ? t - 0.5 : t + 0.5);
        } else {
            // FP32 interpolation: a second locate schedules better than the
            // derivation above (measured 5% faster for the full kernel)
            const double xs = x[a] - 0.5;
            double ts;
            i1[a] = locate<F::kPow2>(xs, G.n, G.inv_n, ts);
            t1[a] = cell_offset<F>(xs, ts);
        }
    }
    // Field f (0 potential, 1 + a shift component a) at its cell: the shift
    // component a uses the staggered cell along axis a.  Raw taps: Catmull-Rom
    // weights of the unstaggered offsets are shared by all four fields, each
    // shift component adds the set of its own staggered axis.  Bicubic
    // records: Horner factors of the x and y offsets, Catmull-Rom weights of
    // the z offset.
    auto interp = [&](int f) -> Interp<RI> {
        const int a = f - 1;
        const int ix = (a == 0) ? i1[0] : i0[0];
        const int iy = (a == 1) ? i1[1] : i0[1];
        const int iz = (a == 2) ? i1[2] : i0[2];
        if constexpr (F::kBicubic) {
            static_assert(sizeof(RI) == 4, "bicubic records are FP32");
            // each field's 16 record loads and its evaluation stay together:
            // left free, the scheduler hoists later fields' loads into earlier
            // fields and spills (472 instead of 450 ms at 512^3 / Nside 1024 /
            // 256 steps)
            __builtin_amdgcn_sched_barrier(0);
            const Powers<RI> px = powers<RI>(a == 0 ? t1[0] : t0[0]);
            const Powers<RI> py = powers<RI>(a == 1 ? t1[1] : t0[1]);
            const Axis<RI> az = cr_axis<RI>(a == 2 ? t1[2] : t0[2]);
            return bicubic_column<RI>(G.rec[f] + record_offset(G, ix, iy, iz), px, py, az);
        } else {
            // the same fence as above: FP32 taps 581 instead of 738 ms on the
            // same workload, FP64 taps (with their grouped loads) 1134 instead
            // of 1159 ms
            __builtin_amdgcn_sched_barrier(0);
            const size_t py = (size_t)G.pad, pz = (size_t)G.pad * (size_t)G.pad;
            const size_t off = (size_t)ix + (size_t)iy * py + (size_t)iz * pz;
            const Axis<RI> wx = cr_axis<RI>(a == 0 ? t1[0] : t0[0]);
            const Axis<RI> wy = cr_axis<RI>(a == 1 ? t1[1] : t0[1]);
            const Axis<RI> wz = cr_axis<RI>(a == 2 ? t1[2] : t0[2]);
            constexpr bool kGrouped = F::kWavesPerSimd == 1;  // FP64 and REF
            if constexpr (F::kGrid64)
                return tricubic<RI, kGrouped>((f == 0 ? G.phi64 : G.b64[a]) + off, py, pz, wx, wy, wz);
            else
                return tricubic<RI, kGrouped>((f == 0 ? G.phi : G.b[a]) + off, py, pz, wx, wy, wz);
        }
    };

    // potential at x (.cu:462)
    RA phi, g[3], H[3][3];
    {
        const Interp<RI> r = interp(0);
        phi = (RA)r.v;
        g[0] = (RA)r.gx;
        g[1] = (RA)r.gy;
        g[2] = (RA)r.gz;
        H[0][0] = (RA)r.hxx;
        H[0][1] = H[1][0] = (RA)r.hxy;
        H[0][2] = H[2][0] = (RA)r.hxz;
        H[1][1] = (RA)r.hyy;
        H[1][2] = H[2][1] = (RA)r.hyz;
        H[2][2] = (RA)r.hzz;
    }

    // shift: N_j = sum_a n_a dB_a/dx_j, E = 1/2 sum_a [e_a (n.gB_a) + n_a (e.gB_a)],
    // T_ij = sum_a n_a H^a_ij - 1/2 [(n H^i)_j + (n H^j)_i], L = sum_a n_a tr H^a
    RI N[3] = {RI(0), RI(0), RI(0)};
    RI T[6] = {RI(0), RI(0), RI(0), RI(0), RI(0), RI(0)};  // xx xy xz yy yz zz
    RI L = RI(0), E1 = RI(0), E2 = RI(0);
    RA Bv[3];
    // FP32 forms: with G_aj = dB_a/dx_j, E = 1/2 e . (G n + G^T n) and G^T n = N,
    // so only G n is needed beside N (the FP64 and REF forms keep the
    // reference's per-component order and rounding points)
    constexpr bool kFusedE = sizeof(RA) == 4;
    RA Gn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const Interp<RI> r = interp(1 + a);
        Bv[a] = (RA)r.v;
        const RA gb[3] = {(RA)r.gx, (RA)r.gy, (RA)r.gz};
        const RA Hb[3][3] = {{(RA)r.hxx, (RA)r.hxy, (RA)r.hxz},
                             {(RA)r.hxy, (RA)r.hyy, (RA)r.hyz},
                             {(RA)r.hxz, (RA)r.hyz, (RA)r.hzz}};
        const RA ng = n[0] * gb[0] + n[1] * gb[1] + n[2] * gb[2];
        for (int j = 0; j < 3; ++j) N[j] = (RI)((RA)N[j] + n[a] * gb[j]);
        if constexpr (kFusedE) {
            Gn[a] = ng;
        } else {
            const RA e1g = e1[0] * gb[0] + e1[1] * gb[1] + e1[2] * gb[2];
            const RA e2g = e2[0] * gb[0] + e2[1] * gb[1] + e2[2] * gb[2];
            E1 = (RI)((RA)E1 + RA(0.5) * (e1[a] * ng + n[a] * e1g));
            E2 = (RI)((RA)E2 + RA(0.5) * (e2[a] * ng + n[a] * e2g));
        }
        RA nH[3];
        for (int j = 0; j < 3; ++j) nH[j] = n[0] * Hb[0][j] + n[1] * Hb[1][j] + n[2] * Hb[2][j];
        int k = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j, ++k) {
                RA c = n[a] * Hb[i][j];
                if (i == a) c -= RA(0.5) * nH[j];
                if (j == a) c -= RA(0.5) * nH[i];
                T[k] = (RI)((RA)T[k] + c);
            }
        L = (RI)((RA)L + n[a] * (Hb[0][0] + Hb[1][1] + Hb[2][2]));
    }
    if constexpr (kFusedE) {
        const RA s0 = Gn[0] + (RA)N[0], s1 = Gn[1] + (RA)N[1], s2 = Gn[2] + (RA)N[2];
        E1 = RA(0.5) * (e1[0] * s0 + e1[1] * s1 + e1[2] * s2);
        E2 = RA(0.5) * (e2[0] * s0 + e2[1] * s1 + e2[2] * s2);
    }

    // x' = -n (1 + 2 phi) - B, with the norm of n in FP64: the position decides
    // the stencil cell, and the Catmull-Rom Hessian jumps across cell faces, so
    // the FP64 and mixed forms must track the same position to ~1e-10 cells
    {
        const double nn = nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2];
        double rnd;
        if constexpr (sizeof(RA) == 4) {
            // FP32 reciprocal square root refined by one Newton step in FP64
            // (relative error ~1e-14; |n| stays within 1e-3 of 1)
            const double r0 = (double)__frsqrt_rn((float)nn);
            rnd = r0 * (1.5 - 0.5 * nn * r0 * r0);
        } else {
            rnd = 1.0 / sqrt(nn);
        }
        const double scale = (1.0 + 2.0 * (double)phi) * rnd;
        for (int i = 0; i < 3; ++i) {
            kx[i] = -nd[i] * scale - (double)Bv[i];
            dy[i] = (RA)kx[i];
        }
    }
    // n' = -P v with P = n n^T - 1 and v = 2 grad phi + N:  n' = v - n (n . v)
    {
        const RA v[3] = {RA(2) * g[0] + (RA)N[0], RA(2) * g[1] + (RA)N[1], RA(2) * g[2] + (RA)N[2]};
        const RA nv = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
        for (int i = 0; i < 3; ++i) dy[3 + i] = v[i] - n[i] * nv;
    }
    // e' = -n [2 (e . grad phi) + E]
    {
        const RA c1 = RA(2) * (e1[0] * g[0] + e1[1] * g[1] + e1[2] * g[2]) + (RA)E1;
        const RA c2 = RA(2) * (e2[0] * g[0] + e2[1] * g[1] + e2[2] * g[2]) + (RA)E2;
        for (int i = 0; i < 3; ++i) {
            dy[6 + i] = -n[i] * c1;
            dy[9 + i] = -n[i] * c2;
        }
    }
    // Jacobi map: D'' = -4 D' Gamma - (P : H) D + L D / 2, Gamma = n.grad phi + n.N / 4
    dy[12] = -Dp;
    const RA gam = n[0] * g[0] + n[1] * g[1] + n[2] * g[2] +
                   RA(0.25) * (n[0] * (RA)N[0] + n[1] * (RA)N[1] + n[2] * (RA)N[2]);
    RA nHn;
    {
        RA u[3];
        for (int i = 0; i < 3; ++i) u[i] = H[i][0] * n[0] + H[i][1] * n[1] + H[i][2] * n[2];
        nHn = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
    }
    const RA PH = nHn - (H[0][0] + H[1][1] + H[2][2]);
    RA d13 = RA(-4) * Dp * gam - PH * D + RA(0.5) * (RA)L * D;
    // shear source uses H + T/2 (.cu:484-492)
    RA Hm[3][3];
    {
        int k = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j, ++k) {
                Hm[i][j] = (RA)(RI)(H[i][j] + RA(0.5) * (RA)T[k]);
                Hm[j][i] = Hm[i][j];
            }
    }
    RA S1H, S2H;
    {
        RA u1[3], u2[3];
        for (int i = 0; i < 3; ++i) {
            u1[i] = Hm[i][0] * e1[0] + Hm[i][1] * e1[1] + Hm[i][2] * e1[2];
            u2[i] = Hm[i][0] * e2[0] + Hm[i][1] * e2[1] + Hm[i][2] * e2[2];
        }
        S1H = (e1[0] * u1[0] + e1[1] * u1[1] + e1[2] * u1[2]) - (e2[0] * u2[0] + e2[1] * u2[1] + e2[2] * u2[2]);
        S2H = RA(2) * (e1[0] * u2[0] + e1[1] * u2[1] + e1[2] * u2[2]);
    }
    const RA D2 = D * D;
    dy[14] = RA(4) * s1 * gam + S1H * D2;
    dy[15] = RA(4) * s2 * gam + S2H * D2;
    if (Dd > 0.5) {
        const RA q = sqrt_r<RA>(RA(4) + eps1 * eps1 + eps2 * eps2);
        const RA inv_D2 = rcp_r<RA>(D2);
        if constexpr (sizeof(RA) == 4) {
            d13 += (s1 * s1 + s2 * s2) * inv_D2 * rcp_r<RA>(D);
            dy[18] = -(s1 * eps2 - s2 * eps1) * inv_D2 * rcp_r<RA>(RA(2) + q);
        } else {
            d13 += (s1 * s1 + s2 * s2) * inv_D2 / D;
            dy[18] = -(s1 * eps2 - s2 * eps1) * inv_D2 / (RA(2) + q);
        }
        dy[16] = RA(-2) * s1 * q * inv_D2;
        dy[17] = RA(-2) * s2 * q * inv_D2;
    } else {
        dy[16] = RA(0);
        dy[17] = RA(0);
        dy[18] = RA(0);
    }
    dy[13] = d13;
}

}  // namespace srt
