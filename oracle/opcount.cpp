// Floating-point operation count of the oracle's per-ray algorithm
// (TEST INFRASTRUCTURE ONLY).  The FP64 restatement (srt_oracle.hpp, which
// follows snapshot-raytracer.cu:195-550) is instantiated with a scalar type
// that counts every add/sub, multiply, divide and square root it executes.
// Negation and comparisons are not counted.  A fused multiply-add would count
// as two operations here.
//
// Three algorithms for the same ray-step are counted:
//   variant 0: the reference's own form: every tap row, and every pass of the
//     tricubic, evaluates the Catmull-Rom value, first and second derivative
//     as three separate expanded polynomials (.cu:195-312);
//   variant 1: a work-efficient form of the same interpolant: per axis the
//     Catmull-Rom weight vectors (w, w', w'') are formed once from the cell
//     offset, and the 4 x 4 x 4 taps are contracted x -> y -> z with them
//     (16 rows x 3 + 4 slices x 6 + 10 outputs = 122 four-term dot products,
//     574 flops, plus three weight sets);
//   variant 2: variant 1 with the weight sets shared between the four fields
//     of one right-hand side: B_a is sampled at x - e_a / 2, so it has the
//     potential's cell offsets on the two axes other than a, and one
//     right-hand side needs 6 weight sets instead of 12 (a set is reused when
//     the same offset recurs on an axis).
// Everything else (right-hand side, RK4) is the restatement unchanged.
// WeightsTricubic is checked against the restatement's tricubic
// (orc_tricubic_weights, tests/test_oracle_interp.py).
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace srt_oracle {
struct Cnt;
}
namespace srt_oracle {

struct OpCounts {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0;
};
static thread_local OpCounts g_ops;

struct Cnt {
    double v = 0.0;
    Cnt() = default;
    Cnt(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return float(v); }
};
inline Cnt operator+(Cnt a, Cnt b) { ++g_ops.add; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++g_ops.add; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++g_ops.mul; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++g_ops.div; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline Cnt sqrt(Cnt a) { ++g_ops.sqrt; return Cnt(std::sqrt(a.v)); }
inline Cnt floor(Cnt a) { return Cnt(std::floor(a.v)); }

}  // namespace srt_oracle

#include "srt_oracle.hpp"

namespace srt_oracle {

// Catmull-Rom weights of the taps at offsets -1, 0, 1, 2 for the cell offset
// t (the polynomials of cr_value / cr_d1 / cr_d2 collected by tap):
//   w  = ( -t + 2t^2 - t^3,  2 - 5t^2 + 3t^3,  t + 4t^2 - 3t^3,  -t^2 + t^3 ) / 2
//   w' = ( -1 + 4t - 3t^2,   -10t + 9t^2,      1 + 8t - 9t^2,    -2t + 3t^2 ) / 2
//   w''= (  2 - 3t,          -5 + 9t,          4 - 9t,           -1 + 3t )
template <class T>
struct AxisWeights {
    T w[4], d[4], dd[4];
};

template <class T>
inline AxisWeights<T> cr_weights(T t) {
    AxisWeights<T> a;
    const T h = T(0.5);
    const T t2 = t * t;
    const T ht = h * t, ht2 = h * t2, ht3 = ht2 * t;
    a.w[0] = t2 - ht - ht3;
    a.w[3] = ht3 - ht2;
    a.w[2] = ht + T(2) * t2 - T(3) * ht3;
    a.w[1] = T(1) - a.w[0] - a.w[2] - a.w[3];
    const T t32 = T(1.5) * t2;
    a.d[0] = T(2) * t - h - t32;
    a.d[3] = t32 - t;
    a.d[2] = h + T(4) * t - T(3) * t32;
    a.d[1] = -(a.d[0] + a.d[2] + a.d[3]);
    const T t3 = T(3) * t;
    a.dd[0] = T(2) - t3;
    a.dd[3] = t3 - T(1);
    a.dd[2] = T(4) - T(3) * t3;
    a.dd[1] = -(a.dd[0] + a.dd[2] + a.dd[3]);
    return a;
}

template <class T>
inline T dot4(const T w[4], const T p[4]) {
    return w[0] * p[0] + w[1] * p[1] + w[2] * p[2] + w[3] * p[3];
}

// Variant 2's weight sets, per axis: the last two offsets seen and their sets;
// a repeated offset (the same double) reuses its set, which gives the same
// weights, so no reset between right-hand sides is needed.
static thread_local bool g_share_weights = false;
struct WeightCache {
    double t[2] = {-1.0, -1.0};
    AxisWeights<Cnt> w[2];
    int next = 0;
};
static thread_local WeightCache g_wcache[3];

template <class I>
inline AxisWeights<I> axis_weights(int axis, I t) {
    if constexpr (std::is_same<I, Cnt>::value) {
        if (g_share_weights) {
            WeightCache& c = g_wcache[axis];
            for (int k = 0; k < 2; ++k)
                if (c.t[k] == double(t)) return c.w[k];
            const AxisWeights<I> w = cr_weights(t);
            c.t[c.next] = double(t);
            c.w[c.next] = w;
            c.next ^= 1;
            return w;
        }
    }
    return cr_weights(t);
}

// The oracle's interpolant (value, gradient, Hessian; periodic, the same taps
// and offsets as tricubic) by weight vectors and one x -> y -> z contraction.
struct WeightsTricubic {
    template <class I, class S>
    static Interp<I> eval(const Grid& G, S x, S y, S z) {
        using std::floor;
        const int64_t ix = int64_t(double(floor(x))), iy = int64_t(double(floor(y))), iz = int64_t(double(floor(z)));
        const AxisWeights<I> wx = axis_weights(0, I(x) - I(double(ix)));
        const AxisWeights<I> wy = axis_weights(1, I(y) - I(double(iy)));
        const AxisWeights<I> wz = axis_weights(2, I(z) - I(double(iz)));
        I v[4][4], dx[4][4], dxx[4][4];  // [c][b]
        for (int c = 0; c < 4; ++c)
            for (int b = 0; b < 4; ++b) {
                I p[4];
                for (int a = 0; a < 4; ++a) p[a] = I(G.at(ix - 1 + a, iy - 1 + b, iz - 1 + c));
                v[c][b] = dot4(wx.w, p);
                dx[c][b] = dot4(wx.d, p);
                dxx[c][b] = dot4(wx.dd, p);
            }
        I sv[4], sx[4], sy[4], sxx[4], sxy[4], syy[4];
        for (int c = 0; c < 4; ++c) {
            sv[c] = dot4(wy.w, v[c]);
            sy[c] = dot4(wy.d, v[c]);
            syy[c] = dot4(wy.dd, v[c]);
            sx[c] = dot4(wy.w, dx[c]);
            sxy[c] = dot4(wy.d, dx[c]);
            sxx[c] = dot4(wy.w, dxx[c]);
        }
        Interp<I> r;
        r.v = dot4(wz.w, sv);
        r.g[0] = dot4(wz.w, sx);
        r.g[1] = dot4(wz.w, sy);
        r.g[2] = dot4(wz.d, sv);
        r.h[0] = dot4(wz.w, sxx);
        r.h[1] = dot4(wz.w, sxy);
        r.h[2] = dot4(wz.d, sx);
        r.h[3] = dot4(wz.w, syy);
        r.h[4] = dot4(wz.d, sy);
        r.h[5] = dot4(wz.dd, sv);
        return r;
    }
};

}  // namespace srt_oracle

using namespace srt_oracle;

namespace {

Fields f32_fields(const float* phi, const float* bx, const float* by, const float* bz, int64_t nx, int64_t ny,
                  int64_t nz) {
    Fields F;
    const float* g[4] = {phi, bx, by, bz};
    Grid* dst[4] = {&F.phi, &F.b[0], &F.b[1], &F.b[2]};
    for (int a = 0; a < 4; ++a) {
        dst[a]->f32 = g[a];
        dst[a]->nx = nx;
        dst[a]->ny = ny;
        dst[a]->nz = nz;
    }
    return F;
}

}  // namespace

extern "C" {

// Runs n_steps RK4 steps of the ray through RING pixel `pix` with the FP64
// form and returns the operation counts per ray-step in counts[0..3]
// (add/sub, mul, div, sqrt).  variant 0: the reference's interpolation form;
// 1: the weights form (WeightsTricubic); 2: the weights form with weight sets
// shared between the fields of a right-hand side.
int orc_count_ops(const float* phi, const float* bx, const float* by, const float* bz, int64_t nx, int64_t ny,
                  int64_t nz, int64_t nside, int64_t pix, int n_steps, const double* obs, double* counts,
                  int variant) {
    if (n_steps <= 0 || variant < 0 || variant > 2) return 1;
    const Fields F = f32_fields(phi, bx, by, bz, nx, ny, nz);
    double z, s, ph;
    pix2zphi_ring(nside, pix, z, s, ph);
    Cnt y[19];
    initial_state<Cnt>(obs, z, s, ph, y);
    g_ops = OpCounts();
    g_share_weights = variant == 2;
    for (WeightCache& c : g_wcache) c = WeightCache();
    if (variant == 0)
        rk4_integrate<Cnt, Cnt>(F, y, n_steps);
    else
        rk4_integrate<Cnt, Cnt, WeightsTricubic>(F, y, n_steps);
    g_share_weights = false;
    counts[0] = double(g_ops.add) / n_steps;
    counts[1] = double(g_ops.mul) / n_steps;
    counts[2] = double(g_ops.div) / n_steps;
    counts[3] = double(g_ops.sqrt) / n_steps;
    return 0;
}

// The weights-form interpolation in FP64 at n positions: out n x 10 = value,
// grad[3], Hessian (xx, xy, xz, yy, yz, zz), like orc_tricubic's FP64 form.
int orc_tricubic_weights(const float* grid, int64_t nx, int64_t ny, int64_t nz, const double* pos, int64_t n,
                         double* out) {
    Grid G;
    G.f32 = grid;
    G.nx = nx;
    G.ny = ny;
    G.nz = nz;
    for (int64_t i = 0; i < n; ++i) {
        const Interp<double> r =
            WeightsTricubic::eval<double, double>(G, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        double* o = out + 10 * i;
        o[0] = r.v;
        for (int k = 0; k < 3; ++k) o[1 + k] = r.g[k];
        for (int k = 0; k < 6; ++k) o[4 + k] = r.h[k];
    }
    return 0;
}

}  // extern "C"
