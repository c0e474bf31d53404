// Floating-point operation count of the oracle's per-ray algorithm
// (TEST INFRASTRUCTURE ONLY).  The FP64 restatement (srt_oracle.hpp, which
// follows snapshot-raytracer.cu:195-550) is instantiated with a scalar type
// that counts every add/sub, multiply, divide and square root it executes.
// Negation and comparisons are not counted.  A fused multiply-add would count
// as two operations here.
#include <cmath>
#include <cstdint>

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

using namespace srt_oracle;

extern "C" {

// Runs n_steps RK4 steps of the ray through RING pixel `pix` with the FP64
// form and returns the operation counts per ray-step in counts[0..3]
// (add/sub, mul, div, sqrt).
int orc_count_ops(const float* phi, const float* bx, const float* by, const float* bz, int64_t nx, int64_t ny,
                  int64_t nz, int64_t nside, int64_t pix, int n_steps, const double* obs, double* counts) {
    if (n_steps <= 0) return 1;
    Fields F;
    const float* g[4] = {phi, bx, by, bz};
    Grid* dst[4] = {&F.phi, &F.b[0], &F.b[1], &F.b[2]};
    for (int a = 0; a < 4; ++a) {
        dst[a]->f32 = g[a];
        dst[a]->nx = nx;
        dst[a]->ny = ny;
        dst[a]->nz = nz;
    }
    double z, s, ph;
    pix2zphi_ring(nside, pix, z, s, ph);
    Cnt y[19];
    initial_state<Cnt>(obs, z, s, ph, y);
    g_ops = OpCounts();
    rk4_integrate<Cnt, Cnt>(F, y, n_steps);
    counts[0] = double(g_ops.add) / n_steps;
    counts[1] = double(g_ops.mul) / n_steps;
    counts[2] = double(g_ops.div) / n_steps;
    counts[3] = double(g_ops.sqrt) / n_steps;
    return 0;
}

}  // extern "C"
