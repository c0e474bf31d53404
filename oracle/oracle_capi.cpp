// C entry points of the host oracle (TEST INFRASTRUCTURE ONLY: loaded by the
// tests, the smoke check and the benchmark's CPU-baseline leg through ctypes;
// never linked into the product library).
//
// form: 0 = FP64 oracle, 1 = REF (the reference GPU code's float/double mix).
// Shift grids: three planar float grids (bx, by, bz), or, when b16 is not
// null, three planar binary16 grids of B * 2^27 packed back to back (the
// reference's device storage, .cu:585-597).
#include <chrono>
#include <cstdint>
#include <immintrin.h>
#include <omp.h>

#include "srt_oracle.hpp"

using namespace srt_oracle;

namespace {

Grid make_grid(const float* f32, const uint16_t* f16, int64_t nx, int64_t ny, int64_t nz) {
    Grid g;
    g.f32 = f32;
    g.f16 = f16;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    return g;
}

Fields make_fields(const float* phi, const float* bx, const float* by, const float* bz, const uint16_t* b16,
                   int64_t nx, int64_t ny, int64_t nz) {
    Fields F;
    F.phi = make_grid(phi, nullptr, nx, ny, nz);
    const int64_t nvox = nx * ny * nz;
    const float* b32[3] = {bx, by, bz};
    for (int a = 0; a < 3; ++a)
        F.b[a] = b16 ? make_grid(nullptr, b16 + a * nvox, nx, ny, nz) : make_grid(b32[a], nullptr, nx, ny, nz);
    return F;
}

}  // namespace

extern "C" {

int orc_pix2zphi_ring(int64_t nside, const int64_t* pix, int64_t n, double* z, double* s, double* phi) {
    if (nside <= 0) return 1;
    for (int64_t i = 0; i < n; ++i) {
        if (pix[i] < 0 || pix[i] >= 12 * nside * nside) return 1;
        pix2zphi_ring(nside, pix[i], z[i], s[i], phi[i]);
    }
    return 0;
}

// out: n x 10 = value, grad[3], hessian (xx, xy, xz, yy, yz, zz)
int orc_tricubic(int form, const float* grid, const uint16_t* grid16, int64_t nx, int64_t ny, int64_t nz,
                 const double* pos, int64_t n, double* out) {
    const Grid G = make_grid(grid16 ? nullptr : grid, grid16, nx, ny, nz);
    for (int64_t i = 0; i < n; ++i) {
        double* o = out + 10 * i;
        if (form == 0) {
            const Interp<double> r = tricubic<double, double>(G, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            o[0] = r.v;
            for (int k = 0; k < 3; ++k) o[1 + k] = r.g[k];
            for (int k = 0; k < 6; ++k) o[4 + k] = r.h[k];
        } else {
            const Interp<float> r = tricubic<float, double>(G, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            o[0] = r.v;
            for (int k = 0; k < 3; ++k) o[1 + k] = r.g[k];
            for (int k = 0; k < 6; ++k) o[4 + k] = r.h[k];
        }
    }
    return 0;
}

int orc_rhs(int form, const float* phi, const float* bx, const float* by, const float* bz, const uint16_t* b16,
            int64_t nx, int64_t ny, int64_t nz, const double* y, int64_t n, double* dydt) {
    const Fields F = make_fields(phi, bx, by, bz, b16, nx, ny, nz);
    for (int64_t i = 0; i < n; ++i) {
        if (form == 0)
            rhs<double, double>(F, y + 19 * i, dydt + 19 * i);
        else
            rhs<double, float>(F, y + 19 * i, dydt + 19 * i);
    }
    return 0;
}

// One RK4 run per pixel.  out is map-major: out[m * npix + i], m = dA, eps1,
// eps2, rotation, deflection1, deflection2.  state_out (npix x 19) may be null.
// Returns the wall time of the ray loop in *seconds.
int orc_trace(int form, const float* phi, const float* bx, const float* by, const float* bz, const uint16_t* b16,
              int64_t nx, int64_t ny, int64_t nz, int64_t nside, const int64_t* pix, int64_t npix, int n_steps,
              const double* obs, float* out, double* state_out, int nthreads, double* seconds) {
    if (nside <= 0 || n_steps < 0 || npix < 0) return 1;
    for (int64_t i = 0; i < npix; ++i)
        if (pix[i] < 0 || pix[i] >= 12 * nside * nside) return 1;
    const Fields F = make_fields(phi, bx, by, bz, b16, nx, ny, nz);
    const double o[3] = {obs[0], obs[1], obs[2]};
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < npix; ++i) {
        float m[6];
        double* st = state_out ? state_out + 19 * i : nullptr;
        if (form == 0)
            trace_ray<double>(F, o, nside, pix[i], n_steps, m, st);
        else
            trace_ray<float>(F, o, nside, pix[i], n_steps, m, st);
        for (int k = 0; k < 6; ++k) out[k * npix + i] = m[k];
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

int orc_max_threads() { return omp_get_max_threads(); }

void orc_f32_to_f16_array(const float* in, uint16_t* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = f32_to_f16_rne(in[i]);
}

void orc_f16_to_f32_array(const uint16_t* in, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = f16_to_f32(in[i]);
}

// Exhaustive comparison of f32_to_f16_rne with the x86 F16C conversion
// (vcvtps2ph, round-to-nearest-even) over [lo, hi) of the 2^32 float bit
// patterns.  Returns the number of mismatching patterns.
__attribute__((target("f16c"))) uint64_t orc_f16_check_range(uint64_t lo, uint64_t hi) {
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t b = int64_t(lo); b < int64_t(hi); ++b) {
        const uint32_t bits = uint32_t(b);
        float f;
        std::memcpy(&f, &bits, 4);
        const uint16_t mine = f32_to_f16_rne(f);
        const uint16_t hw = uint16_t(_cvtss_sh(f, _MM_FROUND_TO_NEAREST_INT));
        if (mine != hw) ++bad;
    }
    return bad;
}

}  // extern "C"
