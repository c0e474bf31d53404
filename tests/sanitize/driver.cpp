// Host-side checks under AddressSanitizer / UndefinedBehaviorSanitizer (and,
// for the I/O part alone, ThreadSanitizer): built and run by
// tests/test_sanitize_host.py, which writes the HDF5 inputs first.
//
//   driver <dir> io       the HDF5 readers (raw multi-threaded path and the
//                         H5Dread fallbacks) and the FITS writer
//   driver <dir> oracle   the host oracle's C API (FP64 and REF traces,
//                         tricubic, right-hand side, binary16 conversion)
//
// Files in <dir>: seq.h5 (float potential, value i at flat index i), seq3.h5
// (array[3] of float, value i at flat float index i), seqd.h5 (the potential as
// double), seqc.h5 (chunked), sequb.h5 (with a user block); all n^3 = 160^3.
// Exit status 0 when every check holds; a sanitizer report aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/srt_io.h"

#ifndef SANITIZE_IO_ONLY
extern "C" {
int orc_trace(int form, const float* phi, const float* bx, const float* by, const float* bz, const uint16_t* b16,
              int64_t nx, int64_t ny, int64_t nz, int64_t nside, const int64_t* pix, int64_t npix, int n_steps,
              const double* obs, float* out, double* state_out, int nthreads, double* seconds);
int orc_tricubic(int form, const float* grid, const uint16_t* grid16, int64_t nx, int64_t ny, int64_t nz,
                 const double* pos, int64_t n, double* out);
int orc_rhs(int form, const float* phi, const float* bx, const float* by, const float* bz, const uint16_t* b16,
            int64_t nx, int64_t ny, int64_t nz, const double* y, int64_t n, double* dydt);
void orc_f32_to_f16_array(const float* in, uint16_t* out, int64_t n);
void orc_f16_to_f32_array(const uint16_t* in, float* out, int64_t n);
}
#endif

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what.c_str());
        ++failures;
    }
}

// The sequence file's floats are 0, 1, 2, ... (exact in binary32 below 2^24).
bool is_sequence(const std::vector<float>& v) {
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i] != (float)i) return false;
    return true;
}

void io_checks(const std::string& dir) {
    char err[256];
    for (const char* name : {"seq.h5", "seqd.h5", "seqc.h5", "sequb.h5"}) {
        const std::string path = dir + "/" + name;
        uint64_t dims[3] = {0, 0, 0};
        check(srtio_read_scalar_field(path.c_str(), "/data", nullptr, 0, dims, err, sizeof err) == 0,
              std::string("dims of ") + name);
        const uint64_t n = dims[0] * dims[1] * dims[2];
        std::vector<float> v(n, -1.0f);
        check(srtio_read_scalar_field(path.c_str(), "/data", v.data(), n, dims, err, sizeof err) == 0,
              std::string("read ") + name);
        check(is_sequence(v), std::string("values of ") + name);
        std::vector<float> small(n - 1, -1.0f);
        check(srtio_read_scalar_field(path.c_str(), "/data", small.data(), n - 1, dims, err, sizeof err) == 1 &&
                  std::string(err) == "output buffer too small",
              std::string("small buffer refused for ") + name);
    }
    {
        const std::string path = dir + "/seq3.h5";
        uint64_t dims[3] = {0, 0, 0};
        check(srtio_read_vector_field(path.c_str(), "/data", nullptr, 0, dims, err, sizeof err) == 0, "dims of seq3");
        std::vector<float> v(3 * dims[0] * dims[1] * dims[2], -1.0f);
        check(srtio_read_vector_field(path.c_str(), "/data", v.data(), v.size(), dims, err, sizeof err) == 0,
              "read seq3");
        check(is_sequence(v), "values of seq3");
        check(srtio_read_vector_field((dir + "/seq.h5").c_str(), "/data", v.data(), v.size(), dims, err,
                                      sizeof err) == 1,
              "scalar dataset refused as a vector field");
    }
    check(srtio_read_scalar_field((dir + "/missing.h5").c_str(), "/data", nullptr, 0, nullptr, err, sizeof err) == 1,
          "null dims refused");
    uint64_t dims[3];
    check(srtio_read_scalar_field((dir + "/missing.h5").c_str(), "/data", nullptr, 0, dims, err, sizeof err) == 1,
          "missing file refused");
    // a map written to FITS: a whole number of 2880-byte blocks
    const int64_t nside = 16, npix = 12 * nside * nside;
    std::vector<float> map(npix);
    for (int64_t i = 0; i < npix; ++i) map[i] = (float)std::sin(0.01 * (double)i);
    const std::string fits = dir + "/m.fits";
    check(srtio_write_healpix_map(fits.c_str(), map.data(), nside, 0, "C", err, sizeof err) == 0, "FITS write");
    FILE* f = std::fopen(fits.c_str(), "rb");
    check(f != nullptr, "FITS file exists");
    if (f) {
        std::fseek(f, 0, SEEK_END);
        check(std::ftell(f) % 2880 == 0, "FITS size");
        std::fclose(f);
    }
}

#ifndef SANITIZE_IO_ONLY
void oracle_checks() {
    const int64_t n = 12;
    const size_t vox = (size_t)(n * n * n);
    std::vector<float> phi(vox), b[3] = {std::vector<float>(vox), std::vector<float>(vox), std::vector<float>(vox)};
    for (size_t i = 0; i < vox; ++i) {
        phi[i] = 1e-5f * (float)std::sin(0.37 * (double)i);
        for (int a = 0; a < 3; ++a) b[a][i] = 1e-7f * (float)std::cos(0.23 * (double)i + a);
    }
    std::vector<uint16_t> b16(3 * vox);
    std::vector<float> scaled(3 * vox);
    for (int a = 0; a < 3; ++a)
        for (size_t i = 0; i < vox; ++i) scaled[a * vox + i] = b[a][i] * 134217728.0f;
    orc_f32_to_f16_array(scaled.data(), b16.data(), (int64_t)scaled.size());
    std::vector<float> back(3 * vox);
    orc_f16_to_f32_array(b16.data(), back.data(), (int64_t)back.size());
    for (size_t i = 0; i < back.size(); ++i)
        if (!(std::fabs(back[i] - scaled[i]) <= 1e-3f * std::fabs(scaled[i]) + 1e-7f)) {
            check(false, "binary16 round trip");
            break;
        }
    const int64_t nside = 2, npix = 48;
    std::vector<int64_t> pix(npix);
    for (int64_t i = 0; i < npix; ++i) pix[i] = i;
    const double obs[3] = {6.0, 6.0, 6.0}, obs2[3] = {-3.25, 15.5, 7.75};
    for (int form = 0; form < 2; ++form)
        for (const double* o : {obs, obs2}) {
            std::vector<float> out(6 * npix);
            std::vector<double> state(19 * npix);
            double sec = 0.0;
            const uint16_t* codes = form == 1 ? b16.data() : nullptr;
            check(orc_trace(form, phi.data(), b[0].data(), b[1].data(), b[2].data(), codes, n, n, n, nside, pix.data(),
                            npix, 30, o, out.data(), state.data(), 2, &sec) == 0,
                  "oracle trace");
            bool finite = true;
            for (float x : out) finite = finite && std::isfinite(x);
            check(finite, "oracle maps finite");
        }
    std::vector<double> pos(3 * 50), res(10 * 50);
    for (int i = 0; i < 150; ++i) pos[i] = -30.0 + 0.77 * i;  // across the periodic wrap, negative included
    for (int form = 0; form < 2; ++form)
        check(orc_tricubic(form, phi.data(), nullptr, n, n, n, pos.data(), 50, res.data()) == 0, "oracle tricubic");
    std::vector<double> y(19 * 4, 0.0), dy(19 * 4);
    for (int r = 0; r < 4; ++r) {
        y[19 * r + 0] = 1.5 * r;
        y[19 * r + 1] = 11.7 - r;
        y[19 * r + 2] = 0.25 * r;
        y[19 * r + 5] = -1.0;
        y[19 * r + 6] = 1.0;
        y[19 * r + 10] = 1.0;
        y[19 * r + 12] = 0.6 * r;
        y[19 * r + 13] = -1.0;
    }
    for (int form = 0; form < 2; ++form)
        check(orc_rhs(form, phi.data(), b[0].data(), b[1].data(), b[2].data(), nullptr, n, n, n, y.data(), 4,
                      dy.data()) == 0,
              "oracle right-hand side");
}
#endif

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: driver <dir> io|oracle\n");
        return 2;
    }
    const std::string dir = argv[1], what = argv[2];
    if (what == "io") io_checks(dir);
#ifndef SANITIZE_IO_ONLY
    else if (what == "oracle")
        oracle_checks();
#endif
    else
        return 2;
    std::printf("%s checks: %d failures\n", what.c_str(), failures);
    return failures ? 1 : 0;
}
