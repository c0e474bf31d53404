// snapshot-raytracer command line: the drop-in for the reference's program.
//
//   snapshot-raytracer <hdf5_file_potential> <hdf5_file_B> <Nside> <n_steps> [<batch>]
//
// Same arguments, validation, batch semantics, output files and exit codes as
// snapshot-raytracer.cu:600-1006 (README.md:27-49):
//   * reads dataset /data of both files (potential: scalar; shift: array[3] of
//     float per voxel);
//   * batch b traces RING pixels [b Nside^2, (b+1) Nside^2), no batch = full sky
//     (.cu:708-717);
//   * writes dAmap, ellipticity1map, ellipticity2map, rotationmap,
//     deflection1map, deflection2map as <name>_batch<b>.fits in the working
//     directory, full-sky RING maps with zeros outside the batch (.cu:924-987);
//   * returns 0 on success and -1 (exit status 255) on handled errors.
// Differences, all on invalid input: non-numeric Nside/n_steps/batch is an
// error (-1) rather than an uncaught exception; grids must be cubic with equal
// potential and shift dimensions (the reference assumes it, .cu:645, 676).
//
// Environment (optional): SRT_PRECISION = mixed (default) | fp64 | ref selects
// the kernel arithmetic; SRT_SHIFT = f32 (default) | f16 selects the device
// shift storage (f16 is the reference's binary16 x 2^27); SRT_LAYOUT = auto
// (default) | bicubic | taps selects the device layout the mixed kernel reads
// (auto: bicubic cell records if they fit in device memory, else the padded
// cubes); SRT_DEVICE = HIP device ordinal.
#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/srt.h"
#include "srt_io.hpp"

namespace {

bool parse_int64(const char* s, int64_t& v) {
    errno = 0;
    char* end = nullptr;
    const long long r = std::strtoll(s, &end, 10);
    if (errno || end == s || *end != '\0') return false;
    v = r;
    return true;
}

const char* env_or(const char* name, const char* dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? v : dflt;
}

void usage(const char* prog) {
    std::cerr << "Usage: " << prog << " <hdf5_file_potential> <hdf5_file_B> <Nside> <n_steps> [<batch>]\n"
              << "  <hdf5_file_potential>: HDF5 file containing a snapshot of the Weyl potential\n"
              << "  <hdf5_file_B>: HDF5 file containing a snapshot of the frame-dragging potential (shift)\n"
              << "  <Nside>: HEALPix Nside parameter for output maps\n"
              << "  <n_steps>: Number of integration steps for raytracing (1 step = 1 grid unit)\n"
              << "  [<batch>]: Optional batch index for processing a subset of pixels" << std::endl;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 5) {
        usage(argv[0]);
        return -1;
    }
    int64_t nside = 0, n_steps = 0, batch = -1;
    if (!parse_int64(argv[3], nside) || nside <= 0) {
        std::cerr << "Invalid Nside value: " << argv[3] << std::endl;
        return -1;
    }
    if (!parse_int64(argv[4], n_steps) || n_steps <= 0 || n_steps > 0x7fffffff) {
        std::cerr << "Invalid n_steps value: " << argv[4] << std::endl;
        return -1;
    }
    if (argc > 5) {
        if (!parse_int64(argv[5], batch) || batch < 0) {
            std::cerr << "Invalid batch value: " << argv[5] << std::endl;
            return -1;
        }
    }
    const std::string prec_s = env_or("SRT_PRECISION", "mixed");
    int precision = SRT_PREC_MIXED;
    if (prec_s == "fp64")
        precision = SRT_PREC_FP64;
    else if (prec_s == "ref")
        precision = SRT_PREC_REF;
    else if (prec_s != "mixed") {
        std::cerr << "Invalid SRT_PRECISION: " << prec_s << " (mixed, fp64 or ref)" << std::endl;
        return -1;
    }
    const std::string shift_s = env_or("SRT_SHIFT", "f32");
    if (shift_s != "f32" && shift_s != "f16") {
        std::cerr << "Invalid SRT_SHIFT: " << shift_s << " (f32 or f16)" << std::endl;
        return -1;
    }
    const std::string layout_s = env_or("SRT_LAYOUT", "auto");
    if (layout_s != "auto" && layout_s != "bicubic" && layout_s != "taps") {
        std::cerr << "Invalid SRT_LAYOUT: " << layout_s << " (auto, bicubic or taps)" << std::endl;
        return -1;
    }
    int64_t device = 0;
    if (!parse_int64(env_or("SRT_DEVICE", "0"), device) || device < 0) {
        std::cerr << "Invalid SRT_DEVICE" << std::endl;
        return -1;
    }

    std::cout << "Nside: " << nside << std::endl;
    std::cout << "n_steps: " << n_steps << std::endl;
    // wall time of each phase, printed before the final line
    using clock = std::chrono::steady_clock;
    const auto t_begin = clock::now();
    auto seconds = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    // The HIP runtime and the library context start on a second thread while
    // the snapshot is read: runtime start-up takes a few tenths of a second,
    // about as long as reading a 512^3 snapshot.
    srt_context* ctx = nullptr;
    int create_rc = SRT_OK;
    std::string create_err;  // srt_last_error is per thread
    auto start_device = [&] {
        create_rc = srt_create((int)device, &ctx);
        if (create_rc != SRT_OK) create_err = srt_last_error();
    };
    std::thread starter;
    try {
        starter = std::thread(start_device);
    } catch (...) {
        start_device();
    }
    auto stop_device = [&] {
        if (starter.joinable()) starter.join();
        srt_destroy(ctx);
        ctx = nullptr;
    };
    // the readers fill buffers that start uninitialised (no pass that zeroes
    // gigabytes first); a failed allocation leaves the buffer null
    auto sink_into = [](std::unique_ptr<float[]>& buf) {
        return srt_io::FloatSink([&buf](uint64_t count) -> float* {
            buf.reset(new (std::nothrow) float[count]);
            return buf.get();
        });
    };
    std::cout << "Loading potential data from HDF5 file " << argv[1] << " ..." << std::endl;
    std::string err;
    std::unique_ptr<float[]> phi, shift;
    std::vector<uint64_t> dims;
    const bool read_phi = srt_io::read_scalar_field(argv[1], "/data", sink_into(phi), dims, err);
    if (!read_phi || !phi) {
        std::cerr << (read_phi ? "Out of host memory for the potential" : err) << std::endl;
        stop_device();
        return -1;
    }
    std::cout << "Data dimensions: " << dims[0] << " x " << dims[1] << " x " << dims[2] << std::endl;
    std::cout << "Loading B field data from HDF5 file " << argv[2] << " ..." << std::endl;
    std::vector<uint64_t> bdims;
    const bool read_shift = srt_io::read_vector_field(argv[2], "/data", sink_into(shift), bdims, err);
    if (!read_shift || !shift) {
        std::cerr << (read_shift ? "Out of host memory for the B field" : err) << std::endl;
        stop_device();
        return -1;
    }
    if (bdims != dims) {
        std::cerr << "Potential and B field dimensions differ" << std::endl;
        stop_device();
        return -1;
    }

    const int64_t npix = 12 * nside * nside;
    int64_t istart = 0, count = npix;
    if (batch >= 0) {
        istart = batch * nside * nside;
        count = nside * nside;
        if (istart >= npix) {
            std::cerr << "Batch exceeds total number of pixels" << std::endl;
            stop_device();
            return -1;
        }
    }

    const auto t_read = clock::now();
    std::cout << "Allocating device memory and copying data to device..." << std::endl;
    if (starter.joinable()) starter.join();
    int rc = create_rc;
    if (rc != SRT_OK) {
        std::cerr << "Device setup failed: " << srt_strerror(rc) << ": " << create_err << std::endl;
        stop_device();
        return -1;
    }
    rc = srt_upload(ctx, phi.get(), shift.get(), (int64_t)dims[2], (int64_t)dims[1], (int64_t)dims[0], 0,
                    shift_s == "f16" ? SRT_SHIFT_F16_SCALED : SRT_SHIFT_F32);
    if (rc != SRT_OK) {
        std::cerr << "Device setup failed: " << srt_strerror(rc) << ": " << srt_last_error() << std::endl;
        stop_device();
        return -1;
    }
    phi.reset();
    shift.reset();
    if (precision == SRT_PREC_MIXED && layout_s != "taps") {
        rc = srt_set_layout(ctx, SRT_LAYOUT_BICUBIC, 16);
        if (rc == SRT_ERR_OUT_OF_MEMORY && layout_s == "auto") {
            std::cout << "Bicubic cell records do not fit in device memory; tracing from the padded grids."
                      << std::endl;
        } else if (rc != SRT_OK) {
            std::cerr << "Device setup failed: " << srt_strerror(rc) << ": " << srt_last_error() << std::endl;
            srt_destroy(ctx);
            return -1;
        }
    }

    // compact sky patches per wavefront (a batch's rays are sorted on the device,
    // up to 2^31 - 1 of them; beyond that the library keeps LINEAR order); the
    // maps do not depend on the lane order
    const int order = SRT_ORDER_NESTED;
    int lane_order = order;
    (void)srt_effective_order(nside, istart, count, 0, order, &lane_order);
    std::cout << "Launching raytracing kernel (" << prec_s << ", "
              << (lane_order == SRT_ORDER_NESTED ? "nested" : "linear") << " lane order)..." << std::endl;
    std::vector<float> maps((size_t)(6 * count));
    const auto t0 = clock::now();
    rc = srt_trace_host(ctx, nside, istart, count, nullptr, (int)n_steps, nullptr, precision, order, maps.data(),
                        nullptr);
    const auto t_trace = clock::now();
    const double sec = seconds(t0, t_trace);
    float kernel_ms = 0.0f;
    if (rc == SRT_OK) rc = srt_last_trace_ms(ctx, &kernel_ms);
    srt_destroy(ctx);
    if (rc != SRT_OK) {
        std::cerr << "Raytracing failed: " << srt_strerror(rc) << ": " << srt_last_error() << std::endl;
        return -1;
    }
    std::cout << "Raytracing completed: " << count << " rays x " << n_steps << " steps, kernel " << kernel_ms
              << " ms (" << (double)count * (double)n_steps / (kernel_ms * 1e-3) << " ray-steps/s), "
              << sec << " s with transfers." << std::endl;

    std::cout << "Saving results to FITS files..." << std::endl;
    const char* names[6] = {"dAmap", "ellipticity1map", "ellipticity2map",
                            "rotationmap", "deflection1map", "deflection2map"};
    const int64_t file_batch = batch < 0 ? 0 : batch;
    // the six files are independent: one thread writes each (a map whose
    // thread cannot be started is written by this one afterwards)
    std::string errs[6];
    bool written[6] = {};
    auto write_map = [&](int m) {
        try {
            std::vector<float> outmap((size_t)npix, 0.0f);
            std::memcpy(outmap.data() + istart, maps.data() + (size_t)m * count, (size_t)count * sizeof(float));
            char fname[256];
            std::snprintf(fname, sizeof(fname), "%s_batch%lld.fits", names[m], (long long)file_batch);
            written[m] = srt_io::write_healpix_map(fname, outmap.data(), nside, false, "C", errs[m]);
        } catch (const std::exception& e) {
            errs[m] = std::string("writing ") + names[m] + ": " + e.what();
        }
    };
    std::vector<std::thread> writers;
    std::vector<int> inline_maps;
    for (int m = 0; m < 6; ++m) {
        try {
            writers.emplace_back(write_map, m);
        } catch (...) {
            inline_maps.push_back(m);
        }
    }
    for (int m : inline_maps) write_map(m);
    for (auto& t : writers) t.join();
    for (int m = 0; m < 6; ++m) {
        if (!written[m]) {
            std::cerr << errs[m] << std::endl;
            return -1;
        }
    }
    const auto t_end = clock::now();
    std::cout << "Timing: HDF5 read " << seconds(t_begin, t_read) << " s, device setup " << seconds(t_read, t0)
              << " s, trace with map download " << sec << " s, FITS write " << seconds(t_trace, t_end)
              << " s, total " << seconds(t_begin, t_end) << " s" << std::endl;
    std::cout << "All done!" << std::endl;
    return 0;
}
