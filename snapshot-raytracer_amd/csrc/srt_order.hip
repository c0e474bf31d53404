// Lane order of partial-sky traces: the rays of a RING pixel range or of a
// caller's pixel list are sorted by srt::ring_order_key (base face, then the
// bit-interleaved in-face coordinates), so each 64-lane wavefront traces a
// compact patch of sky and its rays read neighbouring grid cells.  The
// reference traces RING pixels in index order, one thread each
// (snapshot-raytracer.cu:556-557); the order changes which lane traces which
// ray, never a ray's result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "srt_device.hpp"
#include "srt_order.hpp"

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// bits of the largest key: 12 faces above two interleaved L-bit coordinates
int key_bits(int64_t nside) {
    int L = 0;
    while ((int64_t(1) << L) < nside) ++L;
    return 2 * L + 4;
}

__global__ __launch_bounds__(256) void order_keys_kernel(int64_t nside, int64_t pix_begin, int64_t count,
                                                         const int64_t* __restrict__ pix_list,
                                                         uint64_t* __restrict__ keys, uint32_t* __restrict__ slots,
                                                         uint64_t invalid_key) {
    const int64_t npix = 12 * nside * nside;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pix = pix_list ? pix_list[i] : pix_begin + i;
        keys[i] = (pix >= 0 && pix < npix) ? (uint64_t)srt::ring_order_key(nside, pix) : invalid_key;
        slots[i] = (uint32_t)i;
    }
}

struct Parts {
    size_t keys_in, keys_out, slots_in, temp, temp_bytes, total;
};

hipError_t parts(int64_t nside, int64_t count, Parts& p) {
    p.temp_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, p.temp_bytes, (const uint64_t*)nullptr,
                                                      (uint64_t*)nullptr, (const uint32_t*)nullptr,
                                                      (uint32_t*)nullptr, (int)count, 0, key_bits(nside));
    p.keys_in = 0;
    p.keys_out = align_up(p.keys_in + (size_t)count * sizeof(uint64_t));
    p.slots_in = align_up(p.keys_out + (size_t)count * sizeof(uint64_t));
    p.temp = align_up(p.slots_in + (size_t)count * sizeof(uint32_t));
    p.total = align_up(p.temp + p.temp_bytes);
    return e;
}

}  // namespace

namespace srt {

size_t order_work_bytes(int64_t nside, int64_t count) {
    Parts p;
    if (count <= 0 || count > kMaxOrderedRays || parts(nside, count, p) != hipSuccess) return 0;
    return p.total;
}

hipError_t order_lanes(int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                       uint32_t* lane_slot, void* work, size_t work_bytes, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    if (count > kMaxOrderedRays) return hipErrorInvalidValue;
    Parts p;
    hipError_t e = parts(nside, count, p);
    if (e != hipSuccess) return e;
    if (work_bytes < p.total) return hipErrorInvalidValue;
    char* w = static_cast<char*>(work);
    uint64_t* keys_in = reinterpret_cast<uint64_t*>(w + p.keys_in);
    uint64_t* keys_out = reinterpret_cast<uint64_t*>(w + p.keys_out);
    uint32_t* slots_in = reinterpret_cast<uint32_t*>(w + p.slots_in);
    const int bits = key_bits(nside);
    const uint64_t invalid = uint64_t(12) << (bits - 4);  // above every valid key, below 2^bits
    const int blocks = (int)std::min<int64_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(order_keys_kernel, dim3(blocks), dim3(256), 0, stream, nside, pix_begin, count, pix_list,
                       keys_in, slots_in, invalid);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t temp_bytes = p.temp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(w + p.temp, temp_bytes, keys_in, keys_out, slots_in, lane_slot,
                                              (int)count, 0, bits, stream);
}

}  // namespace srt
