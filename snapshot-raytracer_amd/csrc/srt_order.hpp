// Lane order of partial-sky traces (srt_order.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srt {

// Largest ray count a sorted lane order covers (slots are 32-bit, and the
// radix sort takes an int count).
constexpr int64_t kMaxOrderedRays = (int64_t(1) << 31) - 1;

// Device work bytes order_lanes needs for `count` rays at `nside` (0 if
// count is out of range).
size_t order_work_bytes(int64_t nside, int64_t count);

// lane_slot[lane] = the ray slot lane `lane` traces: slots i in [0, count)
// (pixel pix_list[i], or pix_begin + i if pix_list is null) sorted by
// srt::ring_order_key of their pixel; invalid pixels go last.  Asynchronous
// on `stream`; `work` holds order_work_bytes(nside, count) bytes.
hipError_t order_lanes(int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                       uint32_t* lane_slot, void* work, size_t work_bytes, hipStream_t stream);

}  // namespace srt
