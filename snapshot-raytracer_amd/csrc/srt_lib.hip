// libsrt.so: HIP kernels of the snapshot ray tracer for MI355X (gfx950) and
// the C ABI declared in include/srt.h.
//
// Kernels
//   pad_grid_kernel   dense n^3 (scalar or one component of an interleaved
//                     vector field) -> padded (n+3)^3 periodic cube; optional
//                     binary16(B * 2^27) quantisation, the reference's shift
//                     storage (.cu:585-597, read back at .cu:241-244).
//   trace_kernel<F>   one ray per lane: device HEALPix direction, RK4 over
//                     n_steps with 4 right-hand-side evaluations per step,
//                     each of 4 tricubic interpolations, then the six maps
//                     (.cu:552-583).  F selects FP64, mixed or reference
//                     arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/srt.h"
#include "srt_adjoint.hpp"
#include "srt_device.hpp"
#include "srt_order.hpp"

namespace {

constexpr int kBlock = 256;

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define SRT_HIP(call)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            return fail(e_ == hipErrorOutOfMemory ? SRT_ERR_OUT_OF_MEMORY : SRT_ERR_DEVICE,                 \
                        std::string(#call) + ": " + hipGetErrorName(e_) + " (" + hipGetErrorString(e_) + ")"); \
        }                                                                                                   \
    } while (0)

// ------------------------------------------------------------- binary16 ----
// IEEE-754 binary16, round to nearest even, from the format definition.
__device__ __forceinline__ uint32_t f32_to_f16_bits(float f) {
    const uint32_t x = __float_as_uint(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return ax == 0x7f800000u ? (sign | 0x7c00u) : (sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
    if (ax >= 0x477ff000u) return sign | 0x7c00u;
    if (ax >= 0x38800000u) {
        uint32_t h = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13);
        const uint32_t rem = ax & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        return sign | h;
    }
    if (ax < 0x33000000u) return sign;
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - (ax >> 23);
    uint32_t h = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return sign | h;
}

__device__ __forceinline__ float f16_bits_to_f32(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    if (e == 0) return __uint_as_float(sign | __float_as_uint((float)m * 5.9604644775390625e-08f));
    if (e == 31) return __uint_as_float(sign | 0x7f800000u | (m << 13));
    return __uint_as_float(sign | ((e + 112u) << 23) | (m << 13));
}

// padded(i, j, k) = src(i - 1 mod n, j - 1 mod n, k - 1 mod n), i, j, k in [0, P)
__global__ __launch_bounds__(256) void pad_grid_kernel(const float* __restrict__ src, int64_t stride, int comp,
                                                       int n, int P, int quantise, float* __restrict__ dst) {
    const int64_t total = (int64_t)P * P * P;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % P);
        const int j = (int)((q / P) % P);
        const int k = (int)(q / ((int64_t)P * P));
        const int si = (i + n - 1) % n, sj = (j + n - 1) % n, sk = (k + n - 1) % n;
        const int64_t s = (int64_t)si + (int64_t)n * ((int64_t)sj + (int64_t)n * sk);
        float v = src[s * stride + comp];
        if (quantise) v = f16_bits_to_f32(f32_to_f16_bits(v * 134217728.0f)) / 134217728.0f;
        dst[q] = v;
    }
}

// Elements [e0, e0 + count) of a host field, staged at src[0 .. count), into
// the interior of its padded cubes: element e is component e mod ncomp of
// voxel v = e / ncomp = x + n (y + n z), stored at padded (x + 1, y + 1, z + 1)
// of cube `comp` (cubes pvox floats apart), with the same optional binary16
// quantisation as pad_grid_kernel.
__global__ __launch_bounds__(256) void place_kernel(const float* __restrict__ src, int64_t e0, int64_t count,
                                                    int ncomp, int n, int P, int quantise, float* __restrict__ cube0,
                                                    int64_t pvox) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = e0 + q, v = e / ncomp;
        const int comp = (int)(e - v * ncomp);
        const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((int64_t)n * n));
        float val = src[q];
        if (quantise) val = f16_bits_to_f32(f32_to_f16_bits(val * 134217728.0f)) / 134217728.0f;
        cube0[comp * pvox + (int64_t)(x + 1) + (int64_t)P * ((int64_t)(y + 1) + (int64_t)P * (z + 1))] = val;
    }
}

// The periodic halo of `ncubes` padded cubes (pvox floats apart) from their
// interiors: padded cell (i, j, k) with a coordinate in {0, n + 1, n + 2}
// takes interior cell (w(i), w(j), w(k)), w(i) = (i + n - 1) mod n + 1, the
// value pad_grid_kernel gives it.  Halo cells are enumerated directly: the
// planes k in {0, n + 1, n + 2} (3 P^2 cells), then for interior k the rows
// j in {0, n + 1, n + 2} (3 P n), then for interior j and k the cells
// i in {0, n + 1, n + 2} (3 n^2); every source is an interior cell, so the
// order of the writes does not matter.
__global__ __launch_bounds__(256) void halo_kernel(float* __restrict__ cube0, int ncubes, int n, int P,
                                                   int64_t pvox) {
    const int64_t a = 3 * (int64_t)P * P, b = a + 3 * (int64_t)P * n, per_cube = b + 3 * (int64_t)n * n;
    const int64_t total = per_cube * ncubes;
    auto halo = [n](int h) { return h == 0 ? 0 : n + h; };  // h in {0, 1, 2}
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q / per_cube);
        int64_t r = q - (int64_t)c * per_cube;
        int i, j, k;
        if (r < a) {
            k = halo((int)(r / ((int64_t)P * P)));
            r %= (int64_t)P * P;
            j = (int)(r / P);
            i = (int)(r % P);
        } else if (r < b) {
            r -= a;
            j = halo((int)(r / ((int64_t)P * n)));
            r %= (int64_t)P * n;
            k = 1 + (int)(r / P);
            i = (int)(r % P);
        } else {
            r -= b;
            i = halo((int)(r / ((int64_t)n * n)));
            r %= (int64_t)n * n;
            k = 1 + (int)(r / n);
            j = 1 + (int)(r % n);
        }
        const int si = (i + n - 1) % n + 1, sj = (j + n - 1) % n + 1, sk = (k + n - 1) % n + 1;
        float* const cube = cube0 + c * pvox;
        cube[(int64_t)i + (int64_t)P * ((int64_t)j + (int64_t)P * k)] =
            cube[(int64_t)si + (int64_t)P * ((int64_t)sj + (int64_t)P * sk)];
    }
}

__global__ __launch_bounds__(256) void widen_kernel(const float* __restrict__ src, double* __restrict__ dst,
                                                    int64_t count) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x)
        dst[q] = (double)src[q];
}

// Tile (of rays, or of record columns) taken by workgroup `b` of `nb`.  The dispatcher hands workgroups
// to the chip's XCDs round-robin (b -> XCD b mod 8 on MI355X), and each XCD
// has its own L2.  With g = 8 groups, the workgroups an XCD receives trace
// one contiguous range of tiles, so rays of neighbouring sky patches (which
// read neighbouring cells), or neighbouring record columns, share an L2;
// g = 1 is the identity.
__device__ __forceinline__ int64_t tile_of_block(int64_t b, int64_t nb, int g) {
    if (g <= 1) return b;
    const int64_t x = b % g, idx = b / g, per = nb / g, extra = nb % g;
    return x * per + (x < extra ? x : extra) + idx;
}

// Bicubic cell records of one field from its padded cube (layout: see
// srt::record_offset).  The record of column (i, j), padded plane p holds the
// 4 x 4 taps v(i - 1 + a, j - 1 + b) of that plane, which are padded entries
// (i + a, j + b, p), in the monomial basis
//     c[pw][qw] = sum_{a,b} M[a][pw] M[b][qw] v_ab,  M = srt::kCrMonomial;
// both basis changes run in FP64 and each coefficient is rounded to FP32 once.
//
// A workgroup of kRecRows wavefronts takes a tile of 64 planes p0 .. p0 + 63,
// kRecRows rows j0 .. and kRecCols columns i0 ..: wavefront w builds row
// j0 + w, lane l plane p0 + l, column after column.
//   1. The tile's taps (64 planes x (kRecRows + 3) rows x (kRecCols + 3)
//      columns) are loaded into LDS; each thread issues all of its loads
//      before its first LDS write, so their latencies overlap.
//   2. Each lane keeps the 4 x 4 taps of its current column in registers and
//      moves them one column on with 4 LDS reads (the 12 others are reused).
//   3. A column's 64 records (one per lane) are consecutive in memory (4 KB).
//      They go through a per-wavefront LDS buffer so that each of the 4
//      stores writes 1 KB contiguous (lane l: 16 bytes at 16 l); stored
//      directly, each store instruction would scatter 16 bytes per lane over
//      the 4 KB (3x slower at 512^3).
// Tiles run (plane block, column block, row block) from fastest to slowest
// and are grouped by XCD (tile_of_block): an XCD writes the records of a few
// columns plane block after plane block, close to their order in memory.
// Records are written with non-temporal stores (nothing reads them back
// during the build).  512^3: 2.2 ms per field against 5.5 ms for one thread
// per record in (i, j, p) order.
constexpr int kRecCols = 16, kRecRows = 4;
constexpr int kRecTileCols = kRecCols + 3, kRecTileRows = kRecRows + 3;
constexpr int kRecPlaneStride = kRecTileRows * kRecTileCols + 1;  // odd: lanes (planes) hit distinct banks
constexpr int kRecTileFloats = 64 * kRecTileRows * kRecTileCols;
constexpr int kRecLoadsPerThread = (kRecTileFloats + 64 * kRecRows - 1) / (64 * kRecRows);

__global__ __launch_bounds__(64 * kRecRows) void bicubic_records_kernel(const float* __restrict__ padded, int n,
                                                                       int P, int shift, int nb,
                                                                       float* __restrict__ rec) {
    using f4 = float __attribute__((ext_vector_type(4)));
    __shared__ float taps[64 * kRecPlaneStride];
    __shared__ f4 staged[kRecRows][64 * 4];
    const int npb = (P + 63) / 64, ncb = (n + kRecCols - 1) / kRecCols;
    const int64_t tile = tile_of_block(blockIdx.x, gridDim.x, 8);
    const int pb = (int)(tile % npb);
    const int cb = (int)((tile / npb) % ncb);
    const int rb = (int)(tile / ((int64_t)npb * ncb));
    const int p0 = pb * 64, i0 = cb * kRecCols, j0 = rb * kRecRows;
    {
        float v[kRecLoadsPerThread];
#pragma unroll
        for (int u = 0; u < kRecLoadsPerThread; ++u) {
            const int t = (int)threadIdx.x + u * 64 * kRecRows;
            const int c = t % kRecTileCols, r = (t / kRecTileCols) % kRecTileRows, pl = t / (kRecTileCols * kRecTileRows);
            const int gi = i0 + c, gj = j0 + r, gp = p0 + pl;
            v[u] = 0.0f;
            if (t < kRecTileFloats && gi < P && gj < P && gp < P)
                v[u] = padded[(size_t)gi + (size_t)P * ((size_t)gj + (size_t)P * (size_t)gp)];
        }
#pragma unroll
        for (int u = 0; u < kRecLoadsPerThread; ++u) {
            const int t = (int)threadIdx.x + u * 64 * kRecRows;
            const int c = t % kRecTileCols, r = (t / kRecTileCols) % kRecTileRows, pl = t / (kRecTileCols * kRecTileRows);
            if (t < kRecTileFloats) taps[pl * kRecPlaneStride + r * kRecTileCols + c] = v[u];
        }
    }
    __syncthreads();
    const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int j = j0 + w;
    const int nvalid = min(64, P - p0);  // planes of this block
    const int i_end = min(n, i0 + kRecCols);
    const float* lane_taps = taps + l * kRecPlaneStride + w * kRecTileCols;
    f4* const st = staged[w];
    srt::Grids G;
    G.pad = P;
    G.rec_shift = shift;
    G.rec_nb = nb;
    double win[4][4];  // win[b][a] = tap (i + a, j + b) of the lane's plane
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) win[b][a + 1] = (double)lane_taps[b * kRecTileCols + a];
    for (int i = i0; i < i_end; ++i) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            win[b][0] = win[b][1];
            win[b][1] = win[b][2];
            win[b][2] = win[b][3];
            win[b][3] = (double)lane_taps[b * kRecTileCols + (i - i0) + 3];
        }
        double X[4][4];  // X[b][pw] = sum_a M[a][pw] v_ab
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int pw = 0; pw < 4; ++pw) {
                double c = 0.0;
#pragma unroll
                for (int a = 0; a < 4; ++a) c += srt::kCrMonomial[a][pw] * win[b][a];
                X[b][pw] = c;
            }
        // quarter pw of lane l's record goes to slot 4 l + ((pw + l / 4) mod 4)
        // of the staging buffer (the rotation spreads the 16-byte writes of
        // lanes 4 apart over the banks)
#pragma unroll
        for (int pw = 0; pw < 4; ++pw) {
            f4 c4;
#pragma unroll
            for (int qw = 0; qw < 4; ++qw) {
                double c = 0.0;
#pragma unroll
                for (int b = 0; b < 4; ++b) c += srt::kCrMonomial[b][qw] * X[b][pw];
                c4[qw] = (float)c;
            }
            st[l * 4 + ((pw + (l >> 2)) & 3)] = c4;
        }
        __builtin_amdgcn_wave_barrier();
        if (j < n) {
            f4* const run = reinterpret_cast<f4*>(rec + srt::record_offset(G, i, j, p0));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int r = 16 * k + (l >> 2), q = l & 3;  // record r, quarter q
                const f4 v = st[r * 4 + ((q + (r >> 2)) & 3)];
                if (r < nvalid) __builtin_nontemporal_store(v, run + 4 * r + q);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------- trace ----
struct TraceArgs {
    srt::Grids G;
    int64_t nside;
    int64_t pix_begin;
    int64_t count;
    const int64_t* pix_list;
    int order;
    int n_steps;
    double obs[3];
    float* out;
    double* state;
    int xcd_groups;             // see tile_of_block
    const uint32_t* lane_slot;  // sorted lane -> slot map (srt::order_lanes), or null
    int64_t lane_base;          // first lane of this launch (a trace is split into launches of
                                // at most srt_context::max_launch_rays lanes)
};

// Traces one ray (RING pixel `pix`) and writes its maps to slot `slot` of
// the map-major output (and the final state if `state` is not null).
template <class F, int kLanes = kBlock>
__device__ __forceinline__ void trace_ray(const srt::Grids& G, int64_t nside, int64_t pix, int64_t slot,
                                          int64_t count, int n_steps, const double* __restrict__ obs,
                                          float* __restrict__ out, double* __restrict__ state) {
    using RA = typename F::RA;
    double z, s, ph;
    srt::ring_pixel_direction(nside, pix, z, s, ph);
    double sp, cp;
    sincos(ph, &sp, &cp);

    // state (.cu:560): x, n = -r, e1 = theta-hat, e2 = (sin phi, -cos phi, 0), D = 0, D' = -1.
    // With F::kStateInLds the state and the RK4 stage sum live in LDS (one
    // column per lane, no sharing, no barriers) instead of registers: the
    // right-hand side then has those registers for its interpolations.
    extern __shared__ double lds_state[];
    double y_reg[F::kStateInLds ? 1 : 19];
    RA acc_reg[F::kStateInLds ? 1 : 19];
    double* y_lds = lds_state + threadIdx.x;
    RA* acc_lds = reinterpret_cast<RA*>(lds_state + 19 * kLanes) + threadIdx.x;
    // FP32 forms also keep an FP32 copy of the ten components the right-hand
    // side reads in FP32 (e1, e2, sigma, eps), beside the state, so a stage
    // state is one FP32 multiply-add from it instead of an FP64 one between
    // two conversions
    constexpr bool kYsF32 = sizeof(RA) == 4;
    RA ysf_reg[kYsF32 && !F::kStateInLds ? 10 : 1];
    RA* ysf_lds = acc_lds + 19 * kLanes;
    auto ysf = [&](int j) -> RA& {
        if constexpr (F::kStateInLds) return ysf_lds[j * kLanes];
        else return ysf_reg[j];
    };
    auto y = [&](int j) -> double& {
        if constexpr (F::kStateInLds) return y_lds[j * kLanes];
        else return y_reg[j];
    };
    auto acc = [&](int j) -> RA& {
        if constexpr (F::kStateInLds) return acc_lds[j * kLanes];
        else return acc_reg[j];
    };
    {
        const double init[19] = {obs[0], obs[1], obs[2], -s * cp, -s * sp, -z, z * cp, z * sp, -s, sp, -cp,
                                 0.0,      0.0,      -1.0,     0.0,      0.0,     0.0, 0.0,    0.0};
#pragma unroll
        for (int j = 0; j < 19; ++j) y(j) = init[j];
        if constexpr (kYsF32) {
#pragma unroll
            for (int j = 0; j < 10; ++j) ysf(j) = (RA)init[j < 6 ? 6 + j : 8 + j];
        }
    }
    // RK4 (.cu:513-550).  Position and dA increments stay in FP64 in every
    // form: x' carries the transverse deflection and the FP64 dA decides the
    // dA > 1/2 branch of the right-hand side exactly as the FP64 oracle does.
    // Other derivatives and their stage sums are in RA.
    double accx[3], kx[3], accD, kD;
    RA k[19];
    double comp[4] = {0.0, 0.0, 0.0, 0.0};  // compensated sums of x and dA (mixed form)

    // stage state y + c * k, then the right-hand side into (kxo, ko, kDo)
    auto eval = [&](double c, const double* kxi, const RA* ki, double kDi, double* kxo, RA* ko, double& kDo) {
        double xs[3], ns[3];
        RA ys[10];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            xs[i] = y(i) + c * kxi[i];
            ns[i] = y(3 + i) + c * (double)ki[3 + i];
        }
        const double Ds = y(12) + c * kDi;
        const double Dps = y(13) + c * (double)ki[13];
        if constexpr (kYsF32) {
#pragma unroll
            for (int j = 0; j < 6; ++j) ys[j] = ysf(j) + (RA)c * ki[6 + j];
#pragma unroll
            for (int j = 0; j < 4; ++j) ys[6 + j] = ysf(6 + j) + (RA)c * ki[14 + j];
        } else {
#pragma unroll
            for (int j = 0; j < 6; ++j) ys[j] = (RA)(y(6 + j) + c * (double)ki[6 + j]);
#pragma unroll
            for (int j = 0; j < 4; ++j) ys[6 + j] = (RA)(y(14 + j) + c * (double)ki[14 + j]);
        }
        srt::rhs<F>(G, xs, ns, Ds, Dps, ys, kxo, ko);
        kDo = -Dps;
    };

    for (int step = 0; step < n_steps; ++step) {
        {
            const double zx[3] = {0.0, 0.0, 0.0};
            RA zk[19];
#pragma unroll
            for (int j = 0; j < 19; ++j) zk[j] = RA(0);
            eval(0.0, zx, zk, 0.0, accx, k, accD);  // k1
#pragma unroll
            for (int j = 0; j < 19; ++j) acc(j) = k[j];
        }
        eval(0.5, accx, k, accD, kx, k, kD);  // k2
        {
            const double kx2[3] = {kx[0], kx[1], kx[2]};
            RA k2[19];
#pragma unroll
            for (int j = 0; j < 19; ++j) k2[j] = k[j];
            const double kD2 = kD;
#pragma unroll
            for (int i = 0; i < 3; ++i) accx[i] += 2.0 * kx[i];
#pragma unroll
            for (int j = 0; j < 19; ++j) acc(j) += RA(2) * k[j];
            accD += 2.0 * kD;
            eval(0.5, kx2, k2, kD2, kx, k, kD);  // k3
        }
        {
            const double kx3[3] = {kx[0], kx[1], kx[2]};
            RA k3[19];
#pragma unroll
            for (int j = 0; j < 19; ++j) k3[j] = k[j];
            const double kD3 = kD;
#pragma unroll
            for (int i = 0; i < 3; ++i) accx[i] += 2.0 * kx[i];
#pragma unroll
            for (int j = 0; j < 19; ++j) acc(j) += RA(2) * k[j];
            accD += 2.0 * kD;
            eval(1.0, kx3, k3, kD3, kx, k, kD);  // k4
        }
        constexpr double sixth = 1.0 / 6.0;
        if constexpr (sizeof(RA) == 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = q < 3 ? q : 12;
                const double inc = (q < 3 ? (accx[q] + kx[q]) : (accD + kD)) * sixth - comp[q];
                const double yo = y(j);
                const double t = yo + inc;
                comp[q] = (t - yo) - inc;
                y(j) = t;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 3; ++i) y(i) += (accx[i] + kx[i]) * sixth;
            y(12) += (accD + kD) * sixth;
        }
#pragma unroll
        for (int j = 3; j < 19; ++j)
            if (j != 12) {
                const double t = y(j) + (double)(acc(j) + k[j]) * sixth;
                y(j) = t;
                if constexpr (kYsF32) {
                    if (j >= 6 && j < 12) ysf(j - 6) = (RA)t;
                    if (j >= 14 && j < 18) ysf(j - 8) = (RA)t;
                }
            }
    }
    double yf[19];
#pragma unroll
    for (int j = 0; j < 19; ++j) yf[j] = y(j);

    // deflection: Delta = unit(x - observer) - r (.cu:566-575)
    double d[3] = {yf[0] - obs[0], yf[1] - obs[1], yf[2] - obs[2]};
    const double rn = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] = d[0] * rn - s * cp;
    d[1] = d[1] * rn - s * sp;
    d[2] = d[2] * rn - z;
    const int64_t C = count;
    out[0 * C + slot] = (float)yf[12];
    out[1 * C + slot] = (float)yf[16];
    out[2 * C + slot] = (float)yf[17];
    out[3 * C + slot] = (float)yf[18];
    out[4 * C + slot] = (float)(d[0] * z * cp + d[1] * z * sp - d[2] * s);
    out[5 * C + slot] = (float)(d[0] * sp - d[1] * cp);
    if (state) {
#pragma unroll
        for (int j = 0; j < 19; ++j) state[slot * 19 + j] = yf[j];
    }
}

// Workgroup size of trace_kernel<F>: one wavefront for the record kernel (2 %
// faster than 256 lanes at 512^3 / Nside 1024, 0.5 % faster than 128: a
// workgroup's LDS is free for the next one as soon as its wavefront ends), 256
// lanes for the other forms (which lose 3-13 % at 128).  The persistent batch
// kernel keeps 256 for every form.
template <class F>
constexpr int trace_lanes() {
    return F::kBicubic ? 64 : kBlock;
}

template <class F>
__global__ __launch_bounds__(trace_lanes<F>(), F::kWavesPerSimd) void trace_kernel(TraceArgs A) {
    constexpr int kLanes = trace_lanes<F>();
    const int64_t lane_ray =
        A.lane_base + tile_of_block(blockIdx.x, gridDim.x, A.xcd_groups) * kLanes + threadIdx.x;
    if (lane_ray >= A.count) return;
    int64_t pix, slot;
    const bool nested_sky = A.order == SRT_ORDER_NESTED && !A.lane_slot;
    if (nested_sky) {  // full sky, power-of-two Nside: lane = NESTED pixel
        pix = srt::nest_to_ring(A.nside, lane_ray);
        slot = pix;
    } else {
        slot = A.lane_slot ? (int64_t)A.lane_slot[lane_ray] : lane_ray;
        pix = A.pix_list ? A.pix_list[slot] : A.pix_begin + slot;
    }
    if (pix < 0 || pix >= 12 * A.nside * A.nside) {
        // invalid entry of a caller's pixel list: NaN maps in its slot (in NESTED
        // full-sky order the slot is the pixel itself, so nothing can be written)
        if (!nested_sky) {
            const float nan = __builtin_nanf("");
            for (int m = 0; m < 6; ++m) A.out[m * A.count + slot] = nan;
        }
        return;
    }
    trace_ray<F, kLanes>(A.G, A.nside, pix, slot, A.count, A.n_steps, A.obs, A.out, A.state);
}

// ---------------------------------------------------------------- batch ----
// Many snapshots / observers / pixel ranges / step counts in one persistent
// launch.  Work items ("tiles") of 4 x `tile_rays` consecutive lanes of one
// instance are handed out by atomic counters, one workgroup at a time, so
// instances of very different cost (grid size, step count) balance across the
// chip; wavefront w of the workgroup traces the 64-lane groups w, w + 4, ...
// of the tile, so the four wavefronts of a CU trace neighbouring rays (with
// NESTED lane maps, one 16 x 16 sky block per round) and share its L1, as the
// workgroups of a single trace launch do.  The tiles are split into `groups`
// contiguous ranges with a counter each; workgroup b starts on range
// b mod groups (with 8 groups, the range of its XCD, so neighbouring tiles
// share an L2), then takes tiles of the other ranges once its own is drained.
// Every workgroup leaves after finding all ranges drained.
struct BatchInst {
    srt::Grids G;
    int64_t nside, pix_begin, count;
    int n_steps;
    double obs[3];
    float* out;
    int64_t tile_begin;         // first tile index of this instance
    const uint32_t* lane_slot;  // sorted lane -> slot map, or null (lane i traces slot i)
};

constexpr int kMaxTileGroups = 64;

__global__ void reset_counters_kernel(unsigned long long* counters, int n) {
    if ((int)threadIdx.x < n) counters[threadIdx.x] = 0ull;
}

template <class F>
__global__ __launch_bounds__(kBlock, F::kWavesPerSimd) void batch_kernel(const BatchInst* __restrict__ inst,
                                                                         int n_inst, int64_t total_tiles,
                                                                         int tile_rays, int groups,
                                                                         unsigned long long* __restrict__ counters) {
    __shared__ unsigned long long claimed;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    constexpr int kWaves = kBlock / 64;
    const int home = (int)(blockIdx.x % (unsigned)groups);
    for (int r = 0; r < groups; ++r) {
        const int x = (home + r) % groups;
        const int64_t lo = total_tiles * x / groups, hi = total_tiles * (x + 1) / groups;
        for (;;) {
            if (threadIdx.x == 0) claimed = atomicAdd(counters + x, 1ull);
            __syncthreads();
            const int64_t tile = lo + (int64_t)claimed;
            __syncthreads();  // every wavefront has read `claimed` before the next claim
            if (tile >= hi) break;  // uniform across the workgroup
            int a = 0, b = n_inst - 1;  // last instance with tile_begin <= tile
            while (a < b) {
                const int mid = (a + b + 1) >> 1;
                if (inst[mid].tile_begin <= tile)
                    a = mid;
                else
                    b = mid - 1;
            }
            const BatchInst& I = inst[a];
            const int64_t r0 = (tile - I.tile_begin) * (int64_t)tile_rays * kWaves;
            for (int c = 0; c < tile_rays; c += 64) {
                const int64_t l = r0 + (int64_t)c * kWaves + wave * 64 + lane;
                if (l < I.count) {
                    const int64_t slot = I.lane_slot ? (int64_t)I.lane_slot[l] : l;
                    trace_ray<F>(I.G, I.nside, I.pix_begin + slot, slot, I.count, I.n_steps, I.obs, I.out, nullptr);
                }
            }
        }
    }
}

// The same instances without a persistent loop: one workgroup per tile of
// trace_lanes<F>() consecutive lanes of one instance, balanced by the
// hardware dispatcher as in a single trace launch (no claims, no barriers).
// The dispatcher hands workgroup b to XCD b mod 8; with `groups` ranges,
// workgroup b takes position b / groups of range b mod groups.  Range x holds,
// for every instance in turn, part x of its tiles (tiles [t x / groups,
// t (x + 1) / groups) of its t), so every range carries the same mix of
// instances and each part is a compact run of the instance's lane order.  The
// instances follow in `table[0 .. n_inst)`'s order (longest step count first,
// so that the launch's last workgroups are short ones).  Then
// table[n_inst + x * (n_inst + 1) + j] is the first position of the j-th
// instance of that order in range x, and entry j = n_inst the range's length.
// Workgroups past the end of their range return.
template <class F>
__global__ __launch_bounds__(trace_lanes<F>(), F::kWavesPerSimd) void batch_dispatch_kernel(
    const BatchInst* __restrict__ inst, int n_inst, const int64_t* __restrict__ table, int groups) {
    constexpr int kLanes = trace_lanes<F>();
    const int x = (int)(blockIdx.x % (unsigned)groups);
    const int64_t k = blockIdx.x / (unsigned)groups;
    const int64_t* pb = table + n_inst + (int64_t)x * (n_inst + 1);
    if (k >= pb[n_inst]) return;
    int a = 0, b = n_inst - 1;  // last instance whose part starts at or before k (it is not empty)
    while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (pb[mid] <= k)
            a = mid;
        else
            b = mid - 1;
    }
    const BatchInst& I = inst[table[a]];
    const int64_t tiles = (I.count + kLanes - 1) / kLanes;
    const int64_t l = (tiles * x / groups + (k - pb[a])) * kLanes + threadIdx.x;
    if (l >= I.count) return;
    const int64_t slot = I.lane_slot ? (int64_t)I.lane_slot[l] : l;
    trace_ray<F, kLanes>(I.G, I.nside, I.pix_begin + slot, slot, I.count, I.n_steps, I.obs, I.out, nullptr);
}

// ------------------------------------------------------------- adjoint ----
// The adjoint of the FP64 trace runs per chunk of rays as
//   1. adj_forward_kernel<false>: the forward sweep, storing the state every K
//      steps (checkpoints), the FP64 maps and dL/dy at the end;
// and, per checkpoint segment from the last,
//   2. adj_forward_kernel<true>: the segment's steps again from its
//      checkpoint, storing every stage's state and each field's interpolation
//      outputs up to the third derivatives (a "stage record");
//   3. adj_reverse_kernel: the reverse sweep of the segment's RK4 steps over
//      the stage records (vector-Jacobian products of the stages, pure
//      algebra), writing each field's interpolation adjoints into the records;
//   4. adj_scatter_kernel: the transpose of each record's 64-tap contraction,
//      added to padded FP64 gradient grids.
// Kernels 1 and 2 give a ray two lanes, each interpolating two of its fields
// (potential, B_x, B_y, B_z): the lanes exchange the 40 interpolants through
// LDS and evaluate the right-hand side together; kernel 4 runs one wavefront
// per 64 rays and (step, stage, field).
// Lanes of adj_forward_kernel per ray.  Each lane interpolates 4 / lanes of
// the ray's fields, and every lane evaluates the right-hand side, so a
// wavefront runs the right-hand side once for its 64 / lanes rays.  Two lanes
// per ray run the 128^3 / Nside 64 / 64-step gradient in 12.3 ms against 13.3 ms
// with four (which leave three wavefronts per SIMD for a two-wavefront
// occupancy, and spill), and 397 against 431 ms at 256^3 / Nside 256 / 128
// steps; one lane per ray runs one wavefront per SIMD and is slower still.
constexpr int kAdjLanesPerRay = 2;
constexpr int kAdjFieldsPerLane = 4 / kAdjLanesPerRay;
constexpr int kAdjRays = 64 / kAdjLanesPerRay;  // rays per 64-lane workgroup of adj_forward_kernel
static_assert(kAdjLanesPerRay == 4 || kAdjLanesPerRay == 2, "a ray's lanes split its four fields evenly");
constexpr int kAdjRec = 19 + 4 * 20;            // doubles per stage record: state, then 20 per field
constexpr int64_t kAdjChunkBytes = int64_t(8) << 30;  // per-chunk buffers (records, checkpoints, dL/dy)

struct AdjArgs {
    srt::Grids G;
    int64_t nside, pix_begin;
    const int64_t* pix_list;
    const uint32_t* lane_slot;  // sorted lane -> slot map (srt::order_lanes), or null
    int64_t lane0;              // first lane of this chunk
    int64_t count;              // rays of this chunk
    int64_t C;                  // stride of the chunk's per-ray buffers (component-major, ray fastest)
    int64_t total;              // rays of the whole call: stride of maps64 and map_grad
    int n_steps, K;
    int seg, seg_steps;         // kernels 2-4: the segment and its step count
    int rec_from;               // kernel 1: the first step whose stages it records (the last segment's)
    double obs[3];
    const double* map_grad;  // 6 x total, null = forward only
    double* maps64;          // 6 x total, may be null
    double* ckpt;            // (nseg + 1) x 19 x C
    double* lam;             // 19 x C: dL/dy at the start of the segment being reversed
    double* rec;             // (4 K) x kAdjRec x C stage records
    double* grad[4];         // padded gradient grids: potential, B_x, B_y, B_z
};

// Slot and RING pixel of chunk lane l; false for an invalid entry of a pixel list.
__device__ __forceinline__ bool adj_ray(const AdjArgs& A, int64_t l, int64_t& slot, int64_t& pix) {
    const int64_t lane = A.lane0 + l;
    slot = A.lane_slot ? (int64_t)A.lane_slot[lane] : lane;
    pix = A.pix_list ? A.pix_list[slot] : A.pix_begin + slot;
    return pix >= 0 && pix < 12 * A.nside * A.nside;
}

// Forward sweep (kRecord = false: the whole trace from the pixel's initial
// state, storing checkpoints, maps and the final dL/dy, and the stage records
// of the steps from A.rec_from on: the last segment's, which the reverse
// sweep needs first) or one segment from its checkpoint (kRecord = true:
// storing its stage records).  Lane kAdjLanesPerRay r + h of the workgroup
// serves ray r and interpolates fields kAdjFieldsPerLane h .. (potential and
// B_x, then B_y and B_z, with two lanes per ray).  The ray's state, its two
// stage states and the 40 interpolants live in LDS; component j of a state is
// updated by lane j mod kAdjLanesPerRay of the ray (each lane evaluates the
// whole right-hand side).
template <bool kRecord>
__global__ __launch_bounds__(64, 2) void adj_forward_kernel(AdjArgs A) {
    __shared__ double sy[19][kAdjRays], sbuf[2][19][kAdjRays], sq[40][kAdjRays];
    const int r = threadIdx.x / kAdjLanesPerRay, h = threadIdx.x % kAdjLanesPerRay;
    const int64_t l = (int64_t)blockIdx.x * kAdjRays + r;
    if (l >= A.count) return;  // the lanes of a ray leave together
    int64_t slot, pix;
    if (!adj_ray(A, l, slot, pix)) {
        if (!kRecord && A.maps64 && h == 0)
            for (int m = 0; m < 6; ++m) A.maps64[m * A.total + slot] = __builtin_nan("");
        return;
    }
    const int64_t C = A.C;
    const int K = A.K;
    double z = 0, s = 0, ph = 0, sp = 0, cp = 0;
    if (!kRecord) {
        srt::ring_pixel_direction(A.nside, pix, z, s, ph);
        sincos(ph, &sp, &cp);
        const double init[19] = {A.obs[0], A.obs[1], A.obs[2], -s * cp, -s * sp, -z, z * cp, z * sp, -s, sp, -cp,
                                 0.0,      0.0,      -1.0,     0.0,      0.0,     0.0, 0.0,    0.0};
        if (h == 0)
            for (int j = 0; j < 19; ++j) sy[j][r] = init[j];
    } else {
        const double* ck = A.ckpt + (int64_t)A.seg * 19 * C + l;
        for (int j = h; j < 19; j += kAdjLanesPerRay) sy[j][r] = ck[j * C];
    }
    __syncthreads();
    const size_t py = (size_t)A.G.pad, pz = (size_t)A.G.pad * (size_t)A.G.pad;
    const int steps = kRecord ? A.seg_steps : A.n_steps;
    constexpr int kComps = (19 + kAdjLanesPerRay - 1) / kAdjLanesPerRay;  // state components per lane
    double acc[kComps];  // this lane's components of the RK4 stage sum
    for (int step = 0; step < steps; ++step) {
        if (!kRecord && step % K == 0) {
            double* ck = A.ckpt + (int64_t)(step / K) * 19 * C + l;
            for (int j = h; j < 19; j += kAdjLanesPerRay) ck[j * C] = sy[j][r];
        }
        for (int stage = 0; stage < 4; ++stage) {
            // stage state: y, then y + k1 / 2 in sbuf[0], y + k2 / 2 in sbuf[1], y + k3 in sbuf[0]
            double(*S)[kAdjRays] = stage == 0 ? sy : sbuf[(stage - 1) & 1];
            const double x[3] = {S[0][r], S[1][r], S[2][r]};
            const bool record = kRecord || step >= A.rec_from;
            double* const rq = record ? A.rec + (int64_t)((kRecord ? step : step - A.rec_from) * 4 + stage) *
                                                    kAdjRec * C + l
                                      : nullptr;
            if (record)
                for (int j = h; j < 19; j += kAdjLanesPerRay) rq[j * C] = S[j][r];
            for (int ff = 0; ff < kAdjFieldsPerLane; ++ff) {
                const int f = h * kAdjFieldsPerLane + ff;
                const srt::adj::Stencil st = srt::adj::field_stencil(A.G, x, f);
                const double* grid = srt::adj::field_grid64(A.G, f);
                double o[20];
                if (record)
                    srt::adj::field_interp<true>(grid, py, pz, st, o);
                else
                    srt::adj::field_interp<false>(grid, py, pz, st, o);
                for (int k = 0; k < 10; ++k) sq[10 * f + k][r] = o[k];
                if (record)
                    for (int k = 0; k < 20; ++k) rq[(19 + 20 * f + k) * C] = o[k];
            }
            __syncthreads();  // every field's interpolants are in sq
            double sv[19], dy[19];
            for (int j = 0; j < 19; ++j) sv[j] = S[j][r];
            srt::adj::RhsTape T;
            srt::adj::rhs_forward_from(sv, &sq[0][r], kAdjRays, T, dy);
#pragma unroll
            for (int m = 0; m < kComps; ++m) {
                const int j = h + kAdjLanesPerRay * m;
                if (j >= 19) break;
                // dy[j] by constant indices (a runtime index would put dy in scratch)
                double d;
                if constexpr (kAdjLanesPerRay == 4)
                    d = h == 0 ? dy[4 * m] : (h == 1 ? dy[4 * m + 1] : (h == 2 ? dy[4 * m + 2] : dy[(4 * m + 3) % 19]));
                else
                    d = h == 0 ? dy[2 * m] : dy[(2 * m + 1) % 19];
                if (stage == 0)
                    acc[m] = d;
                else if (stage < 3)
                    acc[m] += 2.0 * d;
                if (stage < 3)
                    sbuf[stage & 1][j][r] = sy[j][r] + (stage == 2 ? 1.0 : 0.5) * d;
                else
                    sy[j][r] = sy[j][r] + (acc[m] + d) * (1.0 / 6.0);
            }
            __syncthreads();  // sq and S are read before they are written again
        }
    }
    if (kRecord) return;
    {
        const int nseg = (A.n_steps + K - 1) / K;
        double* ck = A.ckpt + (int64_t)nseg * 19 * C + l;
        for (int j = h; j < 19; j += kAdjLanesPerRay) ck[j * C] = sy[j][r];
    }
    if (h != 0) return;
    double y[19];
    for (int j = 0; j < 19; ++j) y[j] = sy[j][r];
    // maps (.cu:566-582) in FP64, and dL/dy at the end from dL/dmaps
    const double th[3] = {z * cp, z * sp, -s};
    const double p2[3] = {sp, -cp, 0.0};
    double d[3] = {y[0] - A.obs[0], y[1] - A.obs[1], y[2] - A.obs[2]};
    const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double u[3] = {d[0] / dn, d[1] / dn, d[2] / dn};
    const double rr[3] = {s * cp, s * sp, z};
    double dl1 = 0, dl2 = 0;
    for (int i = 0; i < 3; ++i) {
        dl1 += (u[i] - rr[i]) * th[i];
        dl2 += (u[i] - rr[i]) * p2[i];
    }
    const int64_t T = A.total;
    if (A.maps64) {
        A.maps64[0 * T + slot] = y[12];
        A.maps64[1 * T + slot] = y[16];
        A.maps64[2 * T + slot] = y[17];
        A.maps64[3 * T + slot] = y[18];
        A.maps64[4 * T + slot] = dl1;
        A.maps64[5 * T + slot] = dl2;
    }
    if (!A.map_grad) return;
    double lam[19];
    for (int j = 0; j < 19; ++j) lam[j] = 0.0;
    const double mb[6] = {A.map_grad[0 * T + slot], A.map_grad[1 * T + slot], A.map_grad[2 * T + slot],
                          A.map_grad[3 * T + slot], A.map_grad[4 * T + slot], A.map_grad[5 * T + slot]};
    lam[12] = mb[0];
    lam[16] = mb[1];
    lam[17] = mb[2];
    lam[18] = mb[3];
    // d(defl)/dx = (1 - u u^T) / |d| applied to mb4 theta-hat + mb5 p2
    double w[3], uw = 0;
    for (int i = 0; i < 3; ++i) {
        w[i] = mb[4] * th[i] + mb[5] * p2[i];
        uw += u[i] * w[i];
    }
    for (int i = 0; i < 3; ++i) lam[i] = (w[i] - u[i] * uw) / dn;
    for (int j = 0; j < 19; ++j) A.lam[j * C + l] = lam[j];
}

// Per-lane vectors of the reverse sweep, in LDS (one column per lane, no
// sharing, no barriers): dL/dy of the step being reversed, its running sum
// over the stages, and the stage VJP's input dL/dk and output dL/ds.
constexpr int kAdjRevBlock = 64;
enum { kLam = 0, kYbar = 1, kKbar = 2, kSbar = 3 };
__shared__ double adj_vec[4][19][kAdjRevBlock];

// dL/ds of one stage given dL/dk (adj_vec[kKbar]), into adj_vec[kSbar], from
// the stage record at `rq` (component c at rq[c * C]); the record's per-field
// value, gradient and Hessian slots receive their adjoints.
__device__ __forceinline__ void stage_vjp(double* __restrict__ rq_in, int64_t C) {
    const int l = threadIdx.x;
    const srt::adj::gptr<double> rq = (srt::adj::gptr<double>)rq_in;
    double sv[19], kbar[19], sbar[19], q[40];
    for (int j = 0; j < 19; ++j) {
        sv[j] = rq[j * C];
        kbar[j] = adj_vec[kKbar][j][l];
    }
    for (int f = 0; f < 4; ++f)
        for (int k = 0; k < 10; ++k) q[10 * f + k] = rq[(19 + 20 * f + k) * C];
    srt::adj::RhsTape T;
    double dy[19];
    srt::adj::rhs_forward_from(sv, q, 1, T, dy);
    srt::adj::rhs_vjp_bars(T, kbar, rq + 19 * C, C, sbar);
    for (int j = 0; j < 19; ++j) adj_vec[kSbar][j][l] = sbar[j];
}

// Reverse sweep of one segment, one lane per ray: for each step from the
// last, the transpose of y' = y + (k1 + 2 k2 + 2 k3 + k4) / 6 through the
// four stages' VJPs (k4, k3, k2, k1).
__global__ __launch_bounds__(kAdjRevBlock) void adj_reverse_kernel(AdjArgs A) {
    const int64_t l = (int64_t)blockIdx.x * kAdjRevBlock + threadIdx.x;
    if (l >= A.count) return;
    int64_t slot, pix;
    if (!adj_ray(A, l, slot, pix)) return;
    const int t = threadIdx.x;
    const int64_t C = A.C;
    double(*const vec)[19][kAdjRevBlock] = adj_vec;
    for (int j = 0; j < 19; ++j) vec[kLam][j][t] = A.lam[j * C + l];
    const int64_t stage_stride = (int64_t)kAdjRec * C;
    for (int st = A.seg_steps - 1; st >= 0; --st) {
        double* const rq = A.rec + (int64_t)(st * 4) * kAdjRec * C + l;
        for (int j = 0; j < 19; ++j) {
            vec[kYbar][j][t] = vec[kLam][j][t];
            vec[kKbar][j][t] = vec[kLam][j][t] * (1.0 / 6.0);
        }
        // stages 4, 3, 2, 1 (stage state s4 = y + k3, s3 = y + k2 / 2,
        // s2 = y + k1 / 2, s1 = y): one copy of the stage VJP in a loop, so
        // its registers are allocated once and no call saves them
#pragma nounroll
        for (int stage = 3; stage >= 0; --stage) {
            stage_vjp(rq + stage * stage_stride, C);
            if (stage > 0) {
                const double c_s = stage == 3 ? 1.0 : 0.5;               // ds/dk of the earlier stage
                const double c_lam = stage == 1 ? 1.0 / 6.0 : 1.0 / 3.0;  // RK4 weight of that k
                for (int j = 0; j < 19; ++j) {
                    vec[kYbar][j][t] += vec[kSbar][j][t];
                    vec[kKbar][j][t] = c_s * vec[kSbar][j][t] + vec[kLam][j][t] * c_lam;
                }
            } else {
                for (int j = 0; j < 19; ++j) vec[kLam][j][t] = vec[kYbar][j][t] + vec[kSbar][j][t];
            }
        }
    }
    for (int j = 0; j < 19; ++j) A.lam[j * C + l] = vec[kLam][j][t];
}

// Grid gradient of one segment's interpolation adjoints.  Workgroup b handles
// the 64 rays of tile b mod tiles for job b / tiles = (step, field): the
// records of the step's 4 stages of one field, whose stencils overlap (a ray
// moves about one cell per step), are summed in one LDS tile
// over the box of all of them (in z slabs that fit the tile) and flushed with
// one global atomic per touched cell; only a box wider than the tile in one
// z plane (rays across the periodic wrap) takes each record on its own.  Consecutive workgroups take neighbouring sky patches of
// one job.  The stencils are recomputed from the recorded stage positions
// exactly as the recording kernel located them.
__global__ __launch_bounds__(64) void adj_scatter_kernel(AdjArgs A) {
    const int64_t tiles = (A.count + 63) / 64;
    const int64_t job = blockIdx.x / tiles, tile = blockIdx.x - job * tiles;
    const int64_t l = tile * 64 + threadIdx.x;
    if (l >= A.count) return;
    int64_t slot, pix;
    if (!adj_ray(A, l, slot, pix)) return;
    const int f = (int)(job & 3);
    const int s0 = (int)(job >> 2), s1 = s0 + 1;
    const int64_t C = A.C;
    const size_t py = (size_t)A.G.pad, pz = (size_t)A.G.pad * (size_t)A.G.pad;
    double* const gdst = A.grad[f];
    auto record = [&](int q) {
        return (srt::adj::gptr<const double>)(A.rec + (int64_t)q * kAdjRec * C + l);
    };
    auto stencil = [&](int q) {
        const srt::adj::gptr<const double> rq = record(q);
        const double x[3] = {rq[0], rq[C], rq[2 * C]};
        return srt::adj::field_stencil(A.G, x, f);
    };
    auto taps = [&](int q, const srt::adj::Stencil& st, auto&& add) {
        const srt::adj::gptr<const double> rq = record(q);
        double bar[10];
        for (int k = 0; k < 10; ++k) bar[k] = rq[(19 + 20 * f + k) * C];
        const srt::adj::Axis4 ax = srt::adj::cr_axis4(st.t[0]), ay = srt::adj::cr_axis4(st.t[1]),
                              az = srt::adj::cr_axis4(st.t[2]);
        add(st, ax, ay, az, bar);
    };
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
    for (int q = 4 * s0; q < 4 * s1; ++q) {
        const srt::adj::Stencil st = stencil(q);
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], st.c[k]);
            hi[k] = max(hi[k], st.c[k]);
        }
    }
    const srt::adj::TileBox B = srt::adj::wave_box(lo, hi);
    const int dz = srt::adj::slab_planes(B);
    if (dz > 0) {
        for (int zlo = 0; zlo < B.nz; zlo += dz) {
            const srt::adj::TileBox P = srt::adj::slab(B, zlo, dz);
            srt::adj::tile_clear(P.cells());
            for (int q = 4 * s0; q < 4 * s1; ++q) {
                const srt::adj::Stencil st = stencil(q);
                if (st.c[2] + 3 < P.z0 || st.c[2] >= P.z0 + P.nz) continue;  // no tap in this slab
                taps(q, st, [&](const srt::adj::Stencil& s, const auto& ax, const auto& ay, const auto& az,
                                const double* bar) { srt::adj::tile_add(P, s.c[0], s.c[1], s.c[2], ax, ay, az, bar); });
            }
            srt::adj::tile_flush(gdst, P, py, pz);
        }
    } else {
        for (int q = 4 * s0; q < 4 * s1; ++q)
            taps(q, stencil(q), [&](const srt::adj::Stencil& st, const auto& ax, const auto& ay, const auto& az,
                                    const double* bar) {
                srt::adj::scatter_taps(gdst, st.c[0], st.c[1], st.c[2], py, pz, ax, ay, az, bar);
            });
    }
}

// Folds the padded gradient grid onto the periodic n^3 grid: voxel (i, j, k)
// collects every padded entry whose index minus one is congruent to it.
__global__ __launch_bounds__(256) void fold_grad_kernel(const double* __restrict__ padded, int n, int P,
                                                        double* __restrict__ out, int64_t stride, int comp) {
    const int64_t total = (int64_t)n * n * n;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < total; v += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(v % n), j = (int)((v / n) % n), k = (int)(v / ((int64_t)n * n));
        int px[2], py[2], pz[2], cx = 0, cy = 0, cz = 0;
        // padded index q in [0, P) maps to (q - 1) mod n
        for (int q = i + 1; q < P; q += n) px[cx++] = q;
        if (i == n - 1) px[cx++] = 0;
        for (int q = j + 1; q < P; q += n) py[cy++] = q;
        if (j == n - 1) py[cy++] = 0;
        for (int q = k + 1; q < P; q += n) pz[cz++] = q;
        if (k == n - 1) pz[cz++] = 0;
        double sum = 0.0;
        for (int c = 0; c < cz; ++c)
            for (int b = 0; b < cy; ++b)
                for (int a = 0; a < cx; ++a) sum += padded[(int64_t)px[a] + (int64_t)P * (py[b] + (int64_t)P * pz[c])];
        out[v * stride + comp] = sum;
    }
}

// LDS bytes per workgroup of kLanes lanes tracing with form F
template <class F, int kLanes = kBlock>
constexpr unsigned lds_bytes() {
    using RA = typename F::RA;
    if (!F::kStateInLds) return 0u;
    // y (FP64) and the stage sum (RA), plus the FP32 copy of ten components
    return (unsigned)(kLanes * (19 * (sizeof(double) + sizeof(RA)) + (sizeof(RA) == 4 ? 10 * sizeof(RA) : 0)));
}

// trace_kernel<F> over A.count lanes, specialised for a power-of-two grid
// edge, in launches of at most `max_lanes` lanes (a multiple of 256): one
// launch may not hold 2^32 or more lanes (grid size times workgroup size)
template <class F>
void launch_trace(bool pow2, hipStream_t st, const TraceArgs& A_in, int64_t max_lanes) {
    constexpr int lanes = trace_lanes<F>();
    constexpr unsigned lds = lds_bytes<F, lanes>();
    TraceArgs A = A_in;
    for (int64_t base = 0; base < A.count; base += max_lanes) {
        A.lane_base = base;
        const int64_t n = std::min(max_lanes, A.count - base);
        const unsigned blocks = (unsigned)((n + lanes - 1) / lanes);
        if (pow2)
            hipLaunchKernelGGL(trace_kernel<srt::Pow2Grid<F>>, dim3(blocks), dim3(lanes), lds, st, A);
        else
            hipLaunchKernelGGL(trace_kernel<F>, dim3(blocks), dim3(lanes), lds, st, A);
    }
}

}  // namespace

// ----------------------------------------------------- fixed scratch ----
// Device scratch of srt_trace_noalloc, for callers that cannot allocate
// device memory: the sorted lane order's lane -> slot map (4 B per ray) and
// its sort space (keys, a second key buffer, slot indices and the radix
// sort's temporary storage, about 20 B per ray) for partial-sky NESTED traces
// of up to kNoAllocMaxSortedRays rays.  Every GPU the library runs on holds
// one copy of the array until the process ends.  Host code never names the
// array except to take its address: the address of each device's copy comes
// from hipGetSymbolAddress (g_scratch, filled by srt_create), and only that
// address is passed to kernels.  Calls that use it at the same time on one
// GPU race; nothing in it needs to start at zero (the sort writes every key
// and slot it reads).
constexpr size_t kScratchBytes = size_t(48) << 20;  // 48 MiB
constexpr int64_t kNoAllocMaxSortedRays = int64_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch[kMaxScratchDevices];  // per device: address of its copy of `scratch`, or null

// Takes the address of the current device's copy of `scratch` (once per
// device); `device` must be current.
int setup_scratch(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(SRT_ERR_UNSUPPORTED, "device ordinal beyond the scratch table (" +
                                             std::to_string(kMaxScratchDevices) + " devices)");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return SRT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(SRT_ERR_DEVICE, std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorName(e));
    }
    g_scratch[device] = p;
    return SRT_OK;
}

void* scratch_address(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch[device];
}

}  // namespace

// ------------------------------------------------------------------ C ABI ----
// Host transfers go through kXferRing pinned chunks of kXferChunkBytes, each
// filled (or emptied) by kXferThreads host threads while the DMA engine moves
// the previous one: 55 GB/s host -> device on the MI355X boxes against
// 26-28 GB/s for hipMemcpy from pageable memory (2 GiB: 39 against 76-81 ms);
// one thread copies too slowly to keep up (28 GB/s).  The chunk size weighs
// the first upload of a context, which pins the ring (512^3 from host: 53 ms
// with 16 MiB chunks, 56 with 32, 65 with 64), against later ones (43, 41,
// 41 ms): a one-snapshot process, the reference's use, uploads once.
constexpr int kXferRing = 3, kXferThreads = 4;
constexpr size_t kXferChunkBytes = size_t(16) << 20;

struct srt_context {
    int device = 0;
    int n = 0;
    int pad = 0;
    float* grids = nullptr;  // 4 padded cubes: phi, Bx, By, Bz
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timed = false;
    // adjoint work buffers, grown on demand and kept for later calls
    double* adj_buf = nullptr;
    size_t adj_bytes = 0;
    // FP64 copies of the padded grids, built on the first FP64 trace after an
    // upload (the buffer is kept across uploads of the same grid edge)
    double* grids64 = nullptr;
    bool grids64_valid = false;
    // transfers between pageable host buffers and the device (srt_upload from
    // host memory, srt_trace_host): a ring of pinned host buffers and device
    // buffers of xfer_chunk bytes each, one stream and an event per slot,
    // created by the first such call and kept
    hipStream_t xfer = nullptr;
    size_t xfer_chunk = 0;  // kXferChunkBytes, or less under a per-allocation limit
    char* xfer_host[kXferRing] = {};
    float* xfer_dev[kXferRing] = {};
    hipEvent_t xfer_ev[kXferRing] = {};
    // bicubic cell records of the four fields (SRT_LAYOUT_BICUBIC)
    int layout = SRT_LAYOUT_TAPS;
    int brick_shift = 4;  // 16 x 16 column bricks
    float* rec = nullptr;
    size_t rec_field_floats = 0;
    // workgroup -> tile grouping of trace launches (tile_of_block); from
    // SRT_XCD_GROUPS at context creation, default 8 (the XCDs of an MI355X)
    int xcd_groups = 8;
    // lanes per trace launch (a multiple of 256): a launch may not hold 2^32
    // or more lanes, so larger traces are split; from SRT_MAX_LAUNCH_RAYS at
    // context creation (rounded down to a multiple of 256), so that tests can
    // exercise the split with small traces
    int64_t max_launch_rays = int64_t(1) << 31;
    // largest ray chunk of an adjoint call (0 = as many rays as the chunk's
    // buffers allow in kAdjChunkBytes); from SRT_ADJ_CHUNK_RAYS at context
    // creation (rounded down to a multiple of 64), so that tests can run the
    // chunk loop with small traces
    int64_t adj_chunk_rays = 0;
    // sorted lane order of partial-sky NESTED traces: lane -> slot map and the
    // sort's work space, grown on demand
    void* order_buf = nullptr;
    size_t order_bytes = 0;
    // largest single device allocation the context may make (0 = no limit);
    // from SRT_MAX_ALLOC_BYTES at creation, so that callers and tests can
    // exercise the out-of-memory paths on any device
    size_t max_alloc = 0;
    // bumped whenever a device buffer that a batch plan's table may point at
    // (padded grids, FP64 copies, records) is freed or replaced, so that plans
    // rebuild their tables before the next launch
    uint64_t generation = 0;
};

namespace {

// hipMalloc within the context's allocation limit
hipError_t ctx_malloc(const srt_context* ctx, void** p, size_t bytes) {
    if (ctx->max_alloc && bytes > ctx->max_alloc) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
    }
    return e;
}

template <class T>
hipError_t ctx_malloc(const srt_context* ctx, T** p, size_t bytes) {
    return ctx_malloc(ctx, reinterpret_cast<void**>(p), bytes);
}

// FP64 copies of the padded grids (the FP64 kernel and the adjoint read them)
int ensure_grids64(srt_context* ctx, hipStream_t st) {
    if (ctx->grids64 && ctx->grids64_valid) return SRT_OK;
    const size_t pvox = (size_t)ctx->pad * ctx->pad * ctx->pad;
    if (!ctx->grids64) SRT_HIP(ctx_malloc(ctx, &ctx->grids64, 4 * pvox * sizeof(double)));
    ++ctx->generation;
    const int blocks = (int)std::min<int64_t>(((int64_t)(4 * pvox) + 255) / 256, 65536);
    hipLaunchKernelGGL(widen_kernel, dim3(blocks), dim3(256), 0, st, ctx->grids, ctx->grids64, (int64_t)(4 * pvox));
    SRT_HIP(hipGetLastError());
    ctx->grids64_valid = true;
    return SRT_OK;
}

void set_grids64(const srt_context* ctx, srt::Grids& G) {
    const size_t pvox = (size_t)ctx->pad * ctx->pad * ctx->pad;
    G.phi64 = ctx->grids64;
    for (int a = 0; a < 3; ++a) G.b64[a] = ctx->grids64 ? ctx->grids64 + (size_t)(1 + a) * pvox : nullptr;
}

// memcpy of `bytes` bytes split over kXferThreads host threads (parts whose
// thread cannot be started are copied by the calling thread: no exception
// leaves the C ABI)
void parallel_copy(void* dst, const void* src, size_t bytes) {
    const size_t per = (bytes + kXferThreads - 1) / kXferThreads;
    auto part = [dst, src, bytes, per](int t) {
        const size_t a = (size_t)t * per;
        if (a < bytes) std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, std::min(per, bytes - a));
    };
    std::thread th[kXferThreads - 1];
    bool started[kXferThreads - 1] = {};
    for (int t = 1; t < kXferThreads; ++t) {
        try {
            th[t - 1] = std::thread(part, t);
            started[t - 1] = true;
        } catch (...) {
        }
    }
    part(0);
    for (int t = 1; t < kXferThreads; ++t) {
        if (started[t - 1])
            th[t - 1].join();
        else
            part(t);
    }
}

void free_xfer(srt_context* ctx);

// The transfer ring (see kXferRing), created on first use; all of it or none.
int ensure_xfer(srt_context* ctx) {
    if (ctx->xfer) return SRT_OK;
    ctx->xfer_chunk = kXferChunkBytes;
    if (ctx->max_alloc && ctx->max_alloc < kXferChunkBytes)
        ctx->xfer_chunk = std::max<size_t>(4096, ctx->max_alloc / 4096 * 4096);
    hipError_t e = hipStreamCreateWithFlags(&ctx->xfer, hipStreamNonBlocking);
    for (int k = 0; k < kXferRing && e == hipSuccess; ++k) {
        e = hipHostMalloc(reinterpret_cast<void**>(&ctx->xfer_host[k]), ctx->xfer_chunk, hipHostMallocDefault);
        if (e == hipSuccess) e = ctx_malloc(ctx, &ctx->xfer_dev[k], ctx->xfer_chunk);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->xfer_ev[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        free_xfer(ctx);
        return fail(e == hipErrorOutOfMemory ? SRT_ERR_OUT_OF_MEMORY : SRT_ERR_DEVICE,
                    std::string("host transfer buffers: ") + hipGetErrorName(e));
    }
    return SRT_OK;
}

void free_xfer(srt_context* ctx) {
    for (int k = 0; k < kXferRing; ++k) {
        if (ctx->xfer_host[k]) (void)hipHostFree(ctx->xfer_host[k]);
        if (ctx->xfer_dev[k]) (void)hipFree(ctx->xfer_dev[k]);
        if (ctx->xfer_ev[k]) (void)hipEventDestroy(ctx->xfer_ev[k]);
        ctx->xfer_host[k] = nullptr;
        ctx->xfer_dev[k] = nullptr;
        ctx->xfer_ev[k] = nullptr;
    }
    if (ctx->xfer) (void)hipStreamDestroy(ctx->xfer);
    ctx->xfer = nullptr;
}

// One host field of `elems` floats (ncomp components per voxel) into the
// interior of its padded cubes (place_kernel), chunk after chunk through the
// transfer ring, on the transfer stream.  Returns once every chunk has been
// copied out of the host buffer (the placement may still run).
int upload_host_field(srt_context* ctx, const float* host, int64_t elems, int ncomp, int quantise, int n, int P,
                      float* cube0, int& slot) {
    const int64_t pvox = (int64_t)P * P * P, chunk = (int64_t)(ctx->xfer_chunk / sizeof(float));
    for (int64_t e0 = 0; e0 < elems; e0 += chunk) {
        const int k = slot;
        slot = (slot + 1) % kXferRing;
        const int64_t len = std::min(chunk, elems - e0);
        SRT_HIP(hipEventSynchronize(ctx->xfer_ev[k]));  // the slot's previous chunk has been placed
        parallel_copy(ctx->xfer_host[k], host + e0, (size_t)len * sizeof(float));
        SRT_HIP(hipMemcpyAsync(ctx->xfer_dev[k], ctx->xfer_host[k], (size_t)len * sizeof(float),
                               hipMemcpyHostToDevice, ctx->xfer));
        const int blocks = (int)std::min<int64_t>((len + 255) / 256, 16384);
        hipLaunchKernelGGL(place_kernel, dim3(blocks), dim3(256), 0, ctx->xfer, ctx->xfer_dev[k], e0, len, ncomp, n,
                           P, quantise, cube0, pvox);
        SRT_HIP(hipGetLastError());
        SRT_HIP(hipEventRecord(ctx->xfer_ev[k], ctx->xfer));
    }
    return SRT_OK;
}

// `bytes` bytes of device memory into a pageable host buffer through the
// transfer ring: up to kXferRing chunks in flight, each copied out by
// kXferThreads host threads once it has arrived.  The caller has finished the
// work that writes `dev`.
int download_to_host(srt_context* ctx, void* host, const void* dev, size_t bytes) {
    const int rc = ensure_xfer(ctx);
    if (rc != SRT_OK) return rc;
    const size_t chunk = ctx->xfer_chunk, nchunks = (bytes + chunk - 1) / chunk;
    auto issue = [&](size_t c) -> hipError_t {
        const size_t off = c * chunk, len = std::min(chunk, bytes - off);
        const int k = (int)(c % kXferRing);
        hipError_t e = hipMemcpyAsync(ctx->xfer_host[k], static_cast<const char*>(dev) + off, len,
                                      hipMemcpyDeviceToHost, ctx->xfer);
        if (e == hipSuccess) e = hipEventRecord(ctx->xfer_ev[k], ctx->xfer);
        return e;
    };
    for (size_t c = 0; c < std::min<size_t>(nchunks, kXferRing); ++c) SRT_HIP(issue(c));
    for (size_t c = 0; c < nchunks; ++c) {
        const int k = (int)(c % kXferRing);
        const size_t off = c * chunk, len = std::min(chunk, bytes - off);
        SRT_HIP(hipEventSynchronize(ctx->xfer_ev[k]));
        parallel_copy(static_cast<char*>(host) + off, ctx->xfer_host[k], len);
        if (c + kXferRing < nchunks) SRT_HIP(issue(c + kXferRing));
    }
    return SRT_OK;
}

int brick_columns(int n, int shift) {
    return (n + (1 << shift) - 1) >> shift;
}

void free_records(srt_context* ctx) {
    if (ctx->rec) {
        (void)hipFree(ctx->rec);
        ++ctx->generation;
    }
    ctx->rec = nullptr;
    ctx->rec_field_floats = 0;
}

// Builds the bicubic cell records of the four fields from the padded cubes,
// into the current record buffer when it has the right size (a new brick edge
// that divides the grid edge keeps the size; freeing and allocating hundreds
// of GB costs seconds).
int build_records(srt_context* ctx) {
    const int n = ctx->n, P = ctx->pad, s = ctx->brick_shift, nb = brick_columns(n, s);
    const size_t cols = (size_t)nb * nb << (2 * s);
    const size_t field = cols * (size_t)P * 16;
    if (ctx->rec && ctx->rec_field_floats != field) free_records(ctx);
    if (!ctx->rec) {
        const hipError_t ea = ctx_malloc(ctx, &ctx->rec, 4 * field * sizeof(float));
        if (ea != hipSuccess) {
            return fail(ea == hipErrorOutOfMemory ? SRT_ERR_OUT_OF_MEMORY : SRT_ERR_DEVICE,
                        "bicubic records: hipMalloc of " + std::to_string(4 * field * sizeof(float)) + " bytes: " +
                            hipGetErrorName(ea));
        }
    }
    ctx->rec_field_floats = field;
    ++ctx->generation;
    const size_t pvox = (size_t)P * P * P;
    // tiles: 64 planes x kRecCols columns x kRecRows rows (at most 33 x 128 x 512
    // for the largest grid edge)
    const int64_t tiles = (int64_t)((P + 63) / 64) * ((n + kRecCols - 1) / kRecCols) * ((n + kRecRows - 1) / kRecRows);
    for (int f = 0; f < 4; ++f)
        hipLaunchKernelGGL(bicubic_records_kernel, dim3((unsigned)tiles), dim3(64 * kRecRows), 0, 0,
                           ctx->grids + (size_t)f * pvox, n, P, s, nb, ctx->rec + (size_t)f * field);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_records(ctx);
        return fail(SRT_ERR_DEVICE, std::string("bicubic records: ") + hipGetErrorName(e));
    }
    return SRT_OK;
}

void set_records(const srt_context* ctx, srt::Grids& G) {
    for (int f = 0; f < 4; ++f) G.rec[f] = ctx->rec ? ctx->rec + (size_t)f * ctx->rec_field_floats : nullptr;
    G.rec_shift = ctx->brick_shift;
    G.rec_nb = brick_columns(ctx->n, ctx->brick_shift);
}

// Lane order a trace of `count` rays runs in for a requested order: NESTED over
// the full sky at a power-of-two Nside needs no sort (lane = NESTED pixel);
// any other NESTED request sorts the rays, which covers at most
// srt::kMaxOrderedRays rays, and beyond that the trace runs in LINEAR order
// (the maps do not depend on the order).  `sorted` says whether a sort is run.
int effective_order(int64_t nside, int64_t pix_begin, int64_t count, bool has_list, int requested, bool* sorted) {
    const bool sky = !has_list && pix_begin == 0 && count == 12 * nside * nside && (nside & (nside - 1)) == 0;
    const bool sort = requested == SRT_ORDER_NESTED && count > 0 && !sky;
    if (sort && count > srt::kMaxOrderedRays) {
        if (sorted) *sorted = false;
        return SRT_ORDER_LINEAR;
    }
    if (sorted) *sorted = sort;
    return requested;
}

}  // namespace

extern "C" {

const char* srt_version(void) { return "srt 0.1 (gfx950)"; }

const char* srt_strerror(int status) {
    switch (status) {
        case SRT_OK: return "success";
        case SRT_ERR_INVALID_ARGUMENT: return "invalid argument";
        case SRT_ERR_DEVICE: return "HIP device error";
        case SRT_ERR_OUT_OF_MEMORY: return "out of device memory";
        case SRT_ERR_NOT_UPLOADED: return "no snapshot uploaded";
        case SRT_ERR_SHAPE: return "unsupported grid shape";
        case SRT_ERR_UNSUPPORTED: return "unsupported request";
        default: return "unknown status";
    }
}

const char* srt_last_error(void) { return g_last_error.c_str(); }

int srt_effective_order(int64_t nside, int64_t pix_begin, int64_t count, int has_pix_list, int ray_order,
                        int* effective) {
    if (!effective) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_effective_order: null output");
    if (nside <= 0 || nside > (1 << 20) || count < 0)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_effective_order: nside or count out of range");
    if (ray_order != SRT_ORDER_NESTED && ray_order != SRT_ORDER_LINEAR)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_effective_order: unknown ray order");
    *effective = effective_order(nside, pix_begin, count, has_pix_list != 0, ray_order, nullptr);
    return SRT_OK;
}

int srt_create(int device, srt_context** ctx) {
    if (!ctx) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_create: ctx is null");
    *ctx = nullptr;
    int ndev = 0;
    SRT_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_create: device " + std::to_string(device) + " of " +
                                                  std::to_string(ndev));
    SRT_HIP(hipSetDevice(device));
    {
        const int rc = setup_scratch(device);
        if (rc != SRT_OK) return rc;
    }
    srt_context* c = new srt_context();
    c->device = device;
    if (const char* g = std::getenv("SRT_XCD_GROUPS")) {
        const int v = std::atoi(g);
        if (v >= 1 && v <= 64) c->xcd_groups = v;
    }
    if (const char* m = std::getenv("SRT_MAX_ALLOC_BYTES")) {
        const long long v = std::atoll(m);
        if (v > 0) c->max_alloc = (size_t)v;
    }
    if (const char* m = std::getenv("SRT_ADJ_CHUNK_RAYS")) {
        const long long v = std::atoll(m) / 64 * 64;
        if (v >= 64) c->adj_chunk_rays = v;
    }
    if (const char* m = std::getenv("SRT_MAX_LAUNCH_RAYS")) {
        const long long v = std::atoll(m) / 256 * 256;
        if (v >= 256) c->max_launch_rays = std::min<int64_t>(v, c->max_launch_rays);
    }
    hipError_t e1 = hipEventCreate(&c->ev_start);
    hipError_t e2 = hipEventCreate(&c->ev_stop);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        srt_destroy(c);
        return fail(SRT_ERR_DEVICE, "srt_create: hipEventCreate failed");
    }
    *ctx = c;
    return SRT_OK;
}

int srt_destroy(srt_context* ctx) {
    if (!ctx) return SRT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->grids) (void)hipFree(ctx->grids);
    if (ctx->adj_buf) (void)hipFree(ctx->adj_buf);
    if (ctx->grids64) (void)hipFree(ctx->grids64);
    free_records(ctx);
    free_xfer(ctx);
    if (ctx->order_buf) (void)hipFree(ctx->order_buf);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    delete ctx;
    return SRT_OK;
}

int srt_grid_info(srt_context* ctx, int64_t* nx, int64_t* ny, int64_t* nz, int64_t* padded_edge) {
    if (!ctx) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_grid_info: ctx is null");
    if (!ctx->grids) return fail(SRT_ERR_NOT_UPLOADED, "srt_grid_info: no snapshot uploaded");
    if (nx) *nx = ctx->n;
    if (ny) *ny = ctx->n;
    if (nz) *nz = ctx->n;
    if (padded_edge) *padded_edge = ctx->pad;
    return SRT_OK;
}

int srt_upload(srt_context* ctx, const float* phi, const float* shift_xyz, int64_t nx, int64_t ny, int64_t nz,
               int inputs_on_device, int shift_format) {
    if (!ctx || !phi || !shift_xyz) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_upload: null pointer");
    if (nx != ny || ny != nz)
        return fail(SRT_ERR_SHAPE, "srt_upload: grid must be cubic (got " + std::to_string(nx) + " x " +
                                       std::to_string(ny) + " x " + std::to_string(nz) + ")");
    if (nx < 4 || nx > 2048) return fail(SRT_ERR_SHAPE, "srt_upload: edge must be in [4, 2048]");
    if (shift_format != SRT_SHIFT_F32 && shift_format != SRT_SHIFT_F16_SCALED)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_upload: unknown shift format");
    SRT_HIP(hipSetDevice(ctx->device));
    const int n = (int)nx, P = n + 3;
    const size_t nvox = (size_t)n * n * n, pvox = (size_t)P * P * P;
    ++ctx->generation;
    // Buffers of the same grid edge are kept (kernels still reading them are
    // waited for); others are freed before the new grids are allocated.
    SRT_HIP(hipDeviceSynchronize());
    if (ctx->grids && ctx->n != n) {
        SRT_HIP(hipFree(ctx->grids));
        ctx->grids = nullptr;
        if (ctx->grids64) SRT_HIP(hipFree(ctx->grids64));
        ctx->grids64 = nullptr;
        free_records(ctx);
    }
    if (ctx->layout != SRT_LAYOUT_BICUBIC) free_records(ctx);  // kept only to be rebuilt in place below
    ctx->grids64_valid = false;
    ctx->n = ctx->pad = 0;  // no snapshot until this upload completes
    if (!ctx->grids) SRT_HIP(ctx_malloc(ctx, &ctx->grids, 4 * pvox * sizeof(float)));
    float* const grids = ctx->grids;
    const int quantise = shift_format == SRT_SHIFT_F16_SCALED ? 1 : 0;
    auto abandon = [&](int code, const std::string& msg) {
        // the context holds no snapshot after a failed upload
        (void)hipDeviceSynchronize();
        (void)hipFree(ctx->grids);
        ctx->grids = nullptr;
        if (ctx->grids64) (void)hipFree(ctx->grids64);
        ctx->grids64 = nullptr;
        free_records(ctx);
        return fail(code, msg);
    };
    if (inputs_on_device) {
        const int blocks = (int)std::min<int64_t>(((int64_t)pvox + 255) / 256, 65536);
        hipLaunchKernelGGL(pad_grid_kernel, dim3(blocks), dim3(256), 0, 0, phi, (int64_t)1, 0, n, P, 0, grids);
        for (int a = 0; a < 3; ++a)
            hipLaunchKernelGGL(pad_grid_kernel, dim3(blocks), dim3(256), 0, 0, shift_xyz, (int64_t)3, a, n, P,
                               quantise, grids + (size_t)(1 + a) * pvox);
    } else {
        // host buffers: chunks through the transfer ring into the cubes'
        // interiors, then the periodic halo
        int rc = ensure_xfer(ctx);
        int slot = 0;
        if (rc == SRT_OK) rc = upload_host_field(ctx, phi, (int64_t)nvox, 1, 0, n, P, grids, slot);
        if (rc == SRT_OK)
            rc = upload_host_field(ctx, shift_xyz, 3 * (int64_t)nvox, 3, quantise, n, P, grids + pvox, slot);
        if (rc != SRT_OK) return abandon(rc, "srt_upload: " + g_last_error);
        const int64_t halo = 4 * (3 * (int64_t)P * P + 3 * (int64_t)P * n + 3 * (int64_t)n * n);
        const int blocks = (int)std::min<int64_t>((halo + 255) / 256, 65536);
        hipLaunchKernelGGL(halo_kernel, dim3(blocks), dim3(256), 0, ctx->xfer, grids, 4, n, P, (int64_t)pvox);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return abandon(SRT_ERR_DEVICE, std::string("srt_upload: ") + hipGetErrorName(e));
    ctx->n = n;
    ctx->pad = P;
    if (ctx->layout == SRT_LAYOUT_BICUBIC) {
        // the snapshot is uploaded either way; without records the context falls
        // back to the padded grids and says so
        const int rc = build_records(ctx);
        if (rc != SRT_OK) ctx->layout = SRT_LAYOUT_TAPS;
        return rc;
    }
    return SRT_OK;
}

int srt_set_layout(srt_context* ctx, int layout, int brick) {
    if (!ctx) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_set_layout: null context");
    if (layout != SRT_LAYOUT_TAPS && layout != SRT_LAYOUT_BICUBIC)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_set_layout: unknown layout");
    int shift = -1;
    for (int s = 0; s <= 6; ++s)
        if (brick == (1 << s)) shift = s;
    if (shift < 0) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_set_layout: brick must be a power of two in [1, 64]");
    SRT_HIP(hipSetDevice(ctx->device));
    SRT_HIP(hipDeviceSynchronize());
    const bool rebuild = layout == SRT_LAYOUT_BICUBIC && (!ctx->rec || shift != ctx->brick_shift);
    ctx->layout = layout;
    ctx->brick_shift = shift;
    if (layout == SRT_LAYOUT_TAPS) {
        free_records(ctx);
        return SRT_OK;
    }
    if (ctx->grids && rebuild) {
        const int rc = build_records(ctx);
        if (rc != SRT_OK) ctx->layout = SRT_LAYOUT_TAPS;
        return rc;
    }
    return SRT_OK;
}

int srt_get_layout(srt_context* ctx, int* layout, int* brick, int64_t* record_bytes) {
    if (!ctx) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_get_layout: null context");
    if (layout) *layout = ctx->layout;
    if (brick) *brick = 1 << ctx->brick_shift;
    if (record_bytes) *record_bytes = (int64_t)(4 * ctx->rec_field_floats * sizeof(float));
    return SRT_OK;
}

}  // extern "C"

namespace {

// srt_trace and srt_trace_noalloc.  With no_alloc the sorted lane order's
// lane -> slot map and sort space come from the device array `scratch`
// (the per-device address in g_scratch), and a request that would need an
// allocation (a larger sorted order, the FP64 grid copies) is refused.
int trace_entry(const char* who, bool no_alloc, srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count,
                const int64_t* pix_list, int n_steps, const double* observer, int precision, int ray_order,
                float* out, double* final_state, void* stream) {
    const std::string w(who);
    if (!ctx || !out) return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "null context or output");
    if (!ctx->grids) return fail(SRT_ERR_NOT_UPLOADED, w + ": " + "no snapshot uploaded");
    if (nside <= 0 || nside > (1 << 20)) return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "nside out of range");
    if (n_steps < 0) return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "n_steps < 0");
    const int64_t npix = 12 * nside * nside;
    // a pixel list may repeat pixels (any count); a RING range lies within the sky
    if (count < 0) return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "count < 0");
    if (!pix_list && (count > npix || pix_begin < 0 || pix_begin > npix - count))
        return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "pixel range outside the sky");
    if (ray_order != SRT_ORDER_NESTED && ray_order != SRT_ORDER_LINEAR)
        return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "unknown ray order");
    if (precision != SRT_PREC_FP64 && precision != SRT_PREC_MIXED && precision != SRT_PREC_REF)
        return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "unknown precision");
    SRT_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    // NESTED over the full sky at power-of-two Nside: lanes are NESTED pixels;
    // any other NESTED request sorts the rays by srt::ring_order_key (LINEAR
    // beyond the sort's ray limit)
    bool sorted = false;
    ray_order = effective_order(nside, pix_begin, count, pix_list != nullptr, ray_order, &sorted);
    uint32_t* lane_slot = nullptr;
    void* order_buf = ctx->order_buf;
    size_t order_bytes = ctx->order_bytes;
    if (sorted) {
        const size_t map_bytes = ((size_t)count * sizeof(uint32_t) + 255) / 256 * 256;
        const size_t work = srt::order_work_bytes(nside, count);
        if (!work) return fail(SRT_ERR_DEVICE, w + ": " + "lane-order work size");
        if (no_alloc) {
            if (count > kNoAllocMaxSortedRays || map_bytes + work > kScratchBytes)
                return fail(SRT_ERR_UNSUPPORTED,
                            w + ": a sorted lane order over " + std::to_string(count) + " rays needs " +
                                std::to_string(map_bytes + work) + " bytes of scratch; the fixed scratch holds " +
                                std::to_string(kScratchBytes) + " bytes (" + std::to_string(kNoAllocMaxSortedRays) +
                                " rays): use srt_trace, or SRT_ORDER_LINEAR");
            order_buf = scratch_address(ctx->device);
            if (!order_buf) return fail(SRT_ERR_DEVICE, w + ": no address of the scratch array on this device");
            order_bytes = kScratchBytes;
        } else if (map_bytes + work > ctx->order_bytes) {
            SRT_HIP(hipDeviceSynchronize());  // an earlier launch may still read the old map
            if (ctx->order_buf) SRT_HIP(hipFree(ctx->order_buf));
            ctx->order_buf = nullptr;
            ctx->order_bytes = 0;
            SRT_HIP(ctx_malloc(ctx, &ctx->order_buf, map_bytes + work));
            ctx->order_bytes = map_bytes + work;
            order_buf = ctx->order_buf;
            order_bytes = ctx->order_bytes;
        }
        lane_slot = static_cast<uint32_t*>(order_buf);
    }
    TraceArgs A;
    A.G.phi = ctx->grids;
    const size_t pvox = (size_t)ctx->pad * ctx->pad * ctx->pad;
    for (int a = 0; a < 3; ++a) A.G.b[a] = ctx->grids + (size_t)(1 + a) * pvox;
    if (precision == SRT_PREC_FP64) {
        if (no_alloc && !ctx->grids64)
            return fail(SRT_ERR_UNSUPPORTED, w + ": the FP64 kernel reads FP64 copies of the grids, which this "
                                                 "snapshot does not have yet (srt_trace with SRT_PREC_FP64 "
                                                 "builds them)");
        const int rc = ensure_grids64(ctx, st);
        if (rc != SRT_OK) return rc;
    }
    set_grids64(ctx, A.G);
    set_records(ctx, A.G);
    A.G.n = ctx->n;
    A.G.pad = ctx->pad;
    A.G.inv_n = 1.0 / (double)ctx->n;
    A.nside = nside;
    A.pix_begin = pix_begin;
    A.count = count;
    A.pix_list = pix_list;
    A.order = ray_order;
    A.n_steps = n_steps;
    for (int i = 0; i < 3; ++i) A.obs[i] = observer ? observer[i] : 0.5 * (double)ctx->n;
    A.out = out;
    A.state = final_state;
    A.xcd_groups = ctx->xcd_groups;
    A.lane_slot = lane_slot;
    A.lane_base = 0;
    SRT_HIP(hipEventRecord(ctx->ev_start, st));
    if (lane_slot) {  // the sort is part of the timed trace
        const size_t map_bytes = ((size_t)count * sizeof(uint32_t) + 255) / 256 * 256;
        SRT_HIP(srt::order_lanes(nside, pix_begin, count, pix_list, lane_slot,
                                 static_cast<char*>(order_buf) + map_bytes, order_bytes - map_bytes, st));
    }
    if (count > 0) {
        const bool pow2 = (ctx->n & (ctx->n - 1)) == 0;
        switch (precision) {
            case SRT_PREC_FP64:
                launch_trace<srt::FormFP64>(pow2, st, A, ctx->max_launch_rays);
                break;
            case SRT_PREC_MIXED:
                if (ctx->rec)
                    launch_trace<srt::FormMixedBicubic>(pow2, st, A, ctx->max_launch_rays);
                else
                    launch_trace<srt::FormMixed>(pow2, st, A, ctx->max_launch_rays);
                break;
            case SRT_PREC_REF:
                launch_trace<srt::FormRef>(pow2, st, A, ctx->max_launch_rays);
                break;
            default:
                return fail(SRT_ERR_INVALID_ARGUMENT, w + ": " + "unknown precision");
        }
        SRT_HIP(hipGetLastError());
    }
    SRT_HIP(hipEventRecord(ctx->ev_stop, st));
    ctx->timed = true;
    return SRT_OK;
}

}  // namespace

extern "C" {

int srt_trace(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
              int n_steps, const double* observer, int precision, int ray_order, float* out, double* final_state,
              void* stream) {
    return trace_entry("srt_trace", false, ctx, nside, pix_begin, count, pix_list, n_steps, observer, precision,
                       ray_order, out, final_state, stream);
}

int srt_trace_noalloc(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                      int n_steps, const double* observer, int precision, int ray_order, float* out,
                      double* final_state, void* stream) {
    return trace_entry("srt_trace_noalloc", true, ctx, nside, pix_begin, count, pix_list, n_steps, observer,
                       precision, ray_order, out, final_state, stream);
}

int srt_trace_adjoint(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                      int n_steps, const double* observer, int checkpoint_every, const double* map_grad,
                      double* maps64, double* grad_phi, double* grad_shift_xyz, void* stream) {
    if (!ctx) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: null context");
    if (!ctx->grids) return fail(SRT_ERR_NOT_UPLOADED, "srt_trace_adjoint: no snapshot uploaded");
    if (nside <= 0 || nside > (1 << 20)) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: nside out of range");
    if (n_steps < 1 || checkpoint_every < 1)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: n_steps and checkpoint_every must be >= 1");
    const int64_t npix = 12 * nside * nside;
    // a pixel list may repeat pixels (any count); a RING range lies within the sky
    if (count < 0) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: count < 0");
    if (!pix_list && (count > npix || pix_begin < 0 || pix_begin > npix - count))
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: pixel range outside the sky");
    if (map_grad && (!grad_phi || !grad_shift_xyz))
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: gradient outputs are null");
    if (!map_grad && !maps64) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_adjoint: nothing to compute");
    // at most 2^32 - 64 rays (rounded to whole wavefronts, below 2^32 lanes; the
    // launches themselves are split into chunks far smaller than that)
    if (count > (int64_t(1) << 32) - 64)
        return fail(SRT_ERR_UNSUPPORTED, "srt_trace_adjoint: more rays than one call holds (2^32 lanes)");
    SRT_HIP(hipSetDevice(ctx->device));
    const int K = std::min(checkpoint_every, n_steps);
    const int nseg = (n_steps + K - 1) / K;
    const size_t P = (size_t)ctx->pad, pvox = P * P * P;
    // rays per chunk: the chunk's stage records, checkpoints and dL/dy within kAdjChunkBytes
    const int64_t per_ray = (int64_t)sizeof(double) * ((int64_t)4 * K * kAdjRec + (int64_t)(nseg + 1) * 19 + 19);
    int64_t chunk = std::max<int64_t>(64, kAdjChunkBytes / per_ray / 64 * 64);
    if (ctx->adj_chunk_rays) chunk = std::min(chunk, ctx->adj_chunk_rays);
    chunk = std::min<int64_t>(chunk, (std::max<int64_t>(count, 1) + 63) / 64 * 64);
    const size_t n_ck = (size_t)chunk * (size_t)(nseg + 1) * 19, n_lam = (size_t)chunk * 19,
                 n_rec = map_grad ? (size_t)chunk * (size_t)(4 * K) * kAdjRec : 0, n_gr = map_grad ? 4 * pvox : 0;
    const size_t need = (n_ck + n_lam + n_rec + n_gr) * sizeof(double);
    if (need > ctx->adj_bytes) {
        SRT_HIP(hipDeviceSynchronize());  // an earlier call, on any stream, may still use the old buffers
        if (ctx->adj_buf) SRT_HIP(hipFree(ctx->adj_buf));
        ctx->adj_buf = nullptr;
        ctx->adj_bytes = 0;
        SRT_HIP(ctx_malloc(ctx, &ctx->adj_buf, need));
        ctx->adj_bytes = need;
    }
    hipStream_t st = (hipStream_t)stream;
    AdjArgs A;
    A.G.phi = ctx->grids;
    for (int a = 0; a < 3; ++a) A.G.b[a] = ctx->grids + (size_t)(1 + a) * pvox;
    A.G.n = ctx->n;
    A.G.pad = ctx->pad;
    A.G.inv_n = 1.0 / (double)ctx->n;
    {
        const int rc = ensure_grids64(ctx, st);
        if (rc != SRT_OK) return rc;
    }
    set_grids64(ctx, A.G);
    set_records(ctx, A.G);
    A.nside = nside;
    A.pix_begin = pix_begin;
    A.pix_list = pix_list;
    A.total = count;
    A.C = chunk;
    A.n_steps = n_steps;
    A.K = K;
    for (int i = 0; i < 3; ++i) A.obs[i] = observer ? observer[i] : 0.5 * (double)ctx->n;
    A.map_grad = map_grad;
    A.maps64 = maps64;
    A.lane_slot = nullptr;
    A.rec_from = map_grad ? (nseg - 1) * K : n_steps;  // the forward sweep records the last segment
    A.ckpt = ctx->adj_buf;
    A.lam = A.ckpt + n_ck;
    A.rec = map_grad ? A.lam + n_lam : nullptr;
    for (int f = 0; f < 4; ++f) A.grad[f] = map_grad ? A.lam + n_lam + n_rec + f * pvox : nullptr;
    if (map_grad) SRT_HIP(hipMemsetAsync(A.grad[0], 0, 4 * pvox * sizeof(double), st));
    // rays in NESTED order of their pixels (srt::order_lanes); the results do
    // not depend on it beyond the summation order of the gradient atomics
    size_t map_bytes = 0;
    if (count > 0 && count <= srt::kMaxOrderedRays) {
        map_bytes = ((size_t)count * sizeof(uint32_t) + 255) / 256 * 256;
        const size_t order_work = srt::order_work_bytes(nside, count);
        if (!order_work) return fail(SRT_ERR_DEVICE, "srt_trace_adjoint: lane-order work size");
        if (map_bytes + order_work > ctx->order_bytes) {
            SRT_HIP(hipDeviceSynchronize());  // an earlier launch may still read the old map
            if (ctx->order_buf) SRT_HIP(hipFree(ctx->order_buf));
            ctx->order_buf = nullptr;
            ctx->order_bytes = 0;
            SRT_HIP(ctx_malloc(ctx, &ctx->order_buf, map_bytes + order_work));
            ctx->order_bytes = map_bytes + order_work;
        }
        A.lane_slot = static_cast<const uint32_t*>(ctx->order_buf);
    }
    SRT_HIP(hipEventRecord(ctx->ev_start, st));
    if (A.lane_slot)
        SRT_HIP(srt::order_lanes(nside, pix_begin, count, pix_list, static_cast<uint32_t*>(ctx->order_buf),
                                 static_cast<char*>(ctx->order_buf) + map_bytes, ctx->order_bytes - map_bytes, st));
    for (int64_t l0 = 0; l0 < count; l0 += chunk) {
        A.lane0 = l0;
        A.count = std::min(chunk, count - l0);
        const unsigned fwd_blocks = (unsigned)((A.count + kAdjRays - 1) / kAdjRays);
        hipLaunchKernelGGL(adj_forward_kernel<false>, dim3(fwd_blocks), dim3(64), 0, st, A);
        SRT_HIP(hipGetLastError());
        if (!map_grad) continue;
        const int64_t tiles = (A.count + 63) / 64;
        for (int sgi = nseg - 1; sgi >= 0; --sgi) {
            A.seg = sgi;
            A.seg_steps = std::min(K, n_steps - sgi * K);
            if (sgi < nseg - 1)  // the forward sweep left the last segment's records
                hipLaunchKernelGGL(adj_forward_kernel<true>, dim3(fwd_blocks), dim3(64), 0, st, A);
            hipLaunchKernelGGL(adj_reverse_kernel, dim3((unsigned)tiles), dim3(kAdjRevBlock), 0, st, A);
            hipLaunchKernelGGL(adj_scatter_kernel, dim3((unsigned)(tiles * 4 * A.seg_steps)), dim3(64), 0, st, A);
            SRT_HIP(hipGetLastError());
        }
    }
    SRT_HIP(hipEventRecord(ctx->ev_stop, st));
    ctx->timed = true;
    if (map_grad) {
        const int n = ctx->n;
        const int64_t total = (int64_t)n * n * n;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(fold_grad_kernel, dim3(blocks), dim3(256), 0, st, A.grad[0], n, (int)P, grad_phi,
                           (int64_t)1, 0);
        for (int a = 0; a < 3; ++a)
            hipLaunchKernelGGL(fold_grad_kernel, dim3(blocks), dim3(256), 0, st, A.grad[1 + a], n, (int)P,
                               grad_shift_xyz, (int64_t)3, a);
        SRT_HIP(hipGetLastError());
    }
    return SRT_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batch plans ----
// One kernel launch of a plan: the items traced by the same kernel, with that
// launch's device tables.  A plan runs one launch per kernel its items need: in
// a mixed plan the items whose contexts hold bicubic records run the record
// kernel and the others the padded-cube kernel, and in every plan the items
// whose grid edge is a power of two run the kernels specialised for it.  Each
// item thus runs the kernel of its own srt_trace, and its maps equal that
// trace's.  48 uneven snapshots (edges 128 to 384) ran 68.1 ms in two launches
// against 68.6 ms in one on the general kernel (dispatch schedule), and 80.6
// against 86.7 ms (persistent schedule).
struct PlanLaunch {
    std::vector<int> items;                   // indices into srt_plan::items
    bool bicubic = false;                     // mixed launch over bicubic records
    bool pow2 = false;                        // every item's grid edge is a power of two
    BatchInst* d_inst = nullptr;              // the items' table
    unsigned long long* d_counter = nullptr;  // kMaxTileGroups tile counters (persistent schedule)
    int64_t* d_parts = nullptr;  // batch_dispatch_kernel's table: n + kMaxTileGroups * (n + 1), n items
    int dispatch_lanes = 0;      // lanes per dispatch tile the range table was built for
    int64_t dispatch_blocks = 0;  // workgroups of a dispatch launch
    int64_t total_tiles = 0;      // tiles of the persistent schedule
};

constexpr int kMaxPlanLaunches = 4;

struct srt_plan {
    int device = 0;
    int precision = SRT_PREC_MIXED;
    int n_inst = 0;
    int64_t total_rays = 0, total_ray_steps = 0;
    std::vector<srt_batch_item> items;
    PlanLaunch launches[kMaxPlanLaunches];
    int n_launches = 0;
    int groups = 8;                            // tile ranges (see batch_kernel)
    int order = SRT_ORDER_NESTED;              // lane order within each item
    void* order_buf = nullptr;                 // per-item lane -> slot maps + sort space
    size_t order_bytes = 0;
    std::vector<const uint32_t*> item_map;     // per item: its map in order_buf, or null
    std::vector<uint64_t> item_gen;            // per item: its context's generation when the table was built
    // dispatch by default: at 1024^3 / Nside 2048 / 512 steps a one-item plan takes 3.65 s against 3.77 s
    // persistent (the single trace launch: 3.66 s), and 48 uneven snapshots 68.1 against 80.6 ms
    int schedule = SRT_SCHEDULE_DISPATCH;
    int tile_rays = 64;  // one 64-ray group per claim: in-flight rays of an XCD stay as compact as a trace launch's
    int blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;  // the stream the graph was captured on
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timed = false;
};

namespace {

void plan_drop_graph(srt_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    p->graph_stream = nullptr;
}

// Per-item lane -> slot maps of a NESTED-ordered plan (srt::order_lanes),
// computed once per plan and order.
int plan_build_orders(srt_plan* p) {
    SRT_HIP(hipSetDevice(p->device));
    SRT_HIP(hipDeviceSynchronize());
    if (p->order_buf) SRT_HIP(hipFree(p->order_buf));
    p->order_buf = nullptr;
    p->order_bytes = 0;
    p->item_map.assign(p->n_inst, nullptr);
    if (p->order != SRT_ORDER_NESTED) return SRT_OK;
    std::vector<size_t> off(p->n_inst, 0);
    size_t maps = 0, work = 0;
    for (int i = 0; i < p->n_inst; ++i) {
        const srt_batch_item& it = p->items[i];
        if (it.count <= 0 || it.count > srt::kMaxOrderedRays) continue;
        off[i] = maps + 1;  // + 1: 0 marks an item without a map
        maps += ((size_t)it.count * sizeof(uint32_t) + 255) / 256 * 256;
        work = std::max(work, srt::order_work_bytes(it.nside, it.count));
    }
    if (!maps) return SRT_OK;
    SRT_HIP(hipMalloc(&p->order_buf, maps + work));
    p->order_bytes = maps + work;
    char* base = static_cast<char*>(p->order_buf);
    for (int i = 0; i < p->n_inst; ++i) {
        if (!off[i]) continue;
        const srt_batch_item& it = p->items[i];
        uint32_t* map = reinterpret_cast<uint32_t*>(base + off[i] - 1);
        SRT_HIP(srt::order_lanes(it.nside, it.pix_begin, it.count, nullptr, map, base + maps, work, nullptr));
        p->item_map[i] = map;
    }
    SRT_HIP(hipDeviceSynchronize());
    return SRT_OK;
}

// (re)builds the launches' device tables for the plan's tile size from the
// contexts' current buffers
int plan_build_table(srt_plan* p) {
    for (int i = 0; i < p->n_inst; ++i)
        if (!p->items[i].ctx->grids)
            return fail(SRT_ERR_NOT_UPLOADED, "plan item " + std::to_string(i) + ": context without snapshot");
    // the launches: a mixed plan's items whose contexts hold records, then the others, each split into the
    // items whose grid edge is a power of two and the others
    std::vector<int> members[2][2];  // [records][power of two]
    for (int i = 0; i < p->n_inst; ++i) {
        const int n = p->items[i].ctx->n;
        members[p->precision == SRT_PREC_MIXED && p->items[i].ctx->rec ? 1 : 0][(n & (n - 1)) == 0 ? 1 : 0].push_back(i);
    }
    p->n_launches = 0;
    for (int r = 1; r >= 0; --r)
        for (int q = 1; q >= 0; --q)
            if (!members[r][q].empty()) {
                PlanLaunch& L = p->launches[p->n_launches++];
                L.items = members[r][q];
                L.bicubic = r == 1;
            }
    SRT_HIP(hipSetDevice(p->device));
    SRT_HIP(hipDeviceSynchronize());
    for (int li = 0; li < p->n_launches; ++li) {
        PlanLaunch& L = p->launches[li];
        const int n = (int)L.items.size(), g = p->groups;
        L.pow2 = true;
        for (int idx : L.items) L.pow2 = L.pow2 && (p->items[idx].ctx->n & (p->items[idx].ctx->n - 1)) == 0;
        std::vector<BatchInst> h(n);
        int64_t tiles = 0;
        for (int j = 0; j < n; ++j) {
            const int i = L.items[j];
            const srt_batch_item& it = p->items[i];
            srt_context* c = it.ctx;
            BatchInst& b = h[j];
            const size_t pvox = (size_t)c->pad * c->pad * c->pad;
            b.G.phi = c->grids;
            for (int a = 0; a < 3; ++a) b.G.b[a] = c->grids + (size_t)(1 + a) * pvox;
            b.G.n = c->n;
            b.G.pad = c->pad;
            b.G.inv_n = 1.0 / (double)c->n;
            if (p->precision == SRT_PREC_FP64) {
                const int rc = ensure_grids64(c, nullptr);
                if (rc != SRT_OK) return rc;
            }
            set_grids64(c, b.G);
            set_records(c, b.G);
            b.nside = it.nside;
            b.pix_begin = it.pix_begin;
            b.count = it.count;
            b.n_steps = it.n_steps;
            for (int k = 0; k < 3; ++k) b.obs[k] = it.observer_set ? it.observer[k] : 0.5 * (double)c->n;
            b.out = it.out;
            b.lane_slot = p->item_map.empty() ? nullptr : p->item_map[i];
            b.tile_begin = tiles;
            const int64_t per_tile = (int64_t)p->tile_rays * (kBlock / 64);
            tiles += (it.count + per_tile - 1) / per_tile;
        }
        L.total_tiles = tiles;
        // range table of the dispatch schedule (batch_dispatch_kernel), for the
        // workgroup size of the form this launch runs
        const int lanes = L.bicubic ? 64 : kBlock;
        std::vector<int64_t> parts((size_t)n + (size_t)g * (n + 1));
        for (int j = 0; j < n; ++j) parts[j] = j;
        std::stable_sort(parts.begin(), parts.begin() + n, [&](int64_t a, int64_t b) {
            return p->items[L.items[a]].n_steps > p->items[L.items[b]].n_steps;
        });
        int64_t longest = 0;
        for (int x = 0; x < g; ++x) {
            int64_t* pb = parts.data() + n + (size_t)x * (n + 1);
            int64_t acc = 0;
            for (int j = 0; j < n; ++j) {
                pb[j] = acc;
                const int64_t t = (p->items[L.items[parts[j]]].count + lanes - 1) / lanes;
                acc += t * (x + 1) / g - t * x / g;
            }
            pb[n] = acc;
            longest = std::max(longest, acc);
        }
        if (longest * g * lanes >= (int64_t(1) << 32))
            return fail(SRT_ERR_UNSUPPORTED, "plan: the dispatch schedule's launch would hold 2^32 or more lanes");
        L.dispatch_lanes = lanes;
        L.dispatch_blocks = longest * g;
        SRT_HIP(hipMemcpy(L.d_inst, h.data(), h.size() * sizeof(BatchInst), hipMemcpyHostToDevice));
        SRT_HIP(hipMemcpy(L.d_parts, parts.data(), parts.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    p->item_gen.resize(p->n_inst);
    for (int i = 0; i < p->n_inst; ++i) p->item_gen[i] = p->items[i].ctx->generation;
    return SRT_OK;
}

// Rebuilds the table (and drops the captured graph, whose kernel may change
// with the layout or grid size) if any context's buffers were replaced since
// it was built: a new snapshot, records rebuilt or dropped, FP64 copies made.
int plan_refresh(srt_plan* p) {
    bool stale = p->item_gen.size() != (size_t)p->n_inst;
    for (int i = 0; i < p->n_inst && !stale; ++i) stale = p->items[i].ctx->generation != p->item_gen[i];
    if (!stale) return SRT_OK;
    plan_drop_graph(p);
    return plan_build_table(p);
}

// batch_kernel<F> for launch L of plan p, specialised when every item's grid edge is a power of two
template <class F>
void launch_batch(const srt_plan* p, const PlanLaunch& L, hipStream_t st) {
    const int n = (int)L.items.size();
    if (L.pow2)
        hipLaunchKernelGGL(batch_kernel<srt::Pow2Grid<F>>, dim3(p->blocks), dim3(kBlock), lds_bytes<F>(), st,
                           L.d_inst, n, L.total_tiles, p->tile_rays, p->groups, L.d_counter);
    else
        hipLaunchKernelGGL(batch_kernel<F>, dim3(p->blocks), dim3(kBlock), lds_bytes<F>(), st, L.d_inst, n,
                           L.total_tiles, p->tile_rays, p->groups, L.d_counter);
}

// batch_dispatch_kernel<F> for launch L of plan p (the dispatch schedule)
template <class F>
int launch_batch_dispatch(const srt_plan* p, const PlanLaunch& L, hipStream_t st) {
    constexpr int lanes = trace_lanes<F>();
    constexpr unsigned lds = lds_bytes<F, lanes>();
    if (lanes != L.dispatch_lanes)
        return fail(SRT_ERR_DEVICE, "plan: range table built for another workgroup size");
    if (L.dispatch_blocks == 0) {  // no rays: keep the launch sequence non-empty for graph capture
        hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(kMaxTileGroups), 0, st, L.d_counter, p->groups);
        return SRT_OK;
    }
    const dim3 grid((unsigned)L.dispatch_blocks);
    const int n = (int)L.items.size();
    if (L.pow2)
        hipLaunchKernelGGL(batch_dispatch_kernel<srt::Pow2Grid<F>>, grid, dim3(lanes), lds, st, L.d_inst, n,
                           L.d_parts, p->groups);
    else
        hipLaunchKernelGGL(batch_dispatch_kernel<F>, grid, dim3(lanes), lds, st, L.d_inst, n, L.d_parts, p->groups);
    return SRT_OK;
}

int plan_enqueue_dispatch(srt_plan* p, const PlanLaunch& L, hipStream_t st) {
    int rc;
    switch (p->precision) {
        case SRT_PREC_FP64:
            rc = launch_batch_dispatch<srt::FormFP64>(p, L, st);
            break;
        case SRT_PREC_MIXED:
            rc = L.bicubic ? launch_batch_dispatch<srt::FormMixedBicubic>(p, L, st)
                           : launch_batch_dispatch<srt::FormMixed>(p, L, st);
            break;
        default:
            rc = launch_batch_dispatch<srt::FormRef>(p, L, st);
            break;
    }
    if (rc != SRT_OK) return rc;
    SRT_HIP(hipGetLastError());
    return SRT_OK;
}

// The plan's launches in order on stream st.
int plan_enqueue(srt_plan* p, hipStream_t st) {
    for (int li = 0; li < p->n_launches; ++li) {
        const PlanLaunch& L = p->launches[li];
        if (p->schedule == SRT_SCHEDULE_DISPATCH) {
            const int rc = plan_enqueue_dispatch(p, L, st);
            if (rc != SRT_OK) return rc;
            continue;
        }
        hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(kMaxTileGroups), 0, st, L.d_counter, p->groups);
        switch (p->precision) {
            case SRT_PREC_FP64:
                launch_batch<srt::FormFP64>(p, L, st);
                break;
            case SRT_PREC_MIXED:
                if (L.bicubic)
                    launch_batch<srt::FormMixedBicubic>(p, L, st);
                else
                    launch_batch<srt::FormMixed>(p, L, st);
                break;
            default:
                launch_batch<srt::FormRef>(p, L, st);
                break;
        }
        SRT_HIP(hipGetLastError());
    }
    return SRT_OK;
}

int device_cus(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 256;
    return prop.multiProcessorCount;
}

}  // namespace

extern "C" {

int srt_plan_create(const srt_batch_item* items, int n_items, int precision, srt_plan** plan) {
    if (!plan || !items || n_items <= 0) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_create: bad arguments");
    *plan = nullptr;
    if (precision != SRT_PREC_FP64 && precision != SRT_PREC_MIXED && precision != SRT_PREC_REF)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_create: unknown precision");
    const int device = items[0].ctx ? items[0].ctx->device : -1;
    for (int i = 0; i < n_items; ++i) {
        const srt_batch_item& it = items[i];
        if (!it.ctx || !it.out) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_create: null context or output");
        if (!it.ctx->grids) return fail(SRT_ERR_NOT_UPLOADED, "srt_plan_create: context without snapshot");
        if (it.ctx->device != device)
            return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_create: contexts on different devices");
        if (it.nside <= 0 || it.nside > (1 << 20) || it.n_steps < 0 || it.count < 0 || it.pix_begin < 0 ||
            it.count > 12 * it.nside * it.nside || it.pix_begin > 12 * it.nside * it.nside - it.count)
            return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_create: item " + std::to_string(i) + " out of range");
    }
    SRT_HIP(hipSetDevice(device));
    srt_plan* p = new srt_plan();
    p->device = device;
    p->precision = precision;
    p->n_inst = n_items;
    p->items.assign(items, items + n_items);
    for (const auto& it : p->items) {
        p->total_rays += it.count;
        p->total_ray_steps += it.count * (int64_t)it.n_steps;
    }
    p->blocks = 2 * device_cus(device);
    // tables of every launch sized for all items (a layout change can move items between launches)
    hipError_t e = hipSuccess;
    for (PlanLaunch& L : p->launches) {
        if (e == hipSuccess) e = hipMalloc(&L.d_inst, sizeof(BatchInst) * n_items);
        if (e == hipSuccess) e = hipMalloc(&L.d_counter, kMaxTileGroups * sizeof(unsigned long long));
        if (e == hipSuccess)
            e = hipMalloc(&L.d_parts, ((size_t)n_items + (size_t)kMaxTileGroups * (n_items + 1)) * sizeof(int64_t));
    }
    if (e == hipSuccess) e = hipEventCreate(&p->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&p->ev_stop);
    if (e != hipSuccess) {
        srt_plan_destroy(p);
        return fail(e == hipErrorOutOfMemory ? SRT_ERR_OUT_OF_MEMORY : SRT_ERR_DEVICE,
                    std::string("srt_plan_create: ") + hipGetErrorName(e));
    }
    int rc = plan_build_orders(p);
    if (rc == SRT_OK) rc = plan_build_table(p);
    if (rc != SRT_OK) {
        srt_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return SRT_OK;
}

int srt_plan_set_order(srt_plan* plan, int ray_order, int tile_groups) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_set_order: null plan");
    if (ray_order != SRT_ORDER_LINEAR && ray_order != SRT_ORDER_NESTED)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_set_order: unknown ray order");
    if (tile_groups < 1 || tile_groups > kMaxTileGroups)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_set_order: tile_groups must be in [1, 64]");
    SRT_HIP(hipSetDevice(plan->device));
    SRT_HIP(hipDeviceSynchronize());
    plan_drop_graph(plan);
    plan->groups = tile_groups;
    if (ray_order != plan->order || plan->item_map.size() != (size_t)plan->n_inst) {
        plan->order = ray_order;
        const int rc = plan_build_orders(plan);
        if (rc != SRT_OK) return rc;
    }
    return plan_build_table(plan);
}

int srt_plan_destroy(srt_plan* plan) {
    if (!plan) return SRT_OK;
    (void)hipSetDevice(plan->device);
    plan_drop_graph(plan);
    for (PlanLaunch& L : plan->launches) {
        if (L.d_inst) (void)hipFree(L.d_inst);
        if (L.d_counter) (void)hipFree(L.d_counter);
        if (L.d_parts) (void)hipFree(L.d_parts);
    }
    if (plan->order_buf) (void)hipFree(plan->order_buf);
    if (plan->ev_start) (void)hipEventDestroy(plan->ev_start);
    if (plan->ev_stop) (void)hipEventDestroy(plan->ev_stop);
    delete plan;
    return SRT_OK;
}

int srt_plan_configure(srt_plan* plan, int tile_rays, int blocks) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_configure: null plan");
    if (tile_rays < 64 || tile_rays > 65536 || (tile_rays & 63) || blocks < 1 || blocks > (1 << 20))
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_configure: tile_rays must be a multiple of 64 in "
                                              "[64, 65536] and blocks in [1, 2^20]");
    SRT_HIP(hipSetDevice(plan->device));
    SRT_HIP(hipDeviceSynchronize());
    plan_drop_graph(plan);
    plan->tile_rays = tile_rays;
    plan->blocks = blocks;
    return plan_build_table(plan);
}

int srt_plan_set_schedule(srt_plan* plan, int schedule) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_set_schedule: null plan");
    if (schedule != SRT_SCHEDULE_PERSISTENT && schedule != SRT_SCHEDULE_DISPATCH)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_set_schedule: unknown schedule");
    SRT_HIP(hipSetDevice(plan->device));
    SRT_HIP(hipDeviceSynchronize());
    plan_drop_graph(plan);
    plan->schedule = schedule;
    return SRT_OK;
}

int srt_plan_get_schedule(srt_plan* plan, int* schedule) {
    if (!plan || !schedule) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_get_schedule: null pointer");
    *schedule = plan->schedule;
    return SRT_OK;
}

int srt_plan_get_config(srt_plan* plan, int* tile_rays, int* blocks, int64_t* total_ray_steps) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_get_config: null plan");
    if (tile_rays) *tile_rays = plan->tile_rays;
    if (blocks) *blocks = plan->blocks;
    if (total_ray_steps) *total_ray_steps = plan->total_ray_steps;
    return SRT_OK;
}

int srt_plan_launch(srt_plan* plan, int use_graph, void* stream) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_launch: null plan");
    SRT_HIP(hipSetDevice(plan->device));
    hipStream_t st = (hipStream_t)stream;
    if (use_graph && !st)
        return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_launch: graph replay needs a non-default stream");
    const int rc_refresh = plan_refresh(plan);
    if (rc_refresh != SRT_OK) return rc_refresh;
    if (use_graph && plan->exec && plan->graph_stream != st) plan_drop_graph(plan);
    if (use_graph && !plan->exec) {
        SRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = plan_enqueue(plan, st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != SRT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        SRT_HIP(e);
        plan->graph = g;
        SRT_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
        plan->graph_stream = st;
    }
    SRT_HIP(hipEventRecord(plan->ev_start, st));
    if (use_graph) {
        SRT_HIP(hipGraphLaunch(plan->exec, st));
    } else {
        const int rc = plan_enqueue(plan, st);
        if (rc != SRT_OK) return rc;
    }
    SRT_HIP(hipEventRecord(plan->ev_stop, st));
    plan->timed = true;
    return SRT_OK;
}

int srt_plan_last_ms(srt_plan* plan, float* ms) {
    if (!plan || !ms) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_last_ms: null pointer");
    if (!plan->timed) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_last_ms: no launch recorded");
    SRT_HIP(hipSetDevice(plan->device));
    SRT_HIP(hipEventSynchronize(plan->ev_stop));
    SRT_HIP(hipEventElapsedTime(ms, plan->ev_start, plan->ev_stop));
    return SRT_OK;
}

int srt_plan_autotune(srt_plan* plan, void* stream, int* tile_rays, int* blocks) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_autotune: null plan");
    const int cus = device_cus(plan->device);
    const int tiles[] = {64, 128, 256, 512, 1024};
    const int per_cu[] = {1, 2, 4};
    const int group_counts[] = {1, 8};
    float best = 0.0f;
    int bt = plan->tile_rays, bb = plan->blocks, bg = plan->groups;
    bool have = false;
    int rc = srt_plan_set_schedule(plan, SRT_SCHEDULE_PERSISTENT);
    if (rc != SRT_OK) return rc;
    for (int g : group_counts) {
        rc = srt_plan_set_order(plan, plan->order, g);
        if (rc != SRT_OK) return rc;
        for (int t : tiles)
            for (int w : per_cu) {
                rc = srt_plan_configure(plan, t, w * cus);
                if (rc == SRT_OK) rc = srt_plan_launch(plan, 0, stream);  // warm-up
                float ms = 0.0f;
                if (rc == SRT_OK) rc = srt_plan_launch(plan, 0, stream);
                if (rc == SRT_OK) rc = srt_plan_last_ms(plan, &ms);
                if (rc != SRT_OK) return rc;
                if (!have || ms < best) {
                    best = ms;
                    bt = t;
                    bb = w * cus;
                    bg = g;
                    have = true;
                }
            }
    }
    // the dispatch schedule (no tile size or workgroup count to choose)
    int bs = SRT_SCHEDULE_PERSISTENT, dg = bg;
    rc = srt_plan_set_schedule(plan, SRT_SCHEDULE_DISPATCH);
    for (int g : group_counts) {
        if (rc == SRT_OK) rc = srt_plan_set_order(plan, plan->order, g);
        if (rc == SRT_OK) rc = srt_plan_launch(plan, 0, stream);  // warm-up
        float ms = 0.0f;
        if (rc == SRT_OK) rc = srt_plan_launch(plan, 0, stream);
        if (rc == SRT_OK) rc = srt_plan_last_ms(plan, &ms);
        if (rc != SRT_OK) return rc;
        if (ms < best) {
            best = ms;
            bs = SRT_SCHEDULE_DISPATCH;
            dg = g;
        }
    }
    rc = srt_plan_set_schedule(plan, bs);
    if (rc == SRT_OK) rc = srt_plan_set_order(plan, plan->order, bs == SRT_SCHEDULE_DISPATCH ? dg : bg);
    if (rc == SRT_OK) rc = srt_plan_configure(plan, bt, bb);
    if (rc != SRT_OK) return rc;
    if (tile_rays) *tile_rays = bt;
    if (blocks) *blocks = bb;
    return SRT_OK;
}

int srt_plan_autotune_brick(srt_plan* plan, void* stream, int* brick) {
    if (!plan) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_plan_autotune_brick: null plan");
    SRT_HIP(hipSetDevice(plan->device));
    int rc = plan_refresh(plan);
    if (rc != SRT_OK) return rc;
    bool all_records = plan->n_launches > 0;
    for (int li = 0; li < plan->n_launches; ++li) all_records = all_records && plan->launches[li].bicubic;
    if (!all_records)
        return fail(SRT_ERR_INVALID_ARGUMENT,
                    "srt_plan_autotune_brick: needs a mixed-precision plan whose contexts hold bicubic records");
    std::vector<srt_context*> ctxs;
    int min_n = 1 << 30;
    for (const auto& it : plan->items) {
        if (std::find(ctxs.begin(), ctxs.end(), it.ctx) == ctxs.end()) ctxs.push_back(it.ctx);
        min_n = std::min(min_n, it.ctx->n);
    }
    // every context in bricks of edge b; a context whose records do not fit
    // falls back to the padded grids, and the candidate is skipped
    auto set_all = [&](int b) {
        for (srt_context* c : ctxs) {
            const int r = srt_set_layout(c, SRT_LAYOUT_BICUBIC, b);
            if (r != SRT_OK) return r;
        }
        return (int)SRT_OK;
    };
    const int candidates[] = {1, 4, 8, 16, 32};
    float best = 0.0f;
    int bb = 0;
    for (int b : candidates) {
        if (b > min_n) continue;
        rc = set_all(b);
        if (rc == SRT_ERR_OUT_OF_MEMORY) continue;
        if (rc == SRT_OK) rc = srt_plan_launch(plan, 0, stream);  // warm-up; rebuilds the table
        float ms = 0.0f;
        if (rc == SRT_OK) rc = srt_plan_launch(plan, 0, stream);
        if (rc == SRT_OK) rc = srt_plan_last_ms(plan, &ms);
        if (rc != SRT_OK) return rc;
        if (!bb || ms < best) {
            best = ms;
            bb = b;
        }
    }
    if (!bb) return fail(SRT_ERR_OUT_OF_MEMORY, "srt_plan_autotune_brick: no brick edge fits");
    rc = set_all(bb);
    if (rc == SRT_OK) rc = plan_refresh(plan);
    if (rc != SRT_OK) return rc;
    if (brick) *brick = bb;
    return SRT_OK;
}

int srt_last_trace_ms(srt_context* ctx, float* ms) {
    if (!ctx || !ms) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_last_trace_ms: null pointer");
    if (!ctx->timed) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_last_trace_ms: no trace recorded");
    SRT_HIP(hipSetDevice(ctx->device));
    SRT_HIP(hipEventSynchronize(ctx->ev_stop));
    SRT_HIP(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
    return SRT_OK;
}

int srt_trace_host(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                   int n_steps, const double* observer, int precision, int ray_order, float* out,
                   double* final_state) {
    if (!ctx || !out) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_host: null context or output");
    if (count < 0) return fail(SRT_ERR_INVALID_ARGUMENT, "srt_trace_host: count < 0");
    SRT_HIP(hipSetDevice(ctx->device));
    const size_t c = (size_t)count;
    float* d_out = nullptr;
    int64_t* d_pix = nullptr;
    double* d_state = nullptr;
    int rc = SRT_OK;
    auto cleanup = [&]() {
        if (d_out) (void)hipFree(d_out);
        if (d_pix) (void)hipFree(d_pix);
        if (d_state) (void)hipFree(d_state);
    };
    hipError_t e = ctx_malloc(ctx, &d_out, std::max<size_t>(1, 6 * c) * sizeof(float));
    if (e == hipSuccess && pix_list) e = ctx_malloc(ctx, &d_pix, std::max<size_t>(1, c) * sizeof(int64_t));
    if (e == hipSuccess && final_state) e = ctx_malloc(ctx, &d_state, std::max<size_t>(1, 19 * c) * sizeof(double));
    if (e == hipSuccess && pix_list && c) e = hipMemcpy(d_pix, pix_list, c * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cleanup();
        return fail(e == hipErrorOutOfMemory ? SRT_ERR_OUT_OF_MEMORY : SRT_ERR_DEVICE,
                    std::string("srt_trace_host: ") + hipGetErrorName(e));
    }
    rc = srt_trace(ctx, nside, pix_begin, count, d_pix, n_steps, observer, precision, ray_order, d_out, d_state,
                   nullptr);
    if (rc == SRT_OK) {
        e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = fail(SRT_ERR_DEVICE, std::string("srt_trace_host: ") + hipGetErrorName(e));
    }
    if (rc == SRT_OK && c) rc = download_to_host(ctx, out, d_out, 6 * c * sizeof(float));
    if (rc == SRT_OK && c && final_state) rc = download_to_host(ctx, final_state, d_state, 19 * c * sizeof(double));
    cleanup();
    return rc;
}

}  // extern "C"
