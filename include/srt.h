/*
 * srt.h -- C ABI of the MI355X snapshot ray tracer (libsrt.so).
 *
 * The reference (JulianAdamek/snapshot-raytracer) has no library API: its hot
 * path is the CUDA kernel
 *     raytracing_kernel(const float* data, const __half* B1, B2, B3,
 *                       const int* dims, int64_t Nside, int64_t iend,
 *                       int n_steps, double* theta, double* phi,
 *                       float* dA, eps1, eps2, rotation, defl1, defl2)
 * (snapshot-raytracer.cu:552-555), fed by main() (.cu:600-998), which owns
 * every device buffer, uploads the potential and the shift (.cu:721-855),
 * computes the RING pixel angles on the host (.cu:857-892), launches the
 * kernel (.cu:897-913) and copies the six maps back (.cu:922-987).  The entry
 * points below replace those steps one for one; each names the reference
 * lines it stands in for.
 *
 * Conventions: every function returns an srt_status (0 = success) and never
 * exits the process.  Calls on one context are serialised by the caller;
 * different contexts are independent.  Grids are x-fastest:
 * value(ix, iy, iz) at ix + nx * (iy + ny * iz), which is HDF5's C order with
 * x the last HDF5 axis (.cu:220, 253-256).
 */
#ifndef SRT_H
#define SRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct srt_context srt_context;

typedef enum {
    SRT_OK = 0,
    SRT_ERR_INVALID_ARGUMENT = 1, /* bad size, index, enum or null pointer */
    SRT_ERR_DEVICE = 2,           /* a HIP runtime call failed */
    SRT_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed */
    SRT_ERR_NOT_UPLOADED = 4,     /* srt_trace before srt_upload */
    SRT_ERR_SHAPE = 5,            /* non-cubic or inconsistent grid dimensions */
    SRT_ERR_UNSUPPORTED = 6       /* valid request the library cannot serve */
} srt_status;

/* Arithmetic of the ray kernel. */
typedef enum {
    /* Everything in FP64: interpolation, right-hand side, state. */
    SRT_PREC_FP64 = 0,
    /* Interpolation and right-hand side (geodesic, Jacobi map, shear) in FP32
     * with FP32 accumulation; the 19-component state and the RK4 combination
     * in FP64, with compensated summation for the position and dA. */
    SRT_PREC_MIXED = 1,
    /* The reference GPU code's arithmetic: FP32 interpolation with the cell
     * offset taken as float(x) - float(floor(x)) (.cu:228-230), FP32
     * shift-coupling intermediates (.cu:366-396), FP64 state (.cu:513-549). */
    SRT_PREC_REF = 2
} srt_precision;

/* Device storage of the shift vector field. */
typedef enum {
    SRT_SHIFT_F32 = 0,       /* float, as read from the file */
    SRT_SHIFT_F16_SCALED = 1 /* binary16 of B * 2^27, the reference's storage (.cu:585-597, 241-244) */
} srt_shift_format;

/* Device layout of the snapshot (no counterpart in the reference, which
 * keeps the potential as a linear float array and the shift as three planar
 * binary16 arrays, .cu:738-855). */
typedef enum {
    /* Padded periodic cubes of grid values (4 (n+3)^3 floats); every form
     * interpolates from them. */
    SRT_LAYOUT_TAPS = 0,
    /* In addition, per field, cell column and plane, the 16 monomial
     * coefficients of the Catmull-Rom bicubic of the cell (64 bytes; the
     * 4 x 4 x 4 stencil of a cell is 256 contiguous bytes), columns tiled in
     * brick x brick bricks: 4 * 64 * n^2 * (n + 3) bytes (34.6 GB at 512^3).
     * SRT_PREC_MIXED traces (single and batched) read these records; the
     * other forms keep reading the cubes. */
    SRT_LAYOUT_BICUBIC = 1
} srt_layout;

/* Order in which rays are assigned to GPU lanes (outputs always land in the
 * caller's slot order). */
typedef enum {
    SRT_ORDER_LINEAR = 0, /* lane i traces slot i */
    SRT_ORDER_NESTED = 1  /* lanes walk the rays in NESTED order of their pixels (compact sky
                             patches per wavefront; for a partial sky, a pixel list or a Nside
                             that is not a power of two the rays are first sorted by base face
                             and interleaved in-face coordinates on the device); results are
                             written to the caller's slots, bit-identical to LINEAR */
} srt_ray_order;

/* Replaces the device selection and allocation bookkeeping of main()
 * (.cu:690-772).  device = HIP device ordinal.  Environment read here:
 * SRT_MAX_ALLOC_BYTES (largest single device allocation the context may make;
 * larger requests fail with SRT_ERR_OUT_OF_MEMORY, for exercising those paths),
 * SRT_XCD_GROUPS (workgroup grouping of trace launches, default 8) and
 * SRT_MAX_LAUNCH_RAYS (rays per trace launch, default and maximum 2^31: one
 * launch may not hold 2^32 lanes, so a larger trace is split into several
 * launches; smaller values exercise the split) and SRT_ADJ_CHUNK_RAYS (largest
 * ray chunk of srt_trace_adjoint, default: what fits its 8 GiB of work
 * buffers; smaller values exercise the chunk loop). */
int srt_create(int device, srt_context** ctx);

/* Frees every device buffer of the context (.cu:917-920, 989-998). */
int srt_destroy(srt_context* ctx);

/* Static text for a status code. */
const char* srt_strerror(int status);

/* Text of the last failure on the calling thread (includes HIP error names). */
const char* srt_last_error(void);

/* Uploads one snapshot (.cu:738-855: the potential, and the shift converted to
 * its device format by vector3_to_half, .cu:585-597).
 *   phi:       nx*ny*nz floats (the potential, HDF5 dataset /data, .cu:640)
 *   shift_xyz: nx*ny*nz*3 floats, interleaved (B_x, B_y, B_z) per voxel
 *              (HDF5 /data of H5T_ARRAY{3} float, .cu:676, 89-186)
 *   inputs_on_device: 0 = host pointers, 1 = device pointers of this context's device
 *   shift_format: srt_shift_format
 * The grid must be cubic (nx == ny == nz, 4 <= nx <= 2048); the reference
 * silently assumes equal dimensions for both files (.cu:645, 676, 752-772).
 * Replaces any previous snapshot of the context (a snapshot of the same edge
 * reuses its device buffers).  Host inputs go through the context's pinned
 * transfer ring (3 x 16 MiB of pinned host memory and as much device memory,
 * allocated by the first host upload or srt_trace_host and kept until
 * srt_destroy), copied into it by 4 host threads; the caller's buffers need
 * not be pinned.  If the context's layout is SRT_LAYOUT_BICUBIC and the
 * records do not fit, the snapshot is still uploaded, the context falls back
 * to SRT_LAYOUT_TAPS and the call returns SRT_ERR_OUT_OF_MEMORY.  Rejected
 * arguments leave the previous snapshot in place; any other failure leaves
 * the context without a snapshot. */
int srt_upload(srt_context* ctx, const float* phi, const float* shift_xyz, int64_t nx, int64_t ny, int64_t nz,
               int inputs_on_device, int shift_format);

/* Selects the device layout of this context's snapshot (srt_layout) and the
 * column brick edge (a power of two in [1, 64]; 1 = row-major columns).
 * Builds the layout's data now if a snapshot is uploaded, else at the next
 * srt_upload; the setting persists across uploads.  Replaces the reference's
 * fixed device formats (.cu:738-855).  Fails with SRT_ERR_OUT_OF_MEMORY
 * (and falls back to SRT_LAYOUT_TAPS) if the records do not fit. */
int srt_set_layout(srt_context* ctx, int layout, int brick);

/* Current layout, brick edge and bytes of bicubic records held (0 if none). */
int srt_get_layout(srt_context* ctx, int* layout, int* brick, int64_t* record_bytes);

/* Traces rays and writes the six maps (raytracing_kernel, .cu:552-583, with
 * the pixel angles of .cu:857-892 computed on the device).
 *   Ray slot i traces RING pixel pix_list[i] if pix_list is not null (device
 *   pointer, count int64 values), else pixel pix_begin + i (.cu:708-717).
 *   observer: 3 doubles in grid units (host pointer); null = (nx, ny, nz) / 2 (.cu:560).
 *   count: rays; the library cannot see the output buffers, so the caller
 *        guarantees they hold count rays (a trace of 2^32 rays or more runs
 *        as several launches).  A RING range [pix_begin, pix_begin + count)
 *        must lie within the sky; a pixel list may have any length and repeat
 *        pixels, and its invalid entries give NaN maps.
 *   out: device pointer, 6 * count floats, map-major: out[m * count + i] with
 *        m = dA, ellipticity1, ellipticity2, rotation, deflection1, deflection2
 *        (.cu:577-582).
 *   final_state: optional device pointer, count * 19 doubles: the state after
 *        the last step (x[3], n[3], e1[3], e2[3], dA, dA', sigma1, sigma2,
 *        eps1, eps2, rotation) in the precision of the kernel.
 *   stream: hipStream_t to launch on (null = the default stream).  The call is
 *        asynchronous with respect to the host.
 *   ray_order: srt_ray_order.  SRT_ORDER_NESTED other than over the full sky
 *        at a power-of-two Nside keeps a lane -> slot map (4 bytes per ray)
 *        and sort space in the context: overlapping such calls on one
 *        context must share a stream.  Such a sorted order covers at most
 *        2^31 - 1 rays; a larger request runs in SRT_ORDER_LINEAR order
 *        (srt_effective_order), with the same maps. */
int srt_trace(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
              int n_steps, const double* observer, int precision, int ray_order, float* out,
              double* final_state, void* stream);

/* The lane order srt_trace runs for a request (no counterpart in the
 * reference): *effective = SRT_ORDER_NESTED or SRT_ORDER_LINEAR for a trace
 * of `count` rays from pix_begin (or of a pixel list when has_pix_list != 0)
 * asked for in `ray_order`.  Host-only: needs no device or context. */
int srt_effective_order(int64_t nside, int64_t pix_begin, int64_t count, int has_pix_list, int ray_order,
                        int* effective);

/* srt_trace without device allocations.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and maps as srt_trace.  The scratch of a sorted lane
 * order (SRT_ORDER_NESTED other than over the full sky at a power-of-two
 * Nside) lives in a fixed 48 MiB device array that the library holds on every
 * GPU it runs on until the process ends; it covers at most 2^20 rays, and a
 * larger sorted request fails with SRT_ERR_UNSUPPORTED (use srt_trace, or
 * SRT_ORDER_LINEAR, which needs no scratch).  An SRT_PREC_FP64 trace needs
 * the snapshot's FP64 grid copies, which an srt_trace with SRT_PREC_FP64
 * builds; without them this call fails with SRT_ERR_UNSUPPORTED.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (any
 * context, any stream) race. */
int srt_trace_noalloc(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                      int n_steps, const double* observer, int precision, int ray_order, float* out,
                      double* final_state, void* stream);

/* Synchronous host-memory variant of srt_trace for callers without device
 * buffers (the CLI): pix_list, out (6 * count floats) and final_state are host
 * pointers; pix_list and final_state may be null.  The outputs come back
 * through the context's transfer ring (see srt_upload). */
int srt_trace_host(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                   int n_steps, const double* observer, int precision, int ray_order, float* out,
                   double* final_state);

/* Adjoint of the FP64 trace (no counterpart in the reference; the map-loss
 * gradient extension of raytracing_kernel, .cu:552-583).  For the rays of
 * srt_trace (same pix_begin/count/pix_list/observer meaning), runs the FP64
 * integrator, storing the state every checkpoint_every steps, and returns:
 *   maps64 (optional): the six maps in FP64, map-major 6 * count;
 *   if map_grad (6 * count doubles, dL/d(map), map-major) is not null:
 *     grad_phi       (nx*ny*nz doubles)   dL/d(potential grid value),
 *     grad_shift_xyz (nx*ny*nz*3 doubles) dL/d(shift grid value), interleaved
 *   by a reverse sweep that recomputes each checkpoint segment's stages,
 *   propagates dL/dy through RK4 steps, right-hand sides and tricubic
 *   interpolations, and adds the tap gradients with FP64 atomics (summed per
 *   wavefront in LDS first; the summation order, and so the last bits of the
 *   gradients, may change from call to call).
 * Rays go to lanes in NESTED order of their pixels, as in srt_trace, and are
 * processed in chunks whose work buffers fit 8 GiB.  All pointers are device
 * pointers; asynchronous on `stream`.  Work buffers (checkpoints, one
 * segment's stage records, dL/dy, padded gradient grids, the lane map and its
 * sort space) are owned by the context and reused: overlapping calls on one
 * context must share a stream.  The gradient is with respect to the grid values the
 * context holds (the binary16-quantised ones for SRT_SHIFT_F16_SCALED). */
int srt_trace_adjoint(srt_context* ctx, int64_t nside, int64_t pix_begin, int64_t count, const int64_t* pix_list,
                      int n_steps, const double* observer, int checkpoint_every, const double* map_grad,
                      double* maps64, double* grad_phi, double* grad_shift_xyz, void* stream);

/* ---- batched launch (no counterpart in the reference, whose unit
 * of work is one process per snapshot and batch, README.md:41-42) ---- */

/* One instance of a batched trace: a context (one uploaded snapshot), a RING
 * pixel range, a step count, an observer and a device output of 6 * count
 * floats (map-major, as srt_trace).  observer_set = 0 means the grid centre. */
typedef struct {
    srt_context* ctx;
    int64_t nside;
    int64_t pix_begin;
    int64_t count;
    int n_steps;
    int observer_set;
    double observer[3];
    float* out;
} srt_batch_item;

typedef struct srt_plan srt_plan;

/* Builds a plan for tracing all items in one launch: tiles of consecutive
 * rays of one item are handed to workgroups (srt_plan_schedule), so items of
 * different grid size, step count and ray count share the GPU.
 * All contexts must be on one device and stay alive while the plan is used.
 * A plan follows its contexts: after srt_upload or srt_set_layout on one of
 * them, the next launch rebuilds the plan's device table (and its graph) from
 * the contexts' current buffers, so it traces the new snapshot or layout.
 * Each item's output equals its own srt_trace bit for bit: the plan runs one
 * launch per kernel its items need (in a SRT_PREC_MIXED plan, the record
 * kernel for items whose contexts hold bicubic records and the padded-cube
 * kernel for the others; for items whose grid edge is a power of two, the
 * kernels specialised for it). */
int srt_plan_create(const srt_batch_item* items, int n_items, int precision, srt_plan** plan);
int srt_plan_destroy(srt_plan* plan);

/* Lane order within each item (srt_ray_order; default SRT_ORDER_NESTED: each
 * item's rays sorted into compact sky patches once, when set) and the number
 * of tile ranges the launch splits the work into (1..64; default 8, one per
 * XCD, so neighbouring tiles share an L2; a persistent workgroup then helps
 * the other ranges).  Outputs do not depend on either. */
int srt_plan_set_order(srt_plan* plan, int ray_order, int tile_groups);

/* How a plan's launch hands out its rays. */
typedef enum {
    /* Persistent 256-lane workgroups claim tiles from atomic counters
     * (srt_plan_configure sets the tile size and workgroup count). */
    SRT_SCHEDULE_PERSISTENT = 0,
    /* One workgroup per tile of consecutive lanes of one item (64 lanes for
     * mixed plans over bicubic records, else 256), balanced by the hardware
     * dispatcher as a single srt_trace launch is; each tile range (see
     * srt_plan_set_order) holds the same share of every item, items with
     * more steps first. */
    SRT_SCHEDULE_DISPATCH = 1
} srt_plan_schedule;

/* Schedule of the plan's launches (srt_plan_schedule; default
 * SRT_SCHEDULE_DISPATCH, which measured faster on every workload tried: it
 * has no claim barriers and runs one-wavefront workgroups for mixed plans over
 * records).  Outputs do not depend on it. */
int srt_plan_set_schedule(srt_plan* plan, int schedule);
int srt_plan_get_schedule(srt_plan* plan, int* schedule);

/* Rays per wavefront per claim (multiple of 64; default 64: a 256-lane
 * workgroup claims 4 x 64 neighbouring rays at a time) and number of
 * persistent 256-lane workgroups (default 2 per CU). */
int srt_plan_configure(srt_plan* plan, int tile_rays, int blocks);
int srt_plan_get_config(srt_plan* plan, int* tile_rays, int* blocks, int64_t* total_ray_steps);

/* Times every persistent (tile ranges, tile_rays, blocks per CU) candidate
 * (1 or 8 ranges, 64..1024 rays, 1/2/4 blocks per CU) and the dispatch
 * schedule with 1 or 8 ranges on `stream`, and keeps the fastest; returns the
 * persistent tile size and block count (srt_plan_get_config), and the plan
 * keeps the schedule (srt_plan_get_schedule) and range count. */
int srt_plan_autotune(srt_plan* plan, void* stream, int* tile_rays, int* blocks);

/* Times the plan with its contexts' bicubic records rebuilt in column bricks
 * of edge 1, 4, 8, 16 and 32 (those not larger than the smallest grid edge;
 * an edge whose records do not fit is skipped) and keeps the fastest: every
 * context of the plan is left in SRT_LAYOUT_BICUBIC with that edge, so later
 * srt_trace calls on those contexts use it too.  Needs an SRT_PREC_MIXED plan
 * whose contexts all hold records.  Outputs do not depend on the edge.
 * Synchronous (the records are rebuilt between timed launches). */
int srt_plan_autotune_brick(srt_plan* plan, void* stream, int* brick);

/* Launches the plan on `stream`.  use_graph = 1 captures the launch sequence
 * into a HIP graph on first use (and after srt_plan_configure) and replays it;
 * it needs a non-default stream.  Asynchronous; srt_plan_last_ms times it. */
int srt_plan_launch(srt_plan* plan, int use_graph, void* stream);
int srt_plan_last_ms(srt_plan* plan, float* ms);

/* Milliseconds between HIP events recorded on the launch stream just before
 * and just after the kernel launches of the last srt_trace call of this
 * context (blocks until that work is done). */
int srt_last_trace_ms(srt_context* ctx, float* ms);

/* Information about the uploaded snapshot: dims[3] and padded edge length of
 * the device grids. */
int srt_grid_info(srt_context* ctx, int64_t* nx, int64_t* ny, int64_t* nz, int64_t* padded_edge);

/* Version string of the library. */
const char* srt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SRT_H */
