/*
 * Extended C ABI of the MI355X W-projection gridder (libfpga_gridding_amd.so).
 *
 * oskar_grid_wproj.h holds the drop-in entry points with the reference's
 * exact signatures.  This header adds what a caller that keeps data on the
 * GPU needs, in the reference's own conventions (OSKAR-style trailing
 * `int* status`, negative codes from utils/oskar_global.h:43-64):
 *
 *   - *_status variants of the drop-in calls;
 *   - a last-error query for the void drop-in calls;
 *   - a resident plan: the W-kernel cube, its derived tables and the
 *     workspace stay on the device across calls, and points, visibilities
 *     and the grid are passed as DEVICE pointers (no PCIe in the call);
 *   - device-memory helpers so that callers need no other GPU runtime;
 *   - per-stage device timings of the last call (HIP events on the plan's
 *     stream).
 *
 * Nothing here has a counterpart in the reference beyond the entry points
 * cited per function.
 */
#ifndef FPGA_GRIDDING_AMD_H_
#define FPGA_GRIDDING_AMD_H_

#include <stddef.h>
#include "oskar_grid_wproj.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the reference's values (utils/oskar_global.h:46-64). */
enum gw_status_code
{
    GW_OK                         = 0,
    GW_ERR_INVALID_ARGUMENT       = -3,   /* OSKAR_ERR_INVALID_ARGUMENT */
    GW_ERR_FUNCTION_NOT_AVAILABLE = -4,   /* OSKAR_ERR_FUNCTION_NOT_AVAILABLE */
    GW_ERR_OUT_OF_RANGE           = -5,   /* OSKAR_ERR_OUT_OF_RANGE */
    GW_ERR_MEMORY_ALLOC_FAILURE   = -6,   /* OSKAR_ERR_MEMORY_ALLOC_FAILURE */
    GW_ERR_MEMORY_COPY_FAILURE    = -7,   /* OSKAR_ERR_MEMORY_COPY_FAILURE */
    GW_ERR_KERNEL_LAUNCH_FAILURE  = -18,  /* OSKAR_ERR_KERNEL_LAUNCH_FAILURE */
    GW_ERR_COMPUTE_DEVICES        = -19   /* OSKAR_ERR_COMPUTE_DEVICES */
};

/* Precision of a plan.  GW_SINGLE and GW_DOUBLE have the values of
 * OSKAR_SINGLE / OSKAR_DOUBLE and the reference's arithmetic
 * (oskar_grid_wproj_f / _d).  GW_HALF and GW_BFLOAT16 are the reduced-
 * precision forms, which have no counterpart in the reference: taps and
 * visibilities stored in FP16 / BF16, FP64 index math (the indices of
 * oskar_grid_wproj_d), FP32 products, FP64 per-job sums and a compensated
 * two-float grid; the norm is summed in FP64 over the stored taps. */
enum gw_precision { GW_SINGLE = 0x01, GW_DOUBLE = 0x02, GW_HALF = 0x03,
                    GW_BFLOAT16 = 0x04, GW_SINGLE_INDEX64 = 0x05 };
/* GW_SINGLE_INDEX64 is GW_SINGLE (FP32 inputs, taps, products and grid;
 * the same data layout, batches and server) with the index math of
 * oskar_grid_wproj_d: plane, cell and sub-cell offset computed in FP64 from
 * the float coordinates, so that no FP32 rounding flips an index.  Its
 * results then differ from the FP64 oracle only by FP32 arithmetic (and
 * its num_skipped is oskar_grid_wproj_d's), where GW_SINGLE keeps
 * oskar_grid_wproj_f's FP32 index semantics. */

/* ---------------------------------------------------------------------- */
/* Status-returning drop-in calls                                          */
/* ---------------------------------------------------------------------- */

/* oskar_grid_wproj_f (reference oskar_grid_wproj.h:113-131) with a trailing
 * OSKAR status: returns at once if *status != 0 on entry; on failure sets
 * *status and leaves num_skipped, norm and grid untouched. */
GW_EXPORT
void oskar_grid_wproj_f_status(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, const float cell_size_rad, const float w_scale,
        const int grid_size, size_t* num_skipped, double* norm, float* grid,
        int* status);

/* oskar_grid_wproj_d (reference oskar_grid_wproj.h:68-86) with status. */
GW_EXPORT
void oskar_grid_wproj_d_status(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const double* conv_func, const size_t num_points, const double* uu,
        const double* vv, const double* ww, const double* vis,
        const double* weight, const double cell_size_rad,
        const double w_scale, const int grid_size, size_t* num_skipped,
        double* norm, double* grid, int* status);

/* Code of the last failed void drop-in call on this thread (0 if the last
 * call succeeded), and a message for any status code. */
GW_EXPORT int oskar_grid_wproj_last_error(void);
GW_EXPORT const char* gw_error_string(int status);
/* Detail text of the last error on this thread (HIP error names etc.). */
GW_EXPORT const char* gw_last_error_detail(void);

/* Release the per-thread device buffers the drop-in calls keep between
 * calls (they are otherwise held until the thread exits). */
GW_EXPORT void gw_release_cached(void);

/* ---------------------------------------------------------------------- */
/* Resident plan: device-resident W-kernels, device-pointer data           */
/* ---------------------------------------------------------------------- */

typedef struct gw_plan gw_plan;

/* Create a plan on the current HIP device.  Arguments have the meaning of
 * the reference gridder's (oskar_grid_wproj.h:50-63); conv_func is a HOST
 * array of 2*num_w_planes*conv_size_half^2 values (double for GW_DOUBLE,
 * float otherwise; GW_HALF / GW_BFLOAT16 round them to 16 bits) in the
 * reference layout conv_func[2*((p*C + iy)*C + ix) + {0,1}].  For
 * GW_SINGLE, cell_size_rad and w_scale are rounded to float as the FP32
 * gridder receives them; the other precisions use them as given. */
GW_EXPORT
gw_plan* gw_plan_create(int precision, size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const void* conv_func, double cell_size_rad, double w_scale,
        int grid_size, int* status);

/* As gw_plan_create for GW_SINGLE / GW_SINGLE_INDEX64, with the float
 * kernel cube already in DEVICE memory (for example from
 * gw_wkernel_generate): the tap tables are built on the device, no cube
 * crosses PCIe.  The tables equal those gw_plan_create builds from the same
 * values. */
GW_EXPORT
gw_plan* gw_plan_create_device_cube(int precision, size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* d_conv_func, double cell_size_rad, double w_scale,
        int grid_size, int* status);

/* Generate a W-kernel cube in DEVICE memory (d_conv_func: 2*num_w_planes*
 * conv_size_half^2 floats, the reference layout): per plane p a Gaussian
 * taper exp(-r^2 / (2 sigma^2)), sigma = max(S_p, 1) / 2.5 pixels, times
 * the chirp exp(i pi w_p s r^2), w_p = (p / (num_w_planes - 1))^2 w_max,
 * s = 2 / (w_max max(S)^2), on the quadrant samples r^2 = (ix^2 + iy^2) /
 * oversample^2, normalised to unit complex sum over the footprint taps at
 * offset zero, computed in FP64 and rounded to float.  No reference
 * counterpart (the reference reads precomputed FITS kernels,
 * read_kernel.c:36-118).  Enqueued on hip_stream (NULL: the null stream);
 * returns after completion. */
GW_EXPORT
void gw_wkernel_generate(size_t num_w_planes, const int* support,
        int oversample, int conv_size_half, double w_max, float* d_conv_func,
        void* hip_stream, int* status);

GW_EXPORT void gw_plan_free(gw_plan* plan);

/* Use the given hipStream_t for the plan's work (NULL: the plan's own
 * non-blocking stream).  The caller keeps ownership of the stream. */
GW_EXPORT void gw_plan_set_stream(gw_plan* plan, void* hip_stream);

/* Grid num_points points held in DEVICE memory into a DEVICE grid
 * (accumulated, row-major complex).  Weights are not used, as in the
 * reference.  Asynchronous on the plan's stream: read num_skipped and the
 * norm increment with gw_plan_results().  Element types by precision:
 *   GW_SINGLE   uu/vv/ww float, d_vis 2*num_points float,
 *               d_grid 2*grid_size^2 float (re, im); GW_SINGLE_INDEX64
 *               likewise;
 *   GW_DOUBLE   the same with double;
 *   GW_HALF     uu/vv/ww float, d_vis 2*num_points FP16 (_Float16),
 *               d_grid 4*grid_size^2 float: per cell re_hi, re_lo, im_hi,
 *               im_lo, the value being hi + lo (gw_compensated_resolve);
 *   GW_BFLOAT16 as GW_HALF with d_vis in BF16 (top 16 bits of an FP32). */
GW_EXPORT
void gw_grid_device(gw_plan* plan, size_t num_points, const void* d_uu,
        const void* d_vv, const void* d_ww, const void* d_vis, void* d_grid,
        int* status);

/* gw_grid_device without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as gw_grid_device, but the per-call
 * workspace lives in a static 64 MiB device array of the library (one per
 * GPU, held until the process ends) instead of the plan's own buffers; the
 * call allocates and frees nothing.  A call of more points than that array
 * holds (about 600k points for FP32 plans; the message gives the plan's
 * limit) fails with OSKAR_ERR_OUT_OF_RANGE.  All calls on one GPU share the
 * array: calls that overlap on one GPU (on different streams or from
 * different threads) race and must be serialised by the caller.  The plan
 * itself is created beforehand (gw_plan_create allocates). */
GW_EXPORT
void gw_grid_device_noalloc(gw_plan* plan, size_t num_points,
        const void* d_uu, const void* d_vv, const void* d_ww,
        const void* d_vis, void* d_grid, int* status);

/* Wait for the plan's last gw_grid_device call and return its skip count,
 * the increment to add to the caller's norm, and the number of grid
 * updates (sum over gridded points of (2*support+1)^2). */
GW_EXPORT
void gw_plan_results(gw_plan* plan, size_t* num_skipped, double* norm_delta,
        unsigned long long* num_updates, int* status);

/* Work the scatter of the plan's last gw_grid_device call executed (no
 * counterpart in the reference; a measurement aid): points whose sort keys
 * are equal (same footprint origin, w-plane and sub-cell table) are
 * gridded as one record with their visibility terms summed, so the
 * scatter takes num_records records and applies record_updates taps (sum
 * over records of (2*support+1)^2), against gw_plan_results' num_updates
 * for the points one by one.  Waits for the call like gw_plan_results. */
GW_EXPORT
void gw_plan_records(gw_plan* plan, unsigned long long* num_records,
        unsigned long long* record_updates, int* status);

/* Scatter shape of a plan (no counterpart in the reference; a tuning knob,
 * results do not depend on it): tile side 2^tile_log2 of footprint origins
 * per work item (-1, the default: the largest side up to 64, a power of two
 * or not, whose FP64 LDS region fits; the autotuner also picks such
 * sides), at most
 * slot_cap footprint cells per lane (0: the largest: 8 for FP32-product
 * plans, 6 for GW_DOUBLE), at least min_share points per stream chunk
 * (0: 8).  Waits for the plan's stream. */
GW_EXPORT
void gw_plan_set_tuning(gw_plan* plan, int tile_log2, int slot_cap,
        int min_share, int* status);

/* values[0..6] = tile_log2 (-1 when the side is not a power of two),
 * slot_cap, min_share, LDS region side (0: none), largest tile_log2 whose
 * region fits, tile side, largest tile side whose region fits.  Returns the
 * number written (at most max_n). */
GW_EXPORT int gw_plan_get_tuning(gw_plan* plan, int* values, int max_n);

/* Autotuner: grids the given device points (into d_scratch_grid, sized as
 * for gw_grid_device and overwritten) with scatter shapes around the
 * current one -- tile side, then slots per lane, then chunk minimum, each
 * kept when faster over `reps` timed calls -- and keeps the fastest on the
 * plan.  *best_ms (may be NULL) gets its device milliseconds per call. */
GW_EXPORT
void gw_plan_autotune(gw_plan* plan, size_t num_points, const void* d_uu,
        const void* d_vv, const void* d_ww, const void* d_vis,
        void* d_scratch_grid, int reps, double* best_ms, int* status);

/* Device milliseconds of the stages of the last gw_grid_device call,
 * measured with HIP events on the plan's stream:
 *   ms[0] locate (index math, skip count, histograms, sort keys)
 *   ms[1] sort   (radix sort of the point keys)
 *   ms[2] gather (sorted point records) + work planner
 *   ms[3] grid   (the scatter kernel)
 *   ms[4] total  (first to last event)
 * Returns the number of values written (<= max_n). */
GW_EXPORT int gw_plan_timings(gw_plan* plan, double* ms, int max_n);

/* ---------------------------------------------------------------------- */
/* Batches: many FP32 jobs in one pipeline and one persistent scatter      */
/* launch (no counterpart in the reference; it grids one job per call)     */
/* ---------------------------------------------------------------------- */

typedef struct gw_batch gw_batch;

/* A batch on the current HIP device, with its own non-blocking stream. */
GW_EXPORT gw_batch* gw_batch_create(int* status);
GW_EXPORT void gw_batch_free(gw_batch* batch);

/* Add a job: a GW_SINGLE plan (kernel set, grid geometry, scatter shape;
 * it may serve several jobs and must outlive the batch's use of it) and
 * device arrays as for gw_grid_device.  Returns the job index (-1 on
 * error).  Jobs must not share a grid. */
GW_EXPORT int gw_batch_add(gw_batch* batch, gw_plan* plan, size_t num_points,
        const float* d_uu, const float* d_vv, const float* d_ww,
        const float* d_vis, float* d_grid, int* status);
GW_EXPORT void gw_batch_clear(gw_batch* batch);
GW_EXPORT void gw_batch_set_stream(gw_batch* batch, void* hip_stream);

/* Enqueue the pipeline of all jobs with ordinary launches. */
GW_EXPORT void gw_batch_run(gw_batch* batch, int* status);
/* Capture the pipeline of the current jobs into a HIP graph: only the
 * library's own kernels, no memset nodes, no library sort. */
GW_EXPORT void gw_batch_capture(gw_batch* batch, int* status);
/* Replay the captured graph (captures first if the job list changed). */
GW_EXPORT void gw_batch_launch(gw_batch* batch, int* status);
/* Waits; num_skipped, norm increment and update count of one job of the
 * last run / launch. */
GW_EXPORT void gw_batch_results(gw_batch* batch, int job, size_t* num_skipped,
        double* norm_delta, unsigned long long* num_updates, int* status);
/* Device milliseconds of the last run / launch (waits). */
GW_EXPORT double gw_batch_last_ms(gw_batch* batch);

/* ---------------------------------------------------------------------- */
/* Uniform weighting on device pointers                                    */
/* ---------------------------------------------------------------------- */
/* oskar_grid_weights_write_* / _read_* (oskar_grid_weights.h in this
 * include/ directory) on DEVICE pointers, precision GW_SINGLE (float
 * arrays) or GW_DOUBLE (double arrays); cell_size_rad is converted to the
 * precision of the call.  Enqueued on hip_stream (NULL: a stream of the
 * calling thread's cached workspace); returns after completion.  Results
 * are bit-exact with the reference's loops (utils/oskar_grid_weights.c:
 * 38-172): grid additions of one cell happen in input order. */
GW_EXPORT void gw_grid_weights_write_device(int precision, size_t num_points,
        const void* d_uu, const void* d_vv, const void* d_weight,
        double cell_size_rad, int grid_size, size_t* num_skipped,
        void* d_grid, void* hip_stream, int* status);
GW_EXPORT void gw_grid_weights_read_device(int precision, size_t num_points,
        const void* d_uu, const void* d_vv, const void* d_weight_in,
        void* d_weight_out, double cell_size_rad, int grid_size,
        size_t* num_skipped, const void* d_grid, void* hip_stream,
        int* status);

/* ---------------------------------------------------------------------- */
/* Serving: a request loop with admission by device-memory budget         */
/* ---------------------------------------------------------------------- */
/* A server owns a worker thread on the current HIP device.  Requests (the
 * arguments of oskar_grid_wproj_f, host arrays) are queued FIFO; the worker
 * admits them while the budget allows -- kernel-set tables (plans, cached
 * per kernel set, grid side, cell size and w_scale; idle ones evicted least
 * recently used first), resident grids, staging and batch workspace count
 * against it -- and grids each admitted group as one batch (gw_batch_*) on
 * one of several streams, so that requests arriving while others run start
 * as soon as a stream and memory are free.  No reference counterpart: the
 * reference grids one job per call (oskar_grid_wproj.c:128-243). */

typedef struct gw_server gw_server;

typedef struct gw_request_result
{
    int status;                  /* 0, or the request's error code */
    size_t num_skipped;
    double norm_delta;           /* this request's norm increment */
    unsigned long long num_updates;
    /* Steady-clock seconds: submitted, admitted, enqueued, completed. */
    double t_submit, t_admit, t_start, t_done;
} gw_request_result;

typedef struct gw_server_stats
{
    size_t completed, batches, max_batch, plans_evicted, points;
    size_t device_bytes;         /* held now (tables, grids, staging) */
    size_t resident_plans, pending;
} gw_server_stats;

/* hbm_budget_bytes: device memory the server may hold; max_batch_jobs:
 * most requests gridded together; num_streams: batches in flight (1-8);
 * precision: GW_SINGLE (oskar_grid_wproj_f's index math) or
 * GW_SINGLE_INDEX64 (FP64 index math), for every request. */
GW_EXPORT gw_server* gw_server_create(size_t hbm_budget_bytes,
        int max_batch_jobs, int num_streams, int precision, int* status);
/* Finishes every queued request, then frees all device memory. */
GW_EXPORT void gw_server_free(gw_server* server);
/* Registers a W-kernel set (arguments as oskar_grid_wproj_f's, copied);
 * returns its id. */
GW_EXPORT int gw_server_add_kernel_set(gw_server* server, size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* conv_func, int* status);
/* Registers a kernel set generated on the device by its parameters (as
 * gw_wkernel_generate); the cube stays in device memory (counted against
 * the budget) and plans are built from it there.  Returns its id. */
GW_EXPORT int gw_server_add_wkernel_set(gw_server* server,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, double w_max, int* status);
/* A zeroed resident grid (grid_size^2 complex64) that requests accumulate
 * into; returns its id.  Reading waits until no queued request uses it. */
GW_EXPORT int gw_server_create_grid(gw_server* server, int grid_size,
        int* status);
GW_EXPORT void gw_server_read_grid(gw_server* server, int grid_id,
        float* grid, int* status);
GW_EXPORT void gw_server_free_grid(gw_server* server, int grid_id);
/* Queues a request; returns its id (> 0).  The host arrays must stay
 * valid until the request has been waited for.  grid_id >= 0 adds into
 * that resident grid; grid_id < 0 grids into a fresh zeroed grid that is
 * copied to grid_out (grid_size^2 complex64) on completion. */
GW_EXPORT long long gw_server_submit(gw_server* server, int kernel_set,
        int grid_size, double cell_size_rad, double w_scale,
        size_t num_points, const float* uu, const float* vv, const float* ww,
        const float* vis, int grid_id, float* grid_out, int* status);
/* Waits for a request (timeout_s < 0: without limit).  Returns 1 and fills
 * *result once it is done (the request is then forgotten), 0 on timeout. */
GW_EXPORT int gw_server_wait(gw_server* server, long long request,
        double timeout_s, gw_request_result* result, int* status);
GW_EXPORT void gw_server_get_stats(gw_server* server, gw_server_stats* out);

/* ---------------------------------------------------------------------- */
/* Device memory helpers (current HIP device)                              */
/* ---------------------------------------------------------------------- */
/* The copy and memset helpers first wait for all work on the device (any
 * stream, the plans' and batches' included) and complete before they
 * return, so they order correctly with asynchronous gridding calls. */

GW_EXPORT int   gw_device_count(void);
GW_EXPORT void* gw_device_malloc(size_t bytes, int* status);
GW_EXPORT void  gw_device_free(void* ptr);
/* Page-locked host memory (hipHostMalloc): copies from it, such as a
 * server request's points, run as asynchronous DMA instead of staged
 * copies from pageable memory.  Free with gw_host_free. */
GW_EXPORT void* gw_host_malloc(size_t bytes, int* status);
GW_EXPORT void  gw_host_free(void* ptr);
GW_EXPORT void  gw_memcpy_h2d(void* dst, const void* src, size_t bytes,
        int* status);
GW_EXPORT void  gw_memcpy_d2h(void* dst, const void* src, size_t bytes,
        int* status);
GW_EXPORT void  gw_memset_device(void* ptr, int value, size_t bytes,
        int* status);
GW_EXPORT void  gw_device_synchronize(int* status);
/* Compensated grid (4 floats per cell, re_hi, re_lo, im_hi, im_lo) to a
 * float grid (re, im per cell): d_out[2q + c] = hi + lo.  Enqueued on the
 * given hipStream_t (NULL: the null stream). */
GW_EXPORT void  gw_compensated_resolve(const float* d_comp, size_t num_cells,
        float* d_out, void* hip_stream, int* status);

/* Device-to-device copy with the library's own streaming kernel (k_copy,
 * 16-byte loads and stores); synchronous.  Besides plain use it calibrates
 * the HBM byte counters of a profiler: a copy of B bytes moves exactly B
 * bytes each way. */
GW_EXPORT void  gw_device_copy(void* dst, const void* src, size_t bytes,
        int* status);

#ifdef __cplusplus
}
#endif

#endif /* FPGA_GRIDDING_AMD_H_ */
