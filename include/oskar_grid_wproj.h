/*
 * Drop-in C ABI of the MI355X W-projection gridder.
 *
 * The three entry points below keep the reference's signatures and argument
 * meaning unchanged, so a caller written against OxfordSKA/FPGA_Gridding
 * (main.cpp:516-533, the replacement slot at main.cpp:476-509) links against
 * libfpga_gridding_amd.so instead of oskar_grid_wproj.c without source
 * changes.  All pointers are HOST pointers owned by the caller.  Each call
 * uploads its inputs, grids them on the current HIP device with hand-written
 * gfx950 kernels, and returns with every write complete.
 *
 * Semantics kept from the reference:
 *   - weight[] is not read (the reference fixes the weight to 1,
 *     oskar_grid_wproj.c:81, :185); it may be NULL;
 *   - *num_skipped is reset, then counts points whose footprint crosses the
 *     grid edge (:153, :207-212);
 *   - *norm and grid[] are accumulated, never cleared (:236-241);
 *   - grid is row-major complex, cell (u, v) at grid[2*(v*grid_size + u)].
 *
 * Deliberate deviation: oskar_grid_wproj_f grids all num_points points.  The
 * shipped reference loops over the first 10000 points only
 * (oskar_grid_wproj.c:154-155, a debug leftover); the header contract
 * (oskar_grid_wproj.h:105) and oskar_grid_wproj_d (:65) use num_points.
 *
 * Errors: the reference returns nothing and reports nothing.  These void
 * entry points leave grid, norm and num_skipped untouched on failure and
 * record the code for oskar_grid_wproj_last_error() (fpga_gridding_amd.h);
 * the *_status variants there return it through the OSKAR-style trailing
 * int* status instead.
 */
#ifndef FPGA_GRIDDING_AMD_OSKAR_GRID_WPROJ_H_
#define FPGA_GRIDDING_AMD_OSKAR_GRID_WPROJ_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#  define GW_EXPORT __attribute__((visibility("default")))
#else
#  define GW_EXPORT
#endif

/* Replaces oskar_grid_wproj_d, reference oskar_grid_wproj.h:68-86
 * (implementation oskar_grid_wproj.c:39-126).  FP64 index math and grid. */
GW_EXPORT
void oskar_grid_wproj_d(
        const size_t num_w_planes,
        const int* support,
        const int oversample,
        const int conv_size_half,
        const double* conv_func,
        const size_t num_points,
        const double* uu,
        const double* vv,
        const double* ww,
        const double* vis,
        const double* weight,
        const double cell_size_rad,
        const double w_scale,
        const int grid_size,
        size_t* num_skipped,
        double* norm,
        double* grid);

/* Replaces oskar_grid_wproj_f, reference oskar_grid_wproj.h:113-131
 * (implementation oskar_grid_wproj.c:128-243).  FP32 index math bit-exact
 * with the reference, FP32 products and grid, FP64 norm. */
GW_EXPORT
void oskar_grid_wproj_f(
        const size_t num_w_planes,
        const int* support,
        const int oversample,
        const int conv_size_half,
        const float* conv_func,
        const size_t num_points,
        const float* uu,
        const float* vv,
        const float* ww,
        const float* vis,
        const float* weight,
        const float cell_size_rad,
        const float w_scale,
        const int grid_size,
        size_t* num_skipped,
        double* norm,
        float* grid);

/* Replaces oskar_grid_wproj_f_diagnostics, reference
 * oskar_grid_wproj.c:245-370 (exported there, not declared in its header).
 * Grids like oskar_grid_wproj_f and ADDS into the caller's histograms:
 *   num_vis_per_wsupport[support]  += 1 per point, skipped points included;
 *   num_vis_per_wplane[grid_w]     += 1 per point, grid_w the UNCLAMPED
 *                                     plane index (caller sizes the array);
 *   num_vis_per_grid[cell]         += 1 per footprint tap;
 *   avg_wsupport_per_grid[cell]    += support + grid_w per footprint tap;
 *   avg_wplane_per_grid            is not written (as in the reference).
 * The per-cell sums are formed exactly in integers and added once, which
 * equals the reference's float accumulation while partial sums stay below
 * 2^24 and the array holds integers on entry. */
GW_EXPORT
void oskar_grid_wproj_f_diagnostics(
        const size_t num_w_planes,
        const int* support,
        const int oversample,
        const int conv_size_half,
        const float* conv_func,
        const size_t num_points,
        const float* uu,
        const float* vv,
        const float* ww,
        const float* vis,
        const float* weight,
        const float cell_size_rad,
        const float w_scale,
        const int grid_size,
        size_t* num_skipped,
        double* norm,
        float* grid,
        long long int* num_vis_per_wsupport,
        long long int* num_vis_per_wplane,
        long long int* num_vis_per_grid,
        float* avg_wsupport_per_grid,
        float* avg_wplane_per_grid);

#ifdef __cplusplus
}
#endif

#endif /* FPGA_GRIDDING_AMD_OSKAR_GRID_WPROJ_H_ */
