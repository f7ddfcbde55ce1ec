/*
 * Drop-in C ABI of the uniform-weighting step on the MI355X.
 *
 * The four entry points keep the reference's signatures and argument
 * meaning (utils/oskar_grid_weights.h:59-135; implementation
 * utils/oskar_grid_weights.c:38-172), so the weighting step before the
 * gridder (main.cpp:445-449) links against libfpga_gridding_amd.so without
 * source changes.  All pointers are HOST pointers owned by the caller; each
 * call uploads its inputs, runs on the current HIP device, and returns with
 * every write complete.
 *
 * Semantics kept from the reference, bit for bit:
 *   - the nearest cell: u = (int)round(-uu * grid_size * cell_size_rad) +
 *     grid_size / 2 (v likewise from +vv), in the precision of the call;
 *     points outside [0, grid_size) in u or v are counted in *num_skipped
 *     (reset at entry) and not used;
 *   - write: grid[v * grid_size + u] += weight, the additions of one cell in
 *     input order (the device result equals the sequential loop's);
 *   - read: weight_out = weight_in / grid[cell] if that is non-zero, else 0;
 *     weight_out of skipped points is left as it was.
 *
 * Errors: as the reference, these return nothing; on failure they leave
 * every output untouched and record the code for
 * oskar_grid_wproj_last_error() (fpga_gridding_amd.h).  The device-pointer
 * forms are gw_grid_weights_write_device / gw_grid_weights_read_device.
 */
#ifndef FPGA_GRIDDING_AMD_OSKAR_GRID_WEIGHTS_H_
#define FPGA_GRIDDING_AMD_OSKAR_GRID_WEIGHTS_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#  define GW_EXPORT __attribute__((visibility("default")))
#else
#  define GW_EXPORT
#endif

/* Replaces oskar_grid_weights_write_d, reference
 * utils/oskar_grid_weights.h:59-64 (implementation .c:38-70). */
GW_EXPORT
void oskar_grid_weights_write_d(const size_t num_points, const double* uu,
        const double* vv, const double* weight, const double cell_size_rad,
        const int grid_size, size_t* num_skipped, double* grid);

/* Replaces oskar_grid_weights_read_d, reference
 * utils/oskar_grid_weights.h:83-88 (implementation .c:72-104). */
GW_EXPORT
void oskar_grid_weights_read_d(const size_t num_points, const double* uu,
        const double* vv, const double* weight_in, double* weight_out,
        const double cell_size_rad, const int grid_size,
        size_t* num_skipped, const double* grid);

/* Replaces oskar_grid_weights_write_f, reference
 * utils/oskar_grid_weights.h:106-111 (implementation .c:106-138). */
GW_EXPORT
void oskar_grid_weights_write_f(const size_t num_points, const float* uu,
        const float* vv, const float* weight, const float cell_size_rad,
        const int grid_size, size_t* num_skipped, float* grid);

/* Replaces oskar_grid_weights_read_f, reference
 * utils/oskar_grid_weights.h:130-135 (implementation .c:140-172). */
GW_EXPORT
void oskar_grid_weights_read_f(const size_t num_points, const float* uu,
        const float* vv, const float* weight_in, float* weight_out,
        const float cell_size_rad, const int grid_size,
        size_t* num_skipped, const float* grid);

#ifdef __cplusplus
}
#endif

#endif /* FPGA_GRIDDING_AMD_OSKAR_GRID_WEIGHTS_H_ */
