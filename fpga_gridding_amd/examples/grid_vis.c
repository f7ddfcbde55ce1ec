/*
 * grid_vis: the reference harness's flow (main.cpp:100-605) on this
 * library, as a user of the reference would run it on an MI355X.
 *
 *   grid_vis <input_root> [output_root]
 *
 * Reads <input_root>.vis (OSKAR binary visibilities, single channel and
 * polarisation) and the kernel cube <input_root>_KERNELS_REAL.fits /
 * _KERNELS_IMAG.fits, builds the uniform-weighting grid over all blocks
 * (oskar_grid_weights_write_f), reads the weights back per block
 * (oskar_grid_weights_read_f), grids every block with oskar_grid_wproj_f
 * on the GPU, and writes the grid as <output_root>_REAL.fits /
 * _IMAG.fits (default output_root: vis_grid_new).  Prints the skip count,
 * the norm and the gridding time.
 */
#include "fpga_gridding_amd.h"
#include "fpga_gridding_amd_io.h"
#include "oskar_grid_weights.h"
#include "oskar_grid_wproj.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static int fail(const char* what, int status)
{
    fprintf(stderr, "grid_vis: %s (status %d)\n", what, status);
    return EXIT_FAILURE;
}

int main(int argc, char** argv)
{
    int status = 0;
    if (argc < 2)
    {
        fprintf(stderr, "Usage: grid_vis <input_root> [output_root]\n");
        return EXIT_FAILURE;
    }
    const char* root = argv[1];
    const char* out_root = argc > 2 ? argv[2] : "vis_grid_new";
    char* name = (char*)calloc(strlen(root) + 32, 1);
    sprintf(name, "%s.vis", root);
    gw_binary* h = gw_binary_open(name, &status);
    if (!h) return fail("cannot open the visibility file", status);

    int tags_per_block = 0, vis_type = 0, coord_prec = 0, max_times = 0;
    int times_total = 0, channels = 0, stations = 0;
    double freq = 0.0;
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_NUM_TAGS_PER_BLOCK, 0, &tags_per_block, &status);
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER, GW_VIS_HEADER_TAG_AMP_TYPE, 0,
            &vis_type, &status);
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_COORD_PRECISION, 0, &coord_prec, &status);
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_MAX_TIMES_PER_BLOCK, 0, &max_times, &status);
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_NUM_TIMES_TOTAL, 0, &times_total, &status);
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_NUM_CHANNELS_TOTAL, 0, &channels, &status);
    gw_binary_read_int(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_NUM_STATIONS, 0, &stations, &status);
    gw_binary_read_double(h, GW_VIS_GROUP_HEADER,
            GW_VIS_HEADER_TAG_FREQ_START_HZ, 0, &freq, &status);
    if (status) return fail("cannot read the visibility header", status);
    if (channels != 1 || (vis_type & GW_OSKAR_MATRIX) ||
            (vis_type & 0x0F) != GW_OSKAR_SINGLE ||
            coord_prec != GW_OSKAR_SINGLE)
        return fail("needs single-channel, single-polarisation float data",
                    -3);
    const int num_blocks = (times_total + max_times - 1) / max_times;
    const int num_baselines = stations * (stations - 1) / 2;
    const size_t cap = (size_t)max_times * num_baselines;

    /* Kernels: real and imaginary cubes interleaved (main.cpp:183-201). */
    int C = 0, P = 0, ov = 0, G = 0, *support = NULL;
    double cell = 0.0, w_scale = 0.0;
    sprintf(name, "%s_KERNELS_REAL.fits", root);
    float* kr = (float*)gw_read_kernel(name, GW_OSKAR_SINGLE, &C, &P,
            &support, &ov, &G, &cell, &w_scale, &status);
    sprintf(name, "%s_KERNELS_IMAG.fits", root);
    float* ki = (float*)gw_read_kernel(name, GW_OSKAR_SINGLE, &C, &P, NULL,
            &ov, &G, &cell, &w_scale, &status);
    if (status) return fail("cannot read the kernel cubes", status);
    const size_t ncube = (size_t)C * C * P;
    float* kernels = (float*)malloc(2 * ncube * sizeof(float));
    for (size_t j = 0; j < ncube; ++j)
    {
        kernels[2 * j] = kr[j];
        kernels[2 * j + 1] = ki[j];
    }
    gw_free(kr);
    gw_free(ki);

    float* uu = (float*)malloc(cap * sizeof(float));
    float* vv = (float*)malloc(cap * sizeof(float));
    float* ww = (float*)malloc(cap * sizeof(float));
    float* weight = (float*)malloc(cap * sizeof(float));
    float* weight_1 = (float*)malloc(cap * sizeof(float));
    float* vis = (float*)malloc(2 * cap * sizeof(float));
    for (size_t j = 0; j < cap; ++j) weight_1[j] = 1.0f;
    const size_t cells = (size_t)G * G;
    float* weights_grid = (float*)calloc(cells, sizeof(float));
    float* grid = (float*)calloc(2 * cells, sizeof(float));
    size_t skipped = 0, skipped_total = 0;
    double norm = 0.0;
    int dims[6];

    /* Uniform weighting: the weights grid over all blocks. */
    for (int i = 0; i < num_blocks && !status; ++i)
    {
        gw_binary_set_query_search_start(h, i * tags_per_block, &status);
        gw_binary_read(h, GW_OSKAR_INT, GW_VIS_GROUP_BLOCK,
                GW_VIS_BLOCK_TAG_DIM_START_AND_SIZE, i, sizeof(dims), dims,
                &status);
        const size_t n = (size_t)dims[2] * num_baselines;
        gw_vis_read_coords(h, i, coord_prec, n, uu, vv, ww, &status);
        gw_vis_scale_coords(freq, coord_prec, n, uu, vv, ww);
        oskar_grid_weights_write_f(n, uu, vv, weight_1, (float)cell, G,
                &skipped, weights_grid);
        status = status ? status : oskar_grid_wproj_last_error();
    }
    if (status) return fail("weighting failed", status);

    /* Gridding, block by block. */
    const double t0 = now_s();
    for (int i = 0; i < num_blocks && !status; ++i)
    {
        gw_binary_set_query_search_start(h, i * tags_per_block, &status);
        gw_binary_read(h, GW_OSKAR_INT, GW_VIS_GROUP_BLOCK,
                GW_VIS_BLOCK_TAG_DIM_START_AND_SIZE, i, sizeof(dims), dims,
                &status);
        const size_t n = (size_t)dims[2] * num_baselines;
        gw_binary_read(h, (unsigned char)vis_type, GW_VIS_GROUP_BLOCK,
                GW_VIS_BLOCK_TAG_CROSS_CORRELATIONS, i,
                2 * cap * sizeof(float), vis, &status);
        gw_vis_read_coords(h, i, coord_prec, n, uu, vv, ww, &status);
        gw_vis_scale_coords(freq, coord_prec, n, uu, vv, ww);
        oskar_grid_weights_read_f(n, uu, vv, weight_1, weight, (float)cell,
                G, &skipped, weights_grid);
        if (!status) status = oskar_grid_wproj_last_error();
        oskar_grid_wproj_f((size_t)P, support, ov, C, kernels, n, uu, vv, ww,
                vis, weight, (float)cell, (float)w_scale, G, &skipped, &norm,
                grid);
        if (!status) status = oskar_grid_wproj_last_error();
        skipped_total += skipped;
    }
    const double t1 = now_s();
    gw_binary_free(h);
    if (status) return fail("gridding failed", status);
    printf("blocks %d, points per block <= %zu, grid %d^2, planes %d\n",
           num_blocks, cap, G, P);
    printf("num_skipped %zu, norm %.17g, gridding %.3f ms\n", skipped_total,
           norm, 1e3 * (t1 - t0));
    gw_write_fits_cube(GW_OSKAR_SINGLE | GW_OSKAR_COMPLEX, grid, out_root, G,
            G, 1, 0, &status);
    if (status) return fail("cannot write the grid", status);
    free(name); free(kernels); free(support); free(uu); free(vv); free(ww);
    free(weight); free(weight_1); free(vis); free(weights_grid); free(grid);
    return EXIT_SUCCESS;
}
